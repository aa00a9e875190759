"""Serving loop (continuous batching) on the GPU.

Every request's verdicts must be the same whether it is verified alone
through the plain per-batch ABI call or coalesced with other requests of both
kinds and any size into shared launches, with its public keys resolved
through the device decompression cache (hits, misses, invalid keys) -- and
equal to the workload's known verdicts (1 unless the item was corrupted).
"""
import threading

import numpy as np
import pytest

from ultrafastsecp256k1_amd import KIND_ECDSA, KIND_SCHNORR, VerifyServer
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 700, 1000, 1500, 3, 31, 97]


def _requests(n_ecdsa=8000, n_schnorr=4000, pool_size=32, seed=0x5E):
    """Requests cut from two serving workloads whose keys repeat (pool of
    pool_size keys for 20% of the items), kinds alternating 2:1."""
    e = wl.generate_pool("ecdsa", n_ecdsa, seed, seed + 100, pool_size=pool_size)
    s = wl.generate_pool("schnorr", n_schnorr, seed + 1, seed + 100, pool_size=pool_size)
    reqs, oe, os_ = [], 0, 0
    i = 0
    while True:
        size = SIZES[i % len(SIZES)]
        kind = KIND_SCHNORR if i % 3 == 2 else KIND_ECDSA
        src, off = (e, oe) if kind == KIND_ECDSA else (s, os_)
        if off + size > src.n:
            break
        sl = slice(off, off + size)
        reqs.append((kind, src.msgs[sl].copy(), src.pubkeys[sl].copy(), src.sigs[sl].copy(),
                     1 - src.corrupted[sl]))
        if kind == KIND_ECDSA:
            oe += size
        else:
            os_ += size
        i += 1
    return reqs


def _plain(gpu, kind, m, p, s):
    fn = gpu.ecdsa_verify_batch if kind == KIND_ECDSA else gpu.schnorr_verify_batch
    return fn(m, p, s)


def _submit_all(srv, reqs, threads=4):
    outs = [np.full(len(r[1]), 7, np.uint8) for r in reqs]

    def worker(k):
        for j in range(k, len(reqs), threads):
            kind, m, p, s, _ = reqs[j]
            srv.submit(kind, m, p, s, outs[j])

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    srv.drain()
    return outs


def test_server_results_same_alone_and_in_mix(gpu):
    reqs = _requests()
    with VerifyServer(gpu, key_cache_log2=12) as srv:
        # Warm the key cache with the largest ECDSA request (it holds pool
        # keys), so the requests below hit it however the scheduler groups
        # them: repeats inside one launch group are all misses, since a group
        # resolves its keys before it inserts any.
        big = max((r for r in reqs if r[0] == KIND_ECDSA), key=lambda r: len(r[1]))
        warm = np.zeros(len(big[1]), np.uint8)
        srv.submit(big[0], big[1], big[2], big[3], warm)
        srv.drain()
        assert np.array_equal(warm, big[4])
        srv.stats(reset=True)
        outs = _submit_all(srv, reqs)
        st = srv.stats()
    assert st["requests_completed"] == len(reqs)
    assert st["items_completed"] == sum(len(r[1]) for r in reqs)
    assert st["key_cache_hits"] > 0          # pool keys repeat
    assert st["latency_p99_us"] >= st["latency_p50_us"] > 0
    for (kind, m, p, s, exp), out in zip(reqs, outs):
        alone = _plain(gpu, kind, m, p, s)
        assert np.array_equal(out, alone)
        assert np.array_equal(out, exp)


def test_server_second_pass_hits_cache(gpu):
    """Resubmitting the same requests: keys are now cached, verdicts unchanged."""
    reqs = _requests(n_ecdsa=3000, n_schnorr=1500, pool_size=1 << 16, seed=0x77)
    with VerifyServer(gpu, key_cache_log2=16) as srv:
        first = _submit_all(srv, reqs, threads=2)
        srv.stats(reset=True)
        second = _submit_all(srv, reqs, threads=2)
        st = srv.stats()
    for a, b, r in zip(first, second, reqs):
        assert np.array_equal(a, b)
        assert np.array_equal(a, r[4])
    # nearly every key was inserted on the first pass
    assert st["key_cache_hits"] > 0.8 * st["key_lookups"]


def test_server_cache_keeps_reused_keys_under_one_shot_keys(gpu):
    """Keys found in the cache after their insertion sit in a way that new keys
    never take: 16,384 keys seen once (eight per set of a 2^12-entry table)
    leave a warmed pool of 256 keys resident.  A direct-mapped table would keep
    about e^-4 = 2% of them; here only pool keys that share a set with
    another pool key can be lost."""
    pool = wl.generate_pool("ecdsa", 2048, 0x3A, 0x3B, pool_size=256, pool_permille=1000)
    fresh = wl.generate("ecdsa", 1 << 14, 0x3C)
    req = (KIND_ECDSA, pool.msgs, pool.pubkeys, pool.sigs, 1 - pool.corrupted)
    with VerifyServer(gpu, key_cache_log2=12) as srv:
        for _ in range(3):   # insert, then find (and copy to the kept way)
            assert np.array_equal(_submit_all(srv, [req], threads=1)[0], req[4])
        flood = [(KIND_ECDSA, fresh.msgs[j:j + 4096], fresh.pubkeys[j:j + 4096], fresh.sigs[j:j + 4096],
                  1 - fresh.corrupted[j:j + 4096]) for j in range(0, fresh.n, 4096)]
        for r, out in zip(flood, _submit_all(srv, flood, threads=1)):
            assert np.array_equal(out, r[4])
        srv.stats(reset=True)
        assert np.array_equal(_submit_all(srv, [req], threads=1)[0], req[4])
        st = srv.stats()
    assert st["key_lookups"] == 2048
    assert st["key_cache_hits"] > 0.85 * st["key_lookups"]


def test_server_without_cache(gpu):
    reqs = _requests(n_ecdsa=2000, n_schnorr=1000, seed=0x99)
    with VerifyServer(gpu, key_cache_log2=0) as srv:
        outs = _submit_all(srv, reqs, threads=2)
        st = srv.stats()
    assert st["key_cache_hits"] == 0
    for (kind, m, p, s, exp), out in zip(reqs, outs):
        assert np.array_equal(out, exp)


def test_server_admission_under_small_budget(gpu):
    """A 16 MiB device budget holds a fraction of the submitted bytes: requests
    wait for admission and still complete with exact verdicts."""
    e = wl.generate_pool("ecdsa", 48 * 4096, 0xAD, 0xAE, pool_size=1 << 16)
    reqs = []
    for j in range(48):
        sl = slice(j * 4096, (j + 1) * 4096)
        reqs.append((KIND_ECDSA, e.msgs[sl], e.pubkeys[sl], e.sigs[sl], 1 - e.corrupted[sl]))
    with VerifyServer(gpu, hbm_budget_bytes=16 << 20, inflight_groups=2, key_cache_log2=14,
                      max_group_items=16384) as srv:
        outs = _submit_all(srv, reqs, threads=3)
        st = srv.stats()
    assert st["groups_launched"] >= 12    # at most 4 requests of 4096 per group
    for r, out in zip(reqs, outs):
        assert np.array_equal(out, r[4])


@pytest.mark.parametrize("kind", ["ecdsa", "schnorr"])
def test_server_edge_cases(gpu, golden, kind):
    """Adversarial inputs (invalid keys, out-of-range r/s, high-S ...) alone
    and as part of larger requests, twice (second time through the cache)."""
    cases = golden["edge_cases"][kind]
    k = KIND_ECDSA if kind == "ecdsa" else KIND_SCHNORR
    m = np.stack([np.frombuffer(bytes.fromhex(c["msg"]), np.uint8) for c in cases])
    p = np.stack([np.frombuffer(bytes.fromhex(c["pub"]), np.uint8) for c in cases])
    s = np.stack([np.frombuffer(bytes.fromhex(c["sig"]), np.uint8) for c in cases])
    exp = np.array([c["verdict"] for c in cases], np.uint8)
    b = wl.generate(kind, 300, first=5000)
    mm = np.concatenate([b.msgs[:150], m, b.msgs[150:]])
    pp = np.concatenate([b.pubkeys[:150], p, b.pubkeys[150:]])
    ss = np.concatenate([b.sigs[:150], s, b.sigs[150:]])
    with VerifyServer(gpu, key_cache_log2=10) as srv:
        for _ in range(2):
            reqs = [(k, m, p, s, None), (k, mm, pp, ss, None)] + [(k, m[i:i + 1], p[i:i + 1], s[i:i + 1], None)
                                                                  for i in range(len(cases))]
            outs = _submit_all(srv, reqs, threads=2)
            assert np.array_equal(outs[0], exp)
            assert np.array_equal(outs[1][150:150 + len(cases)], exp)
            assert np.array_equal(outs[1], _plain(gpu, k, mm, pp, ss))
            for i in range(len(cases)):
                assert int(outs[2 + i][0]) == int(exp[i]), cases[i]["name"]


def test_server_mapped_submission(gpu):
    """Inputs in pinned, device-mapped host memory are read in place."""
    reqs = _requests(n_ecdsa=3000, n_schnorr=1500, seed=0x44)
    pinned = []
    for kind, m, p, s, exp in reqs:
        pm, pp, ps = gpu.host_empty(m.shape), gpu.host_empty(p.shape), gpu.host_empty(s.shape)
        pm[:], pp[:], ps[:] = m, p, s
        pinned.append((kind, pm, pp, ps, exp))
    outs = [np.full(len(r[1]), 7, np.uint8) for r in pinned]
    with VerifyServer(gpu, key_cache_log2=12) as srv:
        for (kind, m, p, s, _), out in zip(pinned, outs):
            srv.submit(kind, m, p, s, out, mapped=True)
        srv.drain()
    for (kind, m, p, s, exp), out in zip(pinned, outs):
        assert np.array_equal(out, exp)


def test_server_poisson_arrivals_serving_shape(gpu):
    """The serving configuration's shape: Poisson arrivals of requests whose
    sizes are log-uniform in [1, 8192], 70% ECDSA / 30% BIP-340, 20% of the
    keys from a 2^16-key pool, a 2^21-entry key cache.  Requests are submitted
    from one thread at exponential inter-arrival times (mean 0.4 ms, faster
    than the GPU verifies them, so launch groups overlap and coalesce); every
    request's verdicts equal the known ones, and every 8th equals the same
    request verified alone through the plain ABI."""
    import math
    import random
    import time
    rnd = random.Random(0xC4)
    nreq = 240
    sizes = [min(8192, max(1, int(math.exp(rnd.random() * math.log(8193))))) for _ in range(nreq)]
    kinds = [KIND_ECDSA if rnd.random() < 0.7 else KIND_SCHNORR for _ in range(nreq)]
    ne = sum(s for s, k in zip(sizes, kinds) if k == KIND_ECDSA)
    ns = sum(s for s, k in zip(sizes, kinds) if k == KIND_SCHNORR)
    pools = {KIND_ECDSA: wl.generate_pool("ecdsa", ne, 0xC40, 0xC4F, pool_size=1 << 16, threads=16),
             KIND_SCHNORR: wl.generate_pool("schnorr", ns, 0xC41, 0xC4F, pool_size=1 << 16, threads=16)}
    offs = {KIND_ECDSA: 0, KIND_SCHNORR: 0}
    reqs = []
    for s, k in zip(sizes, kinds):
        src, o = pools[k], offs[k]
        reqs.append((k, src.msgs[o:o + s], src.pubkeys[o:o + s], src.sigs[o:o + s], 1 - src.corrupted[o:o + s]))
        offs[k] += s
    outs = [np.full(len(r[1]), 7, np.uint8) for r in reqs]
    with VerifyServer(gpu, key_cache_log2=21) as srv:
        for (k, m, p, s, _), out in zip(reqs, outs):
            srv.submit(k, m, p, s, out)
            time.sleep(rnd.expovariate(1.0 / 0.4e-3))
        srv.drain()
        st = srv.stats()
    assert st["requests_completed"] == nreq and st["items_completed"] == ne + ns
    assert st["groups_launched"] < nreq            # requests were coalesced
    assert st["latency_p99_us"] >= st["latency_p50_us"] > 0
    for j, ((k, m, p, s, exp), out) in enumerate(zip(reqs, outs)):
        assert np.array_equal(out, exp), j
        if j % 8 == 0:
            assert np.array_equal(out, _plain(gpu, k, m, p, s)), j
