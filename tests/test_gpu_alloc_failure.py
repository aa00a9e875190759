"""Allocation failures forced from the host.

With UFSECP_GPU_FAIL_ALLOC=K in the environment, the library's allocation
number K fails as if the device were out of memory. K counts the library's
device and pinned allocations over the process, the count that
ufsecp_gpu_ext_buffer_alloc_count reports. Each test first finds how many
allocations an operation makes on a fresh context, then fails each of them in
turn. The call must return UFSECP_ERR_GPU_MEMORY with its outputs cleared
and leave none of the library's allocations behind (its live-allocation
count, ufsecp_gpu_ext_buffer_live_count, returns to where it was). The same
context must then run the call without the knob and give exact results. The reference has no such knob;
what is checked is its ABI contract for a non-OK return: outputs cleared,
context still usable (src/cpu/src/ufsecp_gpu_impl.cpp:129-140).
"""
import contextlib
import ctypes
import os

import numpy as np
import pytest

from ultrafastsecp256k1_amd import KIND_ECDSA, KIND_SCHNORR, UfsecpError, UfsecpGpu, VerifyServer, load_library
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu

ERR_GPU_MEMORY = 103
KNOB = "UFSECP_GPU_FAIL_ALLOC"


def _count(lib) -> int:
    return int(lib.ufsecp_gpu_ext_buffer_alloc_count())


@contextlib.contextmanager
def fail_nth(lib, k: int):
    """Fail the k-th library allocation from now (1-based)."""
    os.environ[KNOB] = str(_count(lib) + k)
    try:
        yield
    finally:
        os.environ.pop(KNOB, None)


def _live(lib) -> int:
    return int(lib.ufsecp_gpu_ext_buffer_live_count())


def _buf(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sweep(call, check):
    """call(gpu) -> (rc, outputs); check(outputs) asserts exact results.
    Returns the number of allocations the first call on a context makes."""
    lib = load_library()
    with UfsecpGpu() as g:
        before = _count(lib)
        rc, out = call(g)
        made = _count(lib) - before
        assert rc == 0
        check(out)
    assert made >= 1
    live0 = _live(lib)
    for k in range(1, made + 1):
        with UfsecpGpu() as g:
            with fail_nth(lib, k):
                rc, out = call(g)
            assert rc == ERR_GPU_MEMORY, f"allocation {k} of {made}: rc {rc}"
            for o in out:
                assert not np.any(o), f"allocation {k} of {made}: outputs not cleared"
            rc, out = call(g)
            assert rc == 0, f"allocation {k} of {made}: retry rc {rc}"
            check(out)
    assert _live(lib) == live0, "allocations left live by the failed calls"
    return made


def test_context_creation_fails_cleanly_at_each_allocation(monkeypatch):
    # a comb width no other context of the process holds, so that this
    # context allocates its own G table (the table is shared per device and
    # width)
    monkeypatch.setenv("UFSECP_GCOMB_W", "11")
    lib = load_library()
    before = _count(lib)
    UfsecpGpu().close()
    made = _count(lib) - before
    assert made >= 4          # comb table, midstate, lane slots, work counter
    live0 = _live(lib)
    for k in range(1, made + 1):
        ctx = ctypes.c_void_p()
        with fail_nth(lib, k):
            rc = lib.ufsecp_gpu_ctx_create(ctypes.byref(ctx), 4, 0)
        assert rc == ERR_GPU_MEMORY, k
        assert not ctx.value
    assert _live(lib) == live0, "a failed context creation left allocations live"
    b = wl.generate("schnorr", 1000)
    with UfsecpGpu() as g:
        assert np.array_equal(g.schnorr_verify_batch(b.msgs, b.pubkeys, b.sigs), 1 - b.corrupted)


def test_contexts_share_the_g_table(monkeypatch):
    """A second context on the device with the same comb width takes the
    first one's G table (one allocation fewer), both verify exactly, and the
    table is freed with the last context that holds it."""
    monkeypatch.setenv("UFSECP_GCOMB_W", "12")
    lib = load_library()
    live0 = _live(lib)
    b = wl.generate("ecdsa", 700)
    c0 = _count(lib)
    a = UfsecpGpu()
    first = _count(lib) - c0
    c1 = _count(lib)
    second = UfsecpGpu()
    assert _count(lib) - c1 == first - 1
    for g in (a, second):
        assert np.array_equal(g.ecdsa_verify_batch(b.msgs, b.pubkeys, b.sigs), 1 - b.corrupted)
    a.close()
    assert np.array_equal(second.ecdsa_verify_batch(b.msgs, b.pubkeys, b.sigs), 1 - b.corrupted)
    second.close()
    assert _live(lib) == live0


@pytest.mark.parametrize("kind", ["ecdsa", "schnorr"])
def test_verify_batch_allocation_failures(kind):
    b = wl.generate(kind, 5000, first=300)
    fn = "ufsecp_gpu_ecdsa_verify_batch" if kind == "ecdsa" else "ufsecp_gpu_schnorr_verify_batch"

    def call(g):
        out = np.full(b.n, 7, np.uint8)
        rc = getattr(g.lib, fn)(g.ctx, _buf(b.msgs), _buf(b.pubkeys), _buf(b.sigs), b.n, _buf(out))
        return rc, [out]

    def check(out):
        assert np.array_equal(out[0], 1 - b.corrupted)

    _sweep(call, check)


def test_opaque_rows_allocation_failures():
    b = wl.generate("ecdsa", 3000, first=900)
    stride = 160
    rows = np.zeros(stride * b.n, np.uint8).reshape(b.n, stride)
    rows[:, :32], rows[:, 32:65] = b.msgs, b.pubkeys
    # opaque signature layout: r and s each as little-endian words
    rows[:, 65:97], rows[:, 97:129] = b.sigs[:, 31::-1], b.sigs[:, 63:31:-1]

    def call(g):
        out = np.full(b.n, 7, np.uint8)
        rc = g.lib.ufsecp_gpu_ecdsa_verify_opaque_rows(g.ctx, _buf(rows), stride, b.n, _buf(out))
        return rc, [out]

    def check(out):
        assert np.array_equal(out[0], 1 - b.corrupted)

    _sweep(call, check)


def test_msm_allocation_failures():
    sc, pts, _ = wl.generate_msm(20000, seed=0xA110C)
    lib = load_library()
    with UfsecpGpu() as g:
        expect = g.msm(sc, pts)

    def call(g):
        out = ctypes.create_string_buffer(b"\x07" * 33, 33)
        rc = lib.ufsecp_gpu_msm(g.ctx, _buf(sc), _buf(pts), len(sc), out)
        return rc, [np.frombuffer(out.raw, np.uint8)]

    def check(out):
        assert out[0].tobytes() == expect

    assert _sweep(call, check) >= 10    # one buffer per part of the MSM layout


def test_batch_check_allocation_failures():
    b = wl.generate("schnorr", 4000, first=77)
    good = b.corrupted == 0
    m, p, s = b.msgs[good], b.pubkeys[good], b.sigs[good]

    def call(g):
        ok = ctypes.c_int(7)
        rc = g.lib.ufsecp_gpu_ext_schnorr_batch_check(g.ctx, _buf(m), _buf(p), _buf(s), len(m), ctypes.byref(ok))
        return rc, [np.array([ok.value])]

    def check(out):
        assert out[0][0] == 1

    _sweep(call, check)


def test_multi_batch_allocation_failures():
    """Device-resident multi-batch: the descriptor and tile tables are the
    library's allocations; verdicts are device outputs, so only the retry is
    checked."""
    lib = load_library()
    parts = [wl.generate("ecdsa", 700, first=10), wl.generate("schnorr", 1300, first=20),
             wl.generate("ecdsa", 64, first=30)]

    def run(g, fail_at=None):
        dev, table = [], []
        for b in parts:
            d = (g.to_device(b.msgs), g.to_device(b.pubkeys), g.to_device(b.sigs), g.empty(b.n))
            dev.append(d)
            table.append((KIND_ECDSA if b.pubkeys.shape[1] == 33 else KIND_SCHNORR, b.n,
                          d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr))
        g.synchronize()
        t = g.batch_table(table)
        try:
            if fail_at is None:
                g.verify_multi_device(t, len(table), use_graph=False)
            else:
                with fail_nth(lib, fail_at):
                    g.verify_multi_device(t, len(table), use_graph=False)
            rc = 0
        except UfsecpError as e:
            rc = e.code
        g.synchronize()
        outs = [d[3].download() for d in dev]
        for d in dev:
            for a in d:
                a.free()
        return rc, outs

    with UfsecpGpu() as g:
        before = _count(lib)
        rc, outs = run(g)
        made = _count(lib) - before - 4 * len(parts)      # less the test's own buffers
    assert rc == 0 and made >= 1
    for b, o in zip(parts, outs):
        assert np.array_equal(o, 1 - b.corrupted)
    live0 = _live(lib)
    for k in range(1, made + 1):
        with UfsecpGpu() as g:
            rc, _ = run(g, fail_at=k)
            assert rc == ERR_GPU_MEMORY, k
            rc, outs = run(g)
            assert rc == 0
            for b, o in zip(parts, outs):
                assert np.array_equal(o, 1 - b.corrupted), k
    assert _live(lib) == live0, "a failed multi-batch call left allocations live"


def test_server_creation_fails_cleanly_at_each_allocation():
    """Every ring, cache and launch-group buffer of the serving loop: creation
    fails with UFSECP_ERR_GPU_MEMORY and frees what it took; a server created
    afterwards on the same context admits and verifies requests."""
    lib = load_library()
    cfg = dict(hbm_budget_bytes=16 << 20, inflight_groups=2, key_cache_log2=12)
    e = wl.generate_pool("ecdsa", 6000, 0x5A, 0x5B, pool_size=64)
    with UfsecpGpu() as g:
        before = _count(lib)
        VerifyServer(g, **cfg).close()
        made = _count(lib) - before
        assert made >= 8      # input and verdict rings, cache, per-group buffers
        live0 = _live(lib)
        for k in range(1, made + 1):
            with fail_nth(lib, k):
                with pytest.raises(UfsecpError) as ex:
                    VerifyServer(g, **cfg)
            assert ex.value.code == ERR_GPU_MEMORY, k
        assert _live(lib) == live0, "a failed server creation left allocations live"
        with VerifyServer(g, **cfg) as srv:
            outs = []
            for j in range(0, e.n, 1500):
                o = np.full(min(1500, e.n - j), 7, np.uint8)
                srv.submit(KIND_ECDSA, e.msgs[j:j + 1500], e.pubkeys[j:j + 1500], e.sigs[j:j + 1500], o)
                outs.append(o)
            srv.drain()
        assert np.array_equal(np.concatenate(outs), 1 - e.corrupted)
