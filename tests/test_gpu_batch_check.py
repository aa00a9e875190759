"""Batch checks by a random linear combination on the GPU.

* ufsecp_gpu_ext_schnorr_batch_check: the reference engine's BIP-340 batch
  verification (src/cpu/src/batch_verify.cpp:324-424) as one device MSM.  Its
  answer must equal "every item verifies" under the per-item verdicts, which
  the oracle and the reference fixtures pin.
* The libbitcoin bridge's commitment batch check
  (compat/libbitcoin_bridge/src/ufsecp_libbitcoin.cpp:1214-1262) composed, as
  the bridge does, from two ufsecp_gpu_msm calls of this library.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as orc
from ultrafastsecp256k1_amd import UFSECP_OK, UfsecpError
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu

N = wl.N_ORDER
P = wl.P_FIELD


def _clean(b):
    keep = b.corrupted == 0
    return b.msgs[keep], b.pubkeys[keep], b.sigs[keep]


@pytest.mark.parametrize("n", [1, 2, 255, 4096, 1 << 16])
def test_batch_check_valid_batches(gpu, n):
    b = wl.generate("schnorr", n + n // 50 + 2, wl.SEED_SCHNORR + 90 + n)
    m, p, s = (a[:n] for a in _clean(b))
    assert len(m) == n
    assert bool(orc.verify_batch("schnorr", m, p, s).all())
    assert gpu.schnorr_batch_check(m, p, s) is True


@pytest.mark.parametrize("where", [0, 1, 777, 4095])
def test_batch_check_one_invalid(gpu, where):
    b = wl.generate("schnorr", 4200, wl.SEED_SCHNORR + 91)
    m, p, s = (a[:4096].copy() for a in _clean(b))
    for field in range(4):
        mm, pp, ss = m.copy(), p.copy(), s.copy()
        if field == 0:
            mm[where, 5] ^= 1                                   # wrong message
        elif field == 1:
            ss[where, 40] ^= 0x10                               # wrong s
        elif field == 2:
            ss[where, 32:] = np.frombuffer(N.to_bytes(32, "big"), np.uint8)   # s = n
        else:
            ss[where, :32] = np.frombuffer((P + 1).to_bytes(32, "big"), np.uint8)  # r >= p
        assert not orc.verify_batch("schnorr", mm[where:where + 1], pp[where:where + 1], ss[where:where + 1])[0]
        assert gpu.schnorr_batch_check(mm, pp, ss) is False, field


def test_batch_check_golden_vectors(gpu, golden):
    cases = golden["bip340_vectors"]
    cat = lambda cs, k: np.frombuffer(b"".join(bytes.fromhex(c[k]) for c in cs), np.uint8)  # noqa: E731
    good = [c for c in cases if c["verdict"] == 1]
    assert gpu.schnorr_batch_check(cat(good, "msg"), cat(good, "pubkey"), cat(good, "sig")) is True
    for c in cases:
        if c["verdict"] == 0:
            cs = good + [c]
            assert gpu.schnorr_batch_check(cat(cs, "msg"), cat(cs, "pubkey"), cat(cs, "sig")) is False, c["vector"]
    edge = golden["edge_cases"]["schnorr"]
    for c in edge:
        cs = good + [c]
        got = gpu.schnorr_batch_check(np.frombuffer(b"".join(bytes.fromhex(x["msg"]) for x in good) +
                                                    bytes.fromhex(c["msg"]), np.uint8),
                                      np.frombuffer(b"".join(bytes.fromhex(x["pubkey"]) for x in good) +
                                                    bytes.fromhex(c["pub"]), np.uint8),
                                      np.frombuffer(b"".join(bytes.fromhex(x["sig"]) for x in good) +
                                                    bytes.fromhex(c["sig"]), np.uint8))
        assert got is bool(c["verdict"]), c["name"]


def test_batch_check_counts(gpu):
    ok = __import__("ctypes").c_int(7)
    assert gpu.lib.ufsecp_gpu_ext_schnorr_batch_check(gpu.ctx, None, None, None, 0, __import__("ctypes").byref(ok)) \
        == UFSECP_OK
    assert ok.value == 0


def _commitment_batch_ok(gpu, ix, tw, tx, par):
    """The bridge's ufsecp_lbtc_commitment_batch_ok: Fiat-Shamir weights from
    SHA-256 of the whole batch, then two MSMs; 1, 0, or -1 on an MSM error."""
    n = len(par)
    e = hashlib.sha256(ix.tobytes() + tw.tobytes() + tx.tobytes() + par.tobytes()).digest()
    G33 = bytes.fromhex("0279be667ef9dcbbac55a06295ce870b07029bfcdb2dce28d959f2815b16f81798")
    ls, lp, qp = bytearray(), bytearray(), bytearray()
    g = 0
    for i in range(n):
        r = int.from_bytes(hashlib.sha256(e + i.to_bytes(4, "little")).digest(), "big")
        r = r - N if r >= N else r
        ls += r.to_bytes(32, "big")
        lp += b"\x02" + ix[i].tobytes()
        qp += bytes([3 if par[i] else 2]) + tx[i].tobytes()
        t = int.from_bytes(tw[i].tobytes(), "big")
        g = (g + r * (t - N if t >= N else t)) % N
    lp += G33
    try:
        lhs = gpu.msm(bytes(ls) + g.to_bytes(32, "big"), bytes(lp), n + 1)
        rhs = gpu.msm(bytes(ls), bytes(qp), n)
    except UfsecpError:
        return -1
    return 1 if lhs == rhs else 0


def test_commitment_batch_through_msm(gpu, golden):
    d = [c for c in golden["pubdata_vectors"]["commitment"] if c["verdict"] == 1]
    cat = lambda k: np.frombuffer(b"".join(bytes.fromhex(c[k]) for c in d), np.uint8).reshape(len(d), -1)  # noqa
    ix, tw, tx = cat("internal"), cat("tweak"), cat("tweaked")
    par = np.array([1 if c["parity"] else 0 for c in d], np.uint8)
    assert _commitment_batch_ok(gpu, ix, tw, tx, par) == 1
    bad = par.copy()
    bad[3] ^= 1
    assert _commitment_batch_ok(gpu, ix, tw, tx, bad) == 0
    tx2 = tx.copy()
    tx2[5, 31] ^= 1
    assert _commitment_batch_ok(gpu, ix, tw, tx2, par) in (0, -1)
