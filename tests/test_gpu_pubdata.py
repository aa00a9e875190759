"""Public-data key operations on the GPU vs the reference CPU verdicts.

ufsecp_gpu_pubkey_validate, ufsecp_gpu_xonly_validate and
ufsecp_gpu_commitment_verify (include/ufsecp/ufsecp_gpu.h).  Golden verdicts
(tests/golden/pubdata_vectors.json) come from the reference CPU
ufsecp_pubkey_parse and the libbitcoin bridge's CPU commitment check
(compat/libbitcoin_bridge/src/ufsecp_libbitcoin.cpp:1138-1149), including the
seven BIP-341 key-path vectors; larger random sets are checked against the
oracle.
"""
import random

import numpy as np
import pytest

from oracle import oracle as orc
from ultrafastsecp256k1_amd import UFSECP_ERR_BAD_INPUT, UFSECP_ERR_NULL_ARG, UFSECP_OK
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu

N = wl.N_ORDER
P = wl.P_FIELD


def _cat(items, key):
    return np.frombuffer(b"".join(bytes.fromhex(c[key]) for c in items), np.uint8)


def test_pubkey_validate_golden(gpu, golden):
    d = golden["pubdata_vectors"]["pubkey"]
    v = gpu.pubkey_validate(_cat(d, "key"))
    assert [int(x) for x in v] == [c["verdict"] for c in d], [c["name"] for c, x in zip(d, v) if x != c["verdict"]]


def test_xonly_validate_golden(gpu, golden):
    d = golden["pubdata_vectors"]["xonly"]
    v = gpu.xonly_validate(_cat(d, "x"))
    assert [int(x) for x in v] == [c["verdict"] for c in d], [c["name"] for c, x in zip(d, v) if x != c["verdict"]]


def test_commitment_verify_golden(gpu, golden):
    d = golden["pubdata_vectors"]["commitment"]
    par = np.array([c["parity"] for c in d], np.uint8)
    v = gpu.commitment_verify(_cat(d, "internal"), _cat(d, "tweak"), _cat(d, "tweaked"), par)
    assert [int(x) for x in v] == [c["verdict"] for c in d], [c["name"] for c, x in zip(d, v) if x != c["verdict"]]


@pytest.mark.parametrize("n", [1, 63, 64, 65, 3000])
def test_key_validation_random_vs_oracle(gpu, n):
    rnd = random.Random(n)
    keys = bytearray()
    for i in range(n):
        x = rnd.randrange(1 << 256) if i % 3 else rnd.randrange(P)
        keys += bytes([rnd.choice((2, 3, 2, 3, 4, 0))]) + x.to_bytes(32, "big")
    keys = bytes(keys)
    want = np.array([orc.pubkey_validate(keys[33 * i:33 * i + 33]) for i in range(n)], np.uint8)
    assert np.array_equal(gpu.pubkey_validate(keys), want)
    xs = b"".join(keys[33 * i + 1:33 * i + 33] for i in range(n))
    want = np.array([orc.xonly_validate(xs[32 * i:32 * i + 32]) for i in range(n)], np.uint8)
    assert np.array_equal(gpu.xonly_validate(xs), want)


def test_commitment_random_vs_oracle(gpu):
    """Valid commitments from the serving workload's keys, every fifth
    corrupted (tweaked x, parity, or tweak), against the oracle."""
    rnd = random.Random(41)
    n = 600
    b = wl.generate("schnorr", n, wl.SEED_SCHNORR + 41)   # x-only keys (valid lifts)
    ix = b.pubkeys.reshape(-1, 32)
    tw = np.zeros((n, 32), np.uint8)
    tx = np.zeros((n, 32), np.uint8)
    par = np.zeros(n, np.uint8)
    # tweaked keys from the textbook group law
    from ec_math import G as GEN, add, lift_x, mul
    for i in range(n):
        t = rnd.randrange(N)
        Pp = lift_x(int.from_bytes(ix[i].tobytes(), "big"), odd=False)
        Q = add(Pp, mul(t, GEN))
        qx, qp = Q[0], Q[1] & 1
        if i % 5 == 1:
            qx ^= 1 << rnd.randrange(256)
        elif i % 5 == 3:
            qp ^= 1
        elif i % 5 == 4:
            t = t + N if t + N < (1 << 256) else N
        tw[i] = np.frombuffer(t.to_bytes(32, "big"), np.uint8)
        tx[i] = np.frombuffer((qx % (1 << 256)).to_bytes(32, "big"), np.uint8)
        par[i] = qp
    want = np.array([orc.commitment_verify(ix[i].tobytes(), tw[i].tobytes(), tx[i].tobytes(), int(par[i]))
                     for i in range(n)], np.uint8)
    assert 0 < want.sum() < n
    assert np.array_equal(gpu.commitment_verify(ix, tw, tx, par), want)


def test_key_op_argument_checks(gpu):
    lib, ctx = gpu.lib, gpu.ctx
    buf = np.zeros(64, np.uint8)
    out = np.full(2, 7, np.uint8)
    for fn in (lib.ufsecp_gpu_pubkey_validate, lib.ufsecp_gpu_xonly_validate):
        assert fn(None, buf.ctypes.data, 1, out.ctypes.data) == UFSECP_ERR_NULL_ARG
        assert fn(ctx, buf.ctypes.data, 0, out.ctypes.data) == UFSECP_OK
        assert fn(ctx, None, 1, out.ctypes.data) == UFSECP_ERR_NULL_ARG
        assert fn(ctx, buf.ctypes.data, (1 << 26) + 1, out.ctypes.data) == UFSECP_ERR_BAD_INPUT
    f = lib.ufsecp_gpu_commitment_verify
    a = buf.ctypes.data
    assert f(None, a, a, a, a, 1, out.ctypes.data) == UFSECP_ERR_NULL_ARG
    assert f(ctx, a, a, a, None, 1, out.ctypes.data) == UFSECP_ERR_NULL_ARG
    assert f(ctx, a, a, a, a, 0, out.ctypes.data) == UFSECP_OK
    assert f(ctx, a, a, a, a, (1 << 26) + 1, out.ctypes.data) == UFSECP_ERR_BAD_INPUT
