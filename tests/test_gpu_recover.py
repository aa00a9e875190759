"""ECDSA public-key recovery on the GPU vs the reference CPU results.

ufsecp_gpu_ecrecover_batch (include/ufsecp/ufsecp_gpu.h).  Golden results
(tests/golden/recover_vectors.json) come from the reference CPU
ufsecp_ecdsa_recover: every recovery id of reference-signed signatures, high-S
twins, out-of-range r/s and crafted ids 2/3 items with R.x = r + n.  The
argument checks follow the reference's ufsecp_gpu_impl.cpp:756-786.
"""
import random

import numpy as np
import pytest

from oracle import oracle as orc
from ultrafastsecp256k1_amd import UFSECP_ERR_BAD_INPUT, UFSECP_ERR_NULL_ARG, UFSECP_OK
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu


def _cat(items, key):
    return np.frombuffer(b"".join(bytes.fromhex(c[key]) for c in items), np.uint8)


def test_recover_golden(gpu, golden):
    d = golden["recover_vectors"]
    keys, valid = gpu.ecrecover_batch(_cat(d, "msg"), _cat(d, "sig"), [c["recid"] for c in d])
    bad = [c["name"] for c, k, v in zip(d, keys, valid) if (int(v), k.tobytes().hex()) != (c["valid"], c["key"])]
    assert not bad, bad


@pytest.mark.parametrize("n", [1, 65, 5000])
def test_recover_workload_vs_oracle(gpu, n):
    """ECDSA workload signatures (1% corrupted) under random recovery ids."""
    b = wl.generate("ecdsa", n, wl.SEED_ECDSA + 77 + n)
    rnd = random.Random(n)
    recids = np.array([rnd.randrange(4) if i % 4 else rnd.randrange(2) for i in range(n)], np.int32)
    keys, valid = gpu.ecrecover_batch(b.msgs, b.sigs, recids)
    for i in range(n):
        ok, key = orc.ecdsa_recover(b.msgs[i].tobytes(), b.sigs[i].tobytes(), int(recids[i]))
        assert (int(valid[i]), keys[i].tobytes()) == (ok, key), i
    # the signer's key is among the two low recovery ids of every valid item
    hits = 0
    for i in range(min(n, 200)):
        if b.corrupted[i]:
            continue
        cands = gpu.ecrecover_batch(np.repeat(b.msgs[i:i + 1], 2, 0), np.repeat(b.sigs[i:i + 1], 2, 0), [0, 1])[0]
        hits += any(c.tobytes() == b.pubkeys[i].tobytes() for c in cands)
    assert hits == min(n, 200) - int(b.corrupted[:min(n, 200)].sum())


def test_recover_argument_checks(gpu):
    lib, ctx = gpu.lib, gpu.ctx
    m = np.zeros(64, np.uint8)
    s = np.zeros(128, np.uint8)
    keys = np.full(66, 7, np.uint8)
    valid = np.full(2, 7, np.uint8)
    rec = np.array([0, 4], np.int32)
    f = lib.ufsecp_gpu_ecrecover_batch
    assert f(None, m.ctypes.data, s.ctypes.data, rec.ctypes.data, 2, keys.ctypes.data,
             valid.ctypes.data) == UFSECP_ERR_NULL_ARG
    assert f(ctx, m.ctypes.data, s.ctypes.data, rec.ctypes.data, 0, keys.ctypes.data, valid.ctypes.data) == UFSECP_OK
    # a recovery id outside 0..3 rejects the batch, outputs cleared
    assert f(ctx, m.ctypes.data, s.ctypes.data, rec.ctypes.data, 2, keys.ctypes.data,
             valid.ctypes.data) == UFSECP_ERR_BAD_INPUT
    assert not keys.any() and not valid.any()
    rec[1] = -1
    assert f(ctx, m.ctypes.data, s.ctypes.data, rec.ctypes.data, 2, keys.ctypes.data,
             valid.ctypes.data) == UFSECP_ERR_BAD_INPUT
    assert f(ctx, m.ctypes.data, s.ctypes.data, None, 2, keys.ctypes.data, valid.ctypes.data) == UFSECP_ERR_NULL_ARG
    assert f(ctx, m.ctypes.data, s.ctypes.data, rec.ctypes.data, (1 << 26) + 1, keys.ctypes.data,
             valid.ctypes.data) == UFSECP_ERR_BAD_INPUT
    # all-zero signatures are per-item failures, not errors
    rec[:] = 0
    assert f(ctx, m.ctypes.data, s.ctypes.data, rec.ctypes.data, 2, keys.ctypes.data, valid.ctypes.data) == UFSECP_OK
    assert not keys.any() and not valid.any()
