"""libbitcoin row / column verification on the GPU vs the reference CPU verdicts.

Entry points (include/ufsecp/ufsecp_gpu.h): ufsecp_gpu_ecdsa_verify_opaque_rows,
its ufsecp_gpu_ecdsa_verify_lbtc_rows alias, ufsecp_gpu_ecdsa_verify_lbtc_columns
and ufsecp_gpu_schnorr_verify_lbtc_columns.  Golden verdicts
(tests/golden/opaque_vectors.json) come from the reference CPU
ufsecp_ecdsa_verify_opaque_rows / ufsecp_ecdsa_verify_opaque_batch; the
argument checks follow the reference's audit/test_gpu_abi_gate.cpp:255-309.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from ultrafastsecp256k1_amd import (KIND_ECDSA, KIND_ECDSA_OPAQUE, KIND_SCHNORR, UFSECP_ERR_BAD_INPUT,
                                    UFSECP_ERR_NULL_ARG, UFSECP_OK)
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu

N = wl.N_ORDER


def _cols(items):
    m = np.frombuffer(b"".join(bytes.fromhex(c["msg"]) for c in items), np.uint8).reshape(-1, 32)
    p = np.frombuffer(b"".join(bytes.fromhex(c["pub"]) for c in items), np.uint8).reshape(-1, 33)
    s = np.frombuffer(b"".join(bytes.fromhex(c["sig"]) for c in items), np.uint8).reshape(-1, 64)
    return m, p, s


def _rows(m, p, s, stride, exact=False):
    """(count, stride) rows with a filled tail; exact=True trims the buffer to
    (count - 1) * stride + 129 bytes (no tail after the last row)."""
    n = m.shape[0]
    rows = np.full((n, stride), 0xC3, np.uint8)
    rows[:, :32], rows[:, 32:65], rows[:, 65:129] = m, p, s
    flat = rows.reshape(-1)
    return flat[:(n - 1) * stride + 129].copy() if exact else flat


def _opaque_workload(n, seed):
    """ECDSA workload items in the opaque form, every third with the high-S twin."""
    b = wl.generate("ecdsa", n, seed)
    sigs = np.empty_like(b.sigs)
    for i in range(n):
        r = int.from_bytes(b.sigs[i, :32].tobytes(), "big")
        s = int.from_bytes(b.sigs[i, 32:].tobytes(), "big")
        if i % 3 == 1 and 0 < s < N:
            s = N - s
        sigs[i] = np.frombuffer(r.to_bytes(32, "little") + s.to_bytes(32, "little"), np.uint8)
    return b.msgs, b.pubkeys, sigs


@pytest.mark.parametrize("stride", [129, 137, 160, 258, 259, 4096])
def test_opaque_rows_golden(gpu, golden, stride):
    items = golden["opaque_vectors"]["items"]
    want = [c["verdict"] for c in items]
    m, p, s = _cols(items)
    v = gpu.ecdsa_verify_opaque_rows(_rows(m, p, s, stride), stride, len(items))
    assert [int(x) for x in v] == want, [c["name"] for c, x in zip(items, v) if int(x) != c["verdict"]]
    v = gpu.ecdsa_verify_opaque_rows(_rows(m, p, s, stride, exact=True), stride, len(items), alias=True)
    assert [int(x) for x in v] == want


def test_lbtc_columns_golden(gpu, golden):
    items = golden["opaque_vectors"]["items"]
    m, p, s = _cols(items)
    v = gpu.ecdsa_verify_lbtc_columns(m, p, s)
    assert [int(x) for x in v] == [c["verdict"] for c in items]
    by = {c["name"]: int(x) for c, x in zip(items, v)}
    assert by["high_s"] == 1 and by["small_s_5_plus_n"] == 0


@pytest.mark.parametrize("n,stride", [(4096, 129), (3001, 144), (700, 515)])
def test_opaque_workload_matches_oracle(gpu, n, stride):
    m, p, s = _opaque_workload(n, wl.SEED_ECDSA + n)
    want = orc.verify_batch("ecdsa_opaque", m, p, s)
    assert 0 < want.sum() < n
    assert np.array_equal(gpu.ecdsa_verify_opaque_rows(_rows(m, p, s, stride), stride, n), want)
    assert np.array_equal(gpu.ecdsa_verify_lbtc_columns(m, p, s), want)


def test_schnorr_lbtc_columns(gpu, golden):
    cases = golden["bip340_vectors"]
    cat = lambda k: np.frombuffer(b"".join(bytes.fromhex(c[k]) for c in cases), np.uint8)  # noqa: E731
    v = gpu.schnorr_verify_lbtc_columns(cat("msg"), cat("pubkey"), cat("sig"))
    assert [int(x) for x in v] == [c["verdict"] for c in cases]
    edge = golden["edge_cases"]["schnorr"]
    cat = lambda k: np.frombuffer(b"".join(bytes.fromhex(c[k]) for c in edge), np.uint8)  # noqa: E731
    v = gpu.schnorr_verify_lbtc_columns(cat("msg"), cat("pub"), cat("sig"))
    assert [int(x) for x in v] == [c["verdict"] for c in edge]
    b = wl.generate("schnorr", 2048, wl.SEED_SCHNORR + 3)
    want = orc.verify_batch("schnorr", b.msgs, b.pubkeys, b.sigs)
    assert np.array_equal(gpu.schnorr_verify_lbtc_columns(b.msgs, b.pubkeys, b.sigs), want)


def test_row_argument_checks(gpu):
    lib, ctx = gpu.lib, gpu.ctx
    row = np.zeros(129, np.uint8)
    out = np.full(2, 7, np.uint8)
    for fn in (lib.ufsecp_gpu_ecdsa_verify_opaque_rows, lib.ufsecp_gpu_ecdsa_verify_lbtc_rows):
        assert fn(ctx, row.ctypes.data, 129, 0, out.ctypes.data) == UFSECP_OK
        assert fn(ctx, row.ctypes.data, 128, 1, out.ctypes.data) == UFSECP_ERR_BAD_INPUT
        assert fn(ctx, None, 129, 1, out.ctypes.data) == UFSECP_ERR_NULL_ARG
        assert fn(ctx, row.ctypes.data, 129, 1, None) == UFSECP_ERR_NULL_ARG
        assert fn(ctx, row.ctypes.data, 129, (1 << 26) + 1, out.ctypes.data) == UFSECP_ERR_BAD_INPUT
    # an all-zero row is a per-item 0, not an error; the output is written
    out[:] = 7
    assert lib.ufsecp_gpu_ecdsa_verify_opaque_rows(ctx, row.ctypes.data, 129, 1, out.ctypes.data) == UFSECP_OK
    assert out[0] == 0 and out[1] == 7
    for fn in (lib.ufsecp_gpu_ecdsa_verify_lbtc_columns, lib.ufsecp_gpu_schnorr_verify_lbtc_columns):
        assert fn(ctx, row.ctypes.data, row.ctypes.data, row.ctypes.data, 0, out.ctypes.data) == UFSECP_OK
        assert fn(ctx, None, row.ctypes.data, row.ctypes.data, 1, out.ctypes.data) == UFSECP_ERR_NULL_ARG
        assert fn(ctx, row.ctypes.data, row.ctypes.data, row.ctypes.data, (1 << 26) + 1,
                  out.ctypes.data) == UFSECP_ERR_BAD_INPUT


def test_opaque_kind_in_multi_batch(gpu, golden):
    """Opaque rows and columns as batches of one persistent launch beside
    ECDSA and BIP-340 batches: the same verdicts as the single calls."""
    items = golden["opaque_vectors"]["items"]
    m, p, s = _cols(items)
    want_o = np.array([c["verdict"] for c in items], np.uint8)
    rows = _rows(m, p, s, 150)
    e = wl.generate("ecdsa", 300, wl.SEED_ECDSA + 11)
    k = wl.generate("schnorr", 200, wl.SEED_SCHNORR + 11)
    want_e = orc.verify_batch("ecdsa", e.msgs, e.pubkeys, e.sigs)
    want_k = orc.verify_batch("schnorr", k.msgs, k.pubkeys, k.sigs)
    bufs = [gpu.to_device(a) for a in (rows, m, p, s, e.msgs, e.pubkeys, e.sigs, k.msgs, k.pubkeys, k.sigs)]
    outs = [gpu.empty(len(items)), gpu.empty(len(items)), gpu.empty(300), gpu.empty(200)]
    d_rows, d_m, d_p, d_s, d_em, d_ep, d_es, d_km, d_kp, d_ks = bufs
    n_o = len(items)
    table = gpu.batch_table([
        (KIND_ECDSA_OPAQUE, n_o, d_rows.ptr, d_rows.at(32), d_rows.at(65), outs[0].ptr, 150),
        (KIND_ECDSA, 300, d_em.ptr, d_ep.ptr, d_es.ptr, outs[2].ptr),
        (KIND_ECDSA_OPAQUE, n_o, d_m.ptr, d_p.ptr, d_s.ptr, outs[1].ptr),
        (KIND_SCHNORR, 200, d_km.ptr, d_kp.ptr, d_ks.ptr, outs[3].ptr),
    ])
    for mode in ("multi", "multi_graph", "each"):
        for o in outs:
            o.upload(np.full(o.nbytes, 9, np.uint8))
        if mode == "each":
            gpu.verify_each_device(table, 4)
        else:
            gpu.verify_multi_device(table, 4, use_graph=(mode == "multi_graph"))
        gpu.synchronize()
        assert np.array_equal(outs[0].download(), want_o), mode
        assert np.array_equal(outs[1].download(), want_o), mode
        assert np.array_equal(outs[2].download(), want_e), mode
        assert np.array_equal(outs[3].download(), want_k), mode
    bad = gpu.batch_table([(KIND_ECDSA, n_o, d_rows.ptr, d_rows.at(32), d_rows.at(65), outs[0].ptr, 150)])
    assert gpu.lib.ufsecp_gpu_ext_verify_multi_device(gpu.ctx, ctypes.cast(bad, ctypes.c_void_p), 1, 0,
                                                      None) == UFSECP_ERR_BAD_INPUT
    bad = gpu.batch_table([(KIND_ECDSA_OPAQUE, n_o, d_rows.ptr, d_rows.at(32), d_rows.at(65), outs[0].ptr, 100)])
    assert gpu.lib.ufsecp_gpu_ext_verify_multi_device(gpu.ctx, ctypes.cast(bad, ctypes.c_void_p), 1, 0,
                                                      None) == UFSECP_ERR_BAD_INPUT
    for x in bufs + outs:
        x.free()
