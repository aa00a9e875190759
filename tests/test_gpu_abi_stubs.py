"""The reference header's operations that this backend does not implement
(off the verify / MSM path) answer with the reference's argument checks and
then UFSECP_ERR_GPU_UNSUPPORTED, outputs cleared, on a live context; the
reference's own GPU C-ABI audit programs (audit/test_gpu_host_api_negative.cpp,
audit/test_gpu_abi_gate.cpp, compiled against this library by oracle/ref.mk)
pass with a GPU present, except the gate's check that backend ids are 1..3."""
import ctypes

import pytest

from ultrafastsecp256k1_amd import (UFSECP_ERR_BAD_INPUT, UFSECP_ERR_BAD_PUBKEY, UFSECP_ERR_GPU_UNSUPPORTED,
                                    UFSECP_ERR_NULL_ARG, UFSECP_OK)

pytestmark = pytest.mark.gpu

SZ = ctypes.c_size_t


def _buf(n, fill=0):
    return (ctypes.c_uint8 * n)(*([fill] * n))


def _calls(lib, ctx, count, out, inp):
    """(name, output bytes per item, call) for every unsupported op with
    `count` items, output `out` (cleared by the call) and input `inp`."""
    c = SZ(count)
    return [
        ("generator_mul_batch", 33, lambda: lib.ufsecp_gpu_generator_mul_batch(ctx, inp, c, out)),
        ("ecdh_batch", 32, lambda: lib.ufsecp_gpu_ecdh_batch(ctx, inp, inp, c, out)),
        ("hash160_pubkey_batch", 20, lambda: lib.ufsecp_gpu_hash160_pubkey_batch(ctx, inp, c, out)),
        ("tagged_hash", 32, lambda: lib.ufsecp_gpu_tagged_hash(ctx, inp, inp, SZ(32), c, out)),
        ("hash256", 32, lambda: lib.ufsecp_gpu_hash256(ctx, inp, SZ(64), c, out)),
        ("frost", 1, lambda: lib.ufsecp_gpu_frost_verify_partial_batch(ctx, inp, inp, inp, inp, inp, inp, inp, inp,
                                                                       c, out)),
        ("zk_ecdsa_snark", 760, lambda: lib.ufsecp_gpu_zk_ecdsa_snark_witness_batch(ctx, inp, inp, inp, c, out)),
    ]


def test_unsupported_ops_on_a_live_context(gpu):
    lib, ctx = gpu.lib, gpu.ctx
    inp = _buf(4096, 2)            # every "pubkey" prefix byte 02: passes the prefix checks
    out = _buf(1024, 0xEE)
    for name, width, call in _calls(lib, ctx, 1, out, inp):
        ctypes.memset(out, 0xEE, 1024)
        assert call() == UFSECP_ERR_GPU_UNSUPPORTED, name
        assert bytes(out)[:width] == bytes(width), name       # the item's result was cleared
        assert bytes(out)[width:] == b"\xee" * (1024 - width), name   # and nothing past it written
        assert lib.ufsecp_gpu_last_error(ctx) == UFSECP_ERR_GPU_UNSUPPORTED
        assert b"not supported" in lib.ufsecp_gpu_last_error_msg(ctx)
    for name, _, call in _calls(lib, ctx, 0, None, None):
        assert call() == UFSECP_OK, name                # count 0: OK, nothing read or written
    big = (1 << 26) + 1
    for name, _, call in _calls(lib, ctx, big, None, inp):
        assert call() == UFSECP_ERR_BAD_INPUT, name
    bad = _buf(4096, 5)            # prefix 05 everywhere
    assert lib.ufsecp_gpu_ecdh_batch(ctx, bad, bad, SZ(2), _buf(64)) == UFSECP_ERR_BAD_PUBKEY
    assert lib.ufsecp_gpu_hash160_pubkey_batch(ctx, bad, SZ(2), _buf(64)) == UFSECP_ERR_BAD_PUBKEY
    assert lib.ufsecp_gpu_hash160_pubkey_batch(ctx, None, SZ(2), _buf(64)) == UFSECP_ERR_NULL_ARG
    assert lib.ufsecp_gpu_tagged_hash(ctx, inp, inp, SZ(0), SZ(1), _buf(64)) == UFSECP_ERR_BAD_INPUT
    # the verify path still works after the unsupported calls
    assert gpu.lib.ufsecp_gpu_is_ready(ctx) == 1


def test_msm_count_zero_writes_nothing(gpu):
    """ufsecp_gpu_msm(ctx, NULL, NULL, 0, NULL) is OK (reference
    ufsecp_gpu_impl.cpp:541-558; audit/test_gpu_host_api_negative.cpp:145-150),
    and with buffers it leaves the output untouched."""
    lib, ctx = gpu.lib, gpu.ctx
    assert lib.ufsecp_gpu_msm(ctx, None, None, SZ(0), None) == UFSECP_OK
    out = _buf(33, 0x77)
    assert lib.ufsecp_gpu_msm(ctx, _buf(32), _buf(33), SZ(0), out) == UFSECP_OK
    assert bytes(out) == b"\x77" * 33


@pytest.mark.skipif(not __import__("pathlib").Path(__file__).resolve().parents[1].joinpath(
    "oracle", "_ref", "test_gpu_abi_gate").exists(), reason="reference audit programs not built")
def test_reference_abi_audit_programs_on_gpu():
    from test_abi_host import run_reference_audit
    p, f, fails = run_reference_audit("test_gpu_host_api_negative")
    assert f == 0 and p >= 60, fails
    p, f, fails = run_reference_audit("test_gpu_abi_gate")
    assert f == 1 and fails == ["FAIL: backend id in range [1,3]"], fails
