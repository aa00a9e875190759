"""Host-side checks of the C-ABI library that need no GPU.

* the shared library loads and exports every function declared with
  UFSECP_API in include/ufsecp/*.h;
* discovery calls behave as the reference's (include/ufsecp/ufsecp_gpu.h:109-135,
  src/gpu/src/gpu_registry.cpp:65-77): ids other than ROCm are unavailable,
  backend 0 cannot create a context, NULL arguments give NULL_ARG;
* the workload generator's host logic (plan statistics, corruption schedule).
"""
import ctypes
import pathlib
import re

import numpy as np
import pytest

import ultrafastsecp256k1_amd as ufs
from ultrafastsecp256k1_amd import workload as wl

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in sorted((ROOT / "include" / "ufsecp").glob("*.h")):
        text = h.read_text()
        for m in re.finditer(r"UFSECP_API\s+[^;{]*?\b(ufsecp_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_reference_entry_points():
    names = declared_symbols()
    for ref_name in ("ufsecp_gpu_backend_count", "ufsecp_gpu_ctx_create", "ufsecp_gpu_ecdsa_verify_batch",
                     "ufsecp_gpu_schnorr_verify_batch", "ufsecp_gpu_msm", "ufsecp_gpu_ecdsa_verify_collect",
                     "ufsecp_gpu_schnorr_verify_collect", "ufsecp_gpu_error_str",
                     "ufsecp_gpu_ecdsa_verify_opaque_rows", "ufsecp_gpu_ecdsa_verify_lbtc_rows",
                     "ufsecp_gpu_ecdsa_verify_lbtc_columns", "ufsecp_gpu_schnorr_verify_lbtc_columns"):
        assert ref_name in names


def test_library_exports_every_declared_symbol():
    lib = ufs.load_library()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    for n in ufs.ABI_SYMBOLS:
        assert hasattr(lib, n), n


def test_discovery_without_context():
    lib = ufs.load_library()
    lib.ufsecp_gpu_backend_name.restype = ctypes.c_char_p
    assert lib.ufsecp_gpu_backend_name(4) == b"ROCm"
    assert lib.ufsecp_gpu_backend_name(1) == b"CUDA"
    assert lib.ufsecp_gpu_backend_name(2) == b"OpenCL"
    assert lib.ufsecp_gpu_backend_name(3) == b"Metal"
    for bid in (0, 1, 2, 3, 99):
        assert lib.ufsecp_gpu_is_available(bid) == 0
        assert lib.ufsecp_gpu_device_count(bid) == 0
    ids = (ctypes.c_uint32 * 4)()
    assert lib.ufsecp_gpu_backend_count(ids, 4) == 1
    assert ids[0] == 4
    assert lib.ufsecp_gpu_backend_count(None, 0) == 1
    ctx = ctypes.c_void_p(1)
    assert lib.ufsecp_gpu_ctx_create(ctypes.byref(ctx), 0, 0) == 100  # GPU_UNAVAILABLE
    assert ctx.value is None
    assert lib.ufsecp_gpu_ctx_create(None, 4, 0) == 1  # NULL_ARG
    lib.ufsecp_gpu_error_str.restype = ctypes.c_char_p
    assert lib.ufsecp_gpu_error_str(0) == b"OK"
    assert lib.ufsecp_gpu_error_str(103)


def test_null_context_calls():
    lib = ufs.load_library()
    out = (ctypes.c_uint8 * 4)(7, 7, 7, 7)
    buf = (ctypes.c_uint8 * 256)()
    assert lib.ufsecp_gpu_ecdsa_verify_batch(None, buf, buf, buf, ctypes.c_size_t(4), out) == 1
    assert lib.ufsecp_gpu_schnorr_verify_batch(None, buf, buf, buf, ctypes.c_size_t(4), out) == 1
    assert lib.ufsecp_gpu_msm(None, buf, buf, ctypes.c_size_t(1), out) == 1
    # row / column variants (reference audit/test_gpu_abi_gate.cpp:147-150)
    assert lib.ufsecp_gpu_ecdsa_verify_opaque_rows(None, buf, ctypes.c_size_t(129), ctypes.c_size_t(1), out) == 1
    assert lib.ufsecp_gpu_ecdsa_verify_lbtc_rows(None, buf, ctypes.c_size_t(129), ctypes.c_size_t(1), out) == 1
    assert lib.ufsecp_gpu_ecdsa_verify_lbtc_columns(None, buf, buf, buf, ctypes.c_size_t(1), out) == 1
    assert lib.ufsecp_gpu_schnorr_verify_lbtc_columns(None, buf, buf, buf, ctypes.c_size_t(1), out) == 1
    assert lib.ufsecp_gpu_is_ready(None) == 0
    assert lib.ufsecp_gpu_last_error(None) == 1


def test_mix_plan_statistics():
    plan = wl.mix_plan(10_000, 4096, wl.SEED_MIX)
    kinds = np.array([k for k, _ in plan])
    sizes = np.array([s for _, s in plan])
    assert len(plan) == 10_000
    assert sizes.min() >= 1 and sizes.max() <= 4096
    # log-uniform on [1, 4096]: mean (4096 - 1) / ln(4096) ~ 492
    assert 440 < sizes.mean() < 545
    assert 0.67 < (kinds == ufs.KIND_ECDSA).mean() < 0.73
    # the plan is a pure function of its seed
    assert wl.mix_plan(10_000, 4096, wl.SEED_MIX) == plan


def test_corruption_schedule_rate():
    b = wl.generate("schnorr", 20_000, wl.SEED_SCHNORR)
    rate = b.corrupted.astype(float).mean()
    assert 0.005 < rate < 0.015
    # item i of a batch equals item 0 of the batch started at i
    tail = wl.generate("schnorr", 4, wl.SEED_SCHNORR, first=777)
    assert np.array_equal(tail.sigs, b.sigs[777:781])
    assert np.array_equal(tail.pubkeys, b.pubkeys[777:781])
    assert np.array_equal(tail.corrupted, b.corrupted[777:781])


def test_batch_check_null_context():
    lib = ufs.load_library()
    ok = ctypes.c_int(5)
    buf = (ctypes.c_uint8 * 128)()
    assert lib.ufsecp_gpu_ext_schnorr_batch_check(None, buf, buf, buf, ctypes.c_size_t(1), ctypes.byref(ok)) == 1


REF_GPU_HEADER = pathlib.Path("/root/reference/include/ufsecp/ufsecp_gpu.h")


@pytest.mark.skipif(not REF_GPU_HEADER.exists(), reason="reference sources not mounted")
def test_every_reference_declaration_is_declared_and_exported():
    """Every UFSECP_API function of the reference's include/ufsecp/ufsecp_gpu.h
    is declared in include/ufsecp/ufsecp_gpu.h and exported by the library, so
    a caller compiled or bound against the reference header links here."""
    ref = set(re.findall(r"UFSECP_API\s+[^;(]*?\b(ufsecp_\w+)\s*\(", REF_GPU_HEADER.read_text(), re.S))
    assert len(ref) >= 41
    assert not ref - declared_symbols()
    lib = ufs.load_library()
    assert [n for n in sorted(ref) if not hasattr(lib, n)] == []


def test_unsupported_ops_checks_without_context():
    """Operations off the verify / MSM path keep the reference's argument
    checks (src/cpu/src/ufsecp_gpu_impl.cpp:299-1127): NULL ctx -> NULL_ARG,
    except the SNARK witness calls, which test count first (count 0 -> OK,
    count > 2^26 -> BAD_INPUT, then NULL -> NULL_ARG)."""
    lib = ufs.load_library()
    b = (ctypes.c_uint8 * 1024)()
    one, big = ctypes.c_size_t(1), ctypes.c_size_t((1 << 26) + 1)
    assert lib.ufsecp_gpu_generator_mul_batch(None, b, one, b) == 1
    assert lib.ufsecp_gpu_ecdh_batch(None, b, b, one, b) == 1
    assert lib.ufsecp_gpu_hash160_pubkey_batch(None, b, one, b) == 1
    assert lib.ufsecp_gpu_tagged_hash(None, b, b, ctypes.c_size_t(32), one, b) == 1
    assert lib.ufsecp_gpu_tagged_hash_var(None, b, b, b, ctypes.c_size_t(32), one, b) == 1
    assert lib.ufsecp_gpu_hash256(None, b, ctypes.c_size_t(64), one, b) == 1
    assert lib.ufsecp_gpu_hash256_var(None, b, b, ctypes.c_size_t(64), one, b) == 1
    assert lib.ufsecp_gpu_frost_verify_partial_batch(None, b, b, b, b, b, b, b, b, one, b) == 1
    assert lib.ufsecp_gpu_zk_knowledge_verify_batch(None, b, b, b, one, b) == 1
    assert lib.ufsecp_gpu_zk_dleq_verify_batch(None, b, b, b, b, b, one, b) == 1
    assert lib.ufsecp_gpu_bulletproof_verify_batch(None, b, b, b, one, b) == 1
    assert lib.ufsecp_gpu_bip324_aead_encrypt_batch(None, b, b, b, b, ctypes.c_uint32(8), one, b) == 1
    assert lib.ufsecp_gpu_bip324_aead_decrypt_batch(None, b, b, b, b, ctypes.c_uint32(8), one, b, b) == 1
    assert lib.ufsecp_gpu_bip352_scan_batch(None, b, b, b, one, b) == 1
    for fn in (lib.ufsecp_gpu_zk_ecdsa_snark_witness_batch, lib.ufsecp_gpu_zk_schnorr_snark_witness_batch):
        assert fn(None, b, b, b, ctypes.c_size_t(0), b) == 0
        assert fn(None, b, b, b, big, None) == 5
        assert fn(None, b, b, b, one, None) == 1
    plan = (ctypes.c_uint8 * 264)(*([0xAB] * 264))
    assert lib.ufsecp_bip352_prepare_scan_plan(None, plan) == 1
    assert lib.ufsecp_bip352_prepare_scan_plan(b, plan) == ufs.UFSECP_ERR_GPU_UNSUPPORTED
    assert bytes(plan) == bytes(264)


AUDIT_BIN = ROOT / "oracle" / "_ref"


def run_reference_audit(name):
    """Runs one of the reference's GPU C-ABI audit programs (compiled from its
    audit/ sources by oracle/ref.mk against this library); returns (passed,
    failed, FAIL lines)."""
    import os
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        env = dict(os.environ, SECP256K1_CACHE_DIR=tmp)
        res = subprocess.run([str(AUDIT_BIN / name)], cwd=tmp, env=env, capture_output=True, text=True, timeout=300)
    m = re.search(r"Results: (\d+) passed, (\d+) failed", res.stdout)
    assert m, res.stdout[-2000:] + res.stderr[-2000:]
    fails = [ln.strip() for ln in res.stdout.splitlines() if "FAIL" in ln]
    return int(m.group(1)), int(m.group(2)), fails


@pytest.mark.skipif(not (AUDIT_BIN / "test_gpu_abi_gate").exists(), reason="reference audit programs not built")
def test_reference_abi_audit_programs_host():
    """The reference's audit/test_gpu_host_api_negative.cpp passes in full
    against this library; audit/test_gpu_abi_gate.cpp fails only its check
    that every backend id is 1..3 (this library adds id 4, ROCm).  Without a
    GPU both skip their context sections."""
    p, f, _ = run_reference_audit("test_gpu_host_api_negative")
    assert f == 0 and p >= 30
    p, f, fails = run_reference_audit("test_gpu_abi_gate")
    assert f == 1 and fails == ["FAIL: backend id in range [1,3]"]
