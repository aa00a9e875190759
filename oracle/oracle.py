"""ctypes view of the C oracle (oracle/liboracle.so) -- test infrastructure only.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never
by the product library.  See secp256k1_oracle.c for the reference functions
each entry point restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_LIB = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.is_file():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.orc_ecdsa_verify.argtypes = [vp, vp, vp]
        L.orc_schnorr_verify.argtypes = [vp, vp, vp]
        L.orc_msm.argtypes = [vp, vp, sz, vp]
        L.orc_pubkey33.argtypes = [vp, vp]
        L.orc_ecdsa_sign.argtypes = [vp, vp, vp, vp]
        L.orc_schnorr_sign.argtypes = [vp, vp, vp, vp, vp]
        L.orc_verify_batch.argtypes = [ctypes.c_int, vp, vp, vp, sz, vp, ctypes.c_int]
        L.orc_ecdsa_verify_opaque.argtypes = [vp, vp, vp]
        L.orc_ecdsa_recover.argtypes = [vp, vp, ctypes.c_int, vp]
        L.orc_pubkey_validate.argtypes = [vp]
        L.orc_xonly_validate.argtypes = [vp]
        L.orc_commitment_verify.argtypes = [vp, vp, vp, ctypes.c_int]
        _LIB = L
    return _LIB


def _p(b):
    if isinstance(b, np.ndarray):
        return ctypes.c_void_p(b.ctypes.data)
    if b is None:
        return None
    return ctypes.cast(ctypes.c_char_p(bytes(b)), ctypes.c_void_p)


def ecdsa_verify(msg: bytes, sig: bytes, pub33: bytes) -> int:
    return lib().orc_ecdsa_verify(msg, sig, pub33)


def ecdsa_verify_opaque(msg: bytes, opaque_sig: bytes, pub33: bytes) -> int:
    """Opaque-signature ECDSA (little-endian r, s; high-S accepted)."""
    return lib().orc_ecdsa_verify_opaque(msg, opaque_sig, pub33)


def ecdsa_recover(msg: bytes, sig64: bytes, recid: int):
    """(1, compressed key) or (0, 33 zero bytes)."""
    out = ctypes.create_string_buffer(33)
    ok = lib().orc_ecdsa_recover(bytes(msg), bytes(sig64), int(recid), out)
    return ok, out.raw


def pubkey_validate(pub33: bytes) -> int:
    return lib().orc_pubkey_validate(bytes(pub33))


def xonly_validate(x32: bytes) -> int:
    return lib().orc_xonly_validate(bytes(x32))


def commitment_verify(internal_x32: bytes, tweak32: bytes, tweaked_x32: bytes, parity: int) -> int:
    """BIP-341 tweak-add check: x(lift_x(internal) + t*G) == tweaked_x with the claimed parity."""
    return lib().orc_commitment_verify(bytes(internal_x32), bytes(tweak32), bytes(tweaked_x32), int(parity))


def compact_to_opaque(sig64: bytes) -> bytes:
    """Compact big-endian r||s -> opaque form (each scalar little-endian)."""
    return bytes(sig64[:32])[::-1] + bytes(sig64[32:64])[::-1]


def schnorr_verify(msg: bytes, sig: bytes, px: bytes) -> int:
    return lib().orc_schnorr_verify(msg, sig, px)


def msm(scalars: bytes, points: bytes, n: int):
    out = ctypes.create_string_buffer(33)
    rc = lib().orc_msm(bytes(scalars), bytes(points), n, out)
    return rc, out.raw


def pubkey33(d: bytes) -> bytes:
    out = ctypes.create_string_buffer(33)
    assert lib().orc_pubkey33(d, out) == 1
    return out.raw


def ecdsa_sign(msg: bytes, d: bytes):
    sig = ctypes.create_string_buffer(64)
    k = ctypes.create_string_buffer(32)
    assert lib().orc_ecdsa_sign(msg, d, sig, k) == 1
    return sig.raw, k.raw


def schnorr_sign(msg: bytes, d: bytes, aux: bytes):
    sig = ctypes.create_string_buffer(64)
    px = ctypes.create_string_buffer(32)
    assert lib().orc_schnorr_sign(msg, d, aux, sig, px) == 1
    return sig.raw, px.raw


def verify_batch(kind: str, msgs: np.ndarray, pubs: np.ndarray, sigs: np.ndarray, threads: int = 0) -> np.ndarray:
    n = msgs.shape[0]
    out = np.zeros(n, np.uint8)
    msgs, pubs, sigs = (np.ascontiguousarray(a) for a in (msgs, pubs, sigs))
    threads = threads or min(16, os.cpu_count() or 1)
    code = {"ecdsa": 0, "schnorr": 1, "ecdsa_opaque": 2}[kind]
    lib().orc_verify_batch(code, _p(msgs), _p(pubs), _p(sigs), n, _p(out), threads)
    return out


def verify_rows(rows: np.ndarray) -> np.ndarray:
    """Opaque ECDSA over (count, stride) rows: 32 digest | 33 pubkey | 64 sig."""
    rows = np.ascontiguousarray(rows)
    return verify_batch("ecdsa_opaque", rows[:, :32], rows[:, 32:65], rows[:, 65:129])


REF_LIB_PATH = HERE / "_ref" / "libufsecp_ref.so"


def ref_available() -> bool:
    return REF_LIB_PATH.is_file()


def ref_verify_batch(kind: str, msgs: np.ndarray, pubs: np.ndarray, sigs: np.ndarray, threads: int):
    """Per-item verification through the reference CPU library built by
    oracle/ref.mk (C threads, one ufsecp context each).  Returns (verdicts,
    wall seconds)."""
    os.environ.setdefault("SECP256K1_CACHE_DIR", "/tmp/ufsecp_ref_cache")
    os.makedirs(os.environ["SECP256K1_CACHE_DIR"], exist_ok=True)
    L = lib()
    fn = L.orc_ref_verify_batch
    fn.restype = ctypes.c_double
    fn.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int]
    msgs, pubs, sigs = (np.ascontiguousarray(a) for a in (msgs, pubs, sigs))
    n = msgs.shape[0]
    out = np.zeros(n, np.uint8)
    secs = fn(str(REF_LIB_PATH).encode(), 1 if kind == "schnorr" else 0, _p(msgs), _p(pubs), _p(sigs), n, _p(out),
              threads)
    if secs < 0:
        raise RuntimeError(f"reference driver failed ({secs})")
    return out, secs


def verify_range(kind: str, msgs, pubs, sigs) -> int:
    """Single-threaded verification of a slice; returns the number valid."""
    return int(verify_batch(kind, msgs, pubs, sigs, threads=1).sum())


def ref_verify_opaque(rows=None, stride=0, msgs=None, pubs=None, sigs=None, count=None):
    """Opaque ECDSA through the reference CPU library (oracle/_ref):
    ufsecp_ecdsa_verify_opaque_rows for `rows`, else
    ufsecp_ecdsa_verify_opaque_batch over the three columns.  Returns
    (ABI code, verdicts)."""
    L = lib()
    fn = L.orc_ref_verify_opaque
    vp = ctypes.c_void_p
    fn.argtypes = [ctypes.c_char_p, vp, ctypes.c_size_t, vp, vp, vp, ctypes.c_size_t, vp]
    if rows is not None:
        rows = np.ascontiguousarray(rows)
        n = rows.shape[0] if count is None else count
        out = np.zeros(max(n, 1), np.uint8)
        rc = fn(str(REF_LIB_PATH).encode(), _p(rows), stride, None, None, None, n, _p(out))
    else:
        msgs, pubs, sigs = (np.ascontiguousarray(a) for a in (msgs, pubs, sigs))
        n = msgs.shape[0] if count is None else count
        out = np.zeros(max(n, 1), np.uint8)
        rc = fn(str(REF_LIB_PATH).encode(), None, 0, _p(msgs), _p(pubs), _p(sigs), n, _p(out))
    if rc < 0:
        raise RuntimeError(f"reference driver failed ({rc})")
    return rc, out[:n]


def ref_compact_to_opaque(sig64: bytes):
    """The reference's ufsecp_ecdsa_sig_compact_to_opaque: (code, 64 bytes)."""
    L = lib()
    fn = L.orc_ref_compact_to_opaque
    fn.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
    out = ctypes.create_string_buffer(64)
    rc = fn(str(REF_LIB_PATH).encode(), _p(sig64), out)
    if rc < 0:
        raise RuntimeError(f"reference driver failed ({rc})")
    return rc, out.raw


def ref_msm(scalars: np.ndarray, points: np.ndarray, threads: int):
    """ufsecp_multi_scalar_mul of the reference CPU library (oracle/_ref) over
    `threads` contiguous chunks, partial sums added by one more call.
    Returns (ABI code, 33-byte result, wall seconds)."""
    os.environ.setdefault("SECP256K1_CACHE_DIR", "/tmp/ufsecp_ref_cache")
    os.makedirs(os.environ["SECP256K1_CACHE_DIR"], exist_ok=True)
    L = lib()
    fn = L.orc_ref_msm
    fn.restype = ctypes.c_double
    fn.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_int)]
    scalars, points = np.ascontiguousarray(scalars), np.ascontiguousarray(points)
    out = ctypes.create_string_buffer(33)
    rc = ctypes.c_int(-1)
    secs = fn(str(REF_LIB_PATH).encode(), _p(scalars), _p(points), scalars.shape[0], out, threads, ctypes.byref(rc))
    if secs < 0:
        raise RuntimeError(f"reference driver failed ({secs})")
    return rc.value, out.raw, secs
