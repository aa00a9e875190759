"""MI355X-native secp256k1 batch verification and MSM.

Python host-side mirror of the ufsecp GPU C ABI (``include/ufsecp/ufsecp_gpu.h``)
over ``libufsecp_gpu.so``, in the style of the reference's ctypes wrapper
(bindings/python/ufsecp/__init__.py:643-834, class ``UfsecpGpu``).  Everything
here is plumbing: the arithmetic runs in the HIP kernels of the library, and
the wrapper raises if the library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_int, c_size_t, c_uint32, c_void_p, c_char_p, c_uint8
from pathlib import Path
from typing import Optional

PKG_DIR = Path(__file__).resolve().parent
GPU_LIB_PATH = PKG_DIR / "libufsecp_gpu.so"
WORKLOAD_LIB_PATH = PKG_DIR / "libufsecp_workload.so"

UFSECP_OK = 0
UFSECP_ERR_NULL_ARG = 1
UFSECP_ERR_BAD_PUBKEY = 3
UFSECP_ERR_BAD_INPUT = 5
UFSECP_ERR_ARITH = 7
UFSECP_ERR_INTERNAL = 9
UFSECP_ERR_GPU_UNAVAILABLE = 100
UFSECP_ERR_GPU_DEVICE = 101
UFSECP_ERR_GPU_MEMORY = 103
UFSECP_ERR_GPU_UNSUPPORTED = 104

BACKEND_ROCM = 4

# exported symbols of libufsecp_gpu.so, as declared in include/ufsecp/*.h
ABI_SYMBOLS = (
    "ufsecp_gpu_backend_count", "ufsecp_gpu_backend_name", "ufsecp_gpu_is_available",
    "ufsecp_gpu_device_count", "ufsecp_gpu_device_info", "ufsecp_gpu_ctx_create",
    "ufsecp_gpu_ctx_destroy", "ufsecp_gpu_is_ready", "ufsecp_gpu_last_error",
    "ufsecp_gpu_last_error_msg", "ufsecp_gpu_error_str", "ufsecp_gpu_ecdsa_verify_batch",
    "ufsecp_gpu_schnorr_verify_batch", "ufsecp_gpu_ecdsa_verify_opaque_rows", "ufsecp_gpu_ecdsa_verify_lbtc_rows",
    "ufsecp_gpu_ecdsa_verify_lbtc_columns", "ufsecp_gpu_schnorr_verify_lbtc_columns", "ufsecp_gpu_ecdsa_verify_collect",
    "ufsecp_gpu_schnorr_verify_collect", "ufsecp_gpu_msm",
    "ufsecp_gpu_ext_ecdsa_verify_device", "ufsecp_gpu_ext_schnorr_verify_device",
    "ufsecp_gpu_ext_stream", "ufsecp_gpu_ext_synchronize", "ufsecp_gpu_ext_selftest",
    "ufsecp_gpu_ext_verify_multi_device", "ufsecp_gpu_ext_verify_each_device",
    "ufsecp_gpu_ext_msm", "ufsecp_gpu_ext_msm_device",
    "ufsecp_gpu_ext_alloc", "ufsecp_gpu_ext_free", "ufsecp_gpu_ext_copy_to_device", "ufsecp_gpu_ext_copy_to_host",
    "ufsecp_gpu_ext_event_create", "ufsecp_gpu_ext_event_record", "ufsecp_gpu_ext_event_elapsed",
    "ufsecp_gpu_ext_event_destroy", "ufsecp_gpu_ext_host_alloc", "ufsecp_gpu_ext_host_free",
    "ufsecp_gpu_ext_server_create", "ufsecp_gpu_ext_server_destroy", "ufsecp_gpu_ext_server_submit",
    "ufsecp_gpu_ext_server_submit_mapped",
    "ufsecp_gpu_ext_server_poll", "ufsecp_gpu_ext_server_wait", "ufsecp_gpu_ext_server_drain",
    "ufsecp_gpu_ext_server_stats", "ufsecp_gpu_ext_engine_columns_hook", "ufsecp_gpu_ext_engine_hook_install",
    "ufsecp_gpu_ext_engine_hook_stats", "ufsecp_gpu_pubkey_validate", "ufsecp_gpu_xonly_validate",
    "ufsecp_gpu_commitment_verify", "ufsecp_gpu_ecrecover_batch", "ufsecp_gpu_ext_schnorr_batch_check",
)


class VerifyBatchDesc(ctypes.Structure):
    """ufsecp_gpu_verify_batch_t (include/ufsecp/ufsecp_gpu_ext.h); device pointers."""
    _fields_ = [("kind", c_uint32), ("count", c_uint32), ("msg_hashes32", c_void_p),
                ("pubkeys", c_void_p), ("sigs64", c_void_p), ("out_results", c_void_p),
                ("row_stride", c_uint32), ("reserved", c_uint32)]


KIND_ECDSA = 0
KIND_SCHNORR = 1
KIND_ECDSA_OPAQUE = 2


class ServerConfig(ctypes.Structure):
    """ufsecp_gpu_server_config_t (include/ufsecp/ufsecp_gpu_ext.h)."""
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("inflight_groups", c_uint32),
                ("key_cache_log2", c_uint32), ("max_group_items", c_uint32), ("reserved", c_uint32)]


class ServerStats(ctypes.Structure):
    """ufsecp_gpu_server_stats_t (include/ufsecp/ufsecp_gpu_ext.h)."""
    _fields_ = [("requests_completed", ctypes.c_uint64), ("items_completed", ctypes.c_uint64),
                ("groups_launched", ctypes.c_uint64), ("key_lookups", ctypes.c_uint64),
                ("key_cache_hits", ctypes.c_uint64), ("latency_p50_us", ctypes.c_double),
                ("latency_p99_us", ctypes.c_double), ("latency_mean_us", ctypes.c_double),
                ("latency_max_us", ctypes.c_double), ("busy_seconds", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class EngineHookStats(ctypes.Structure):
    """ufsecp_gpu_engine_hook_stats_t (include/ufsecp/ufsecp_gpu_ext.h)."""
    _fields_ = [("calls", ctypes.c_uint64), ("handled", ctypes.c_uint64), ("declined", ctypes.c_uint64),
                ("items", ctypes.c_uint64)]


def engine_hook_stats(lib: Optional[ctypes.CDLL] = None) -> dict:
    """Counters of the reference-engine column hook served by this library."""
    lib = lib or load_library()
    st = EngineHookStats()
    lib.ufsecp_gpu_ext_engine_hook_stats(byref(st))
    return {k: int(getattr(st, k)) for k, _ in EngineHookStats._fields_}


class UfsecpError(RuntimeError):
    def __init__(self, op: str, code: int, msg: str = ""):
        self.op = op
        self.code = code
        super().__init__(f"{op} failed: code {code}{(': ' + msg) if msg else ''}")


_LIB: Optional[ctypes.CDLL] = None


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    """Load libufsecp_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path or os.environ.get("UFSECP_GPU_LIB") or GPU_LIB_PATH)
    if not p.is_file():
        raise FileNotFoundError(
            f"{p} not found: build it with `python -m ultrafastsecp256k1_amd.build`")
    lib = ctypes.CDLL(str(p))
    _bind(lib)
    if path is None:
        _LIB = lib
    return lib


def _bind(lib: ctypes.CDLL) -> None:
    u8p = POINTER(c_uint8)
    vp = c_void_p
    sig = {
        "ufsecp_gpu_backend_count": (c_uint32, [POINTER(c_uint32), c_uint32]),
        "ufsecp_gpu_backend_name": (c_char_p, [c_uint32]),
        "ufsecp_gpu_is_available": (c_int, [c_uint32]),
        "ufsecp_gpu_device_count": (c_uint32, [c_uint32]),
        "ufsecp_gpu_ctx_create": (c_int, [POINTER(vp), c_uint32, c_uint32]),
        "ufsecp_gpu_ctx_destroy": (None, [vp]),
        "ufsecp_gpu_is_ready": (c_int, [vp]),
        "ufsecp_gpu_last_error": (c_int, [vp]),
        "ufsecp_gpu_last_error_msg": (c_char_p, [vp]),
        "ufsecp_gpu_error_str": (c_char_p, [c_int]),
        "ufsecp_gpu_ecdsa_verify_batch": (c_int, [vp, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_schnorr_verify_batch": (c_int, [vp, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_ecdsa_verify_opaque_rows": (c_int, [vp, vp, c_size_t, c_size_t, vp]),
        "ufsecp_gpu_ecdsa_verify_lbtc_rows": (c_int, [vp, vp, c_size_t, c_size_t, vp]),
        "ufsecp_gpu_ecdsa_verify_lbtc_columns": (c_int, [vp, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_schnorr_verify_lbtc_columns": (c_int, [vp, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_pubkey_validate": (c_int, [vp, vp, c_size_t, vp]),
        "ufsecp_gpu_xonly_validate": (c_int, [vp, vp, c_size_t, vp]),
        "ufsecp_gpu_commitment_verify": (c_int, [vp, vp, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_ecrecover_batch": (c_int, [vp, vp, vp, vp, c_size_t, vp, vp]),
        "ufsecp_gpu_ext_schnorr_batch_check": (c_int, [vp, vp, vp, vp, c_size_t, POINTER(c_int)]),
        "ufsecp_gpu_ecdsa_verify_collect": (c_int, [vp, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_schnorr_verify_collect": (c_int, [vp, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_msm": (c_int, [vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_ext_ecdsa_verify_device": (c_int, [vp, vp, vp, vp, c_size_t, vp, vp]),
        "ufsecp_gpu_ext_schnorr_verify_device": (c_int, [vp, vp, vp, vp, c_size_t, vp, vp]),
        "ufsecp_gpu_ext_verify_multi_device": (c_int, [vp, vp, c_size_t, c_int, vp]),
        "ufsecp_gpu_ext_verify_each_device": (c_int, [vp, vp, c_size_t, vp]),
        "ufsecp_gpu_ext_msm": (c_int, [vp, vp, vp, c_size_t, vp, c_int]),
        "ufsecp_gpu_ext_msm_device": (c_int, [vp, vp, vp, c_size_t, vp, vp, c_int, vp]),
        "ufsecp_gpu_ext_msm_affine_device": (c_int, [vp, vp, vp, c_size_t, vp, vp, c_int, vp]),
        "ufsecp_gpu_ext_msm_device_noalloc": (c_int, [vp, vp, vp, c_size_t, vp, vp, c_int, vp]),
        "ufsecp_gpu_ext_buffer_alloc_count": (ctypes.c_uint64, []),
        "ufsecp_gpu_ext_buffer_live_count": (ctypes.c_int64, []),
        "ufsecp_gpu_ext_alloc": (c_int, [vp, c_size_t, POINTER(vp)]),
        "ufsecp_gpu_ext_free": (c_int, [vp, vp]),
        "ufsecp_gpu_ext_copy_to_device": (c_int, [vp, vp, vp, c_size_t]),
        "ufsecp_gpu_ext_copy_to_host": (c_int, [vp, vp, vp, c_size_t]),
        "ufsecp_gpu_ext_event_create": (c_int, [vp, POINTER(vp)]),
        "ufsecp_gpu_ext_event_record": (c_int, [vp, vp, vp]),
        "ufsecp_gpu_ext_event_elapsed": (c_int, [vp, vp, POINTER(ctypes.c_float)]),
        "ufsecp_gpu_ext_event_destroy": (None, [vp]),
        "ufsecp_gpu_ext_stream": (vp, [vp]),
        "ufsecp_gpu_ext_synchronize": (c_int, [vp]),
        "ufsecp_gpu_ext_verify_lanes": (c_uint32, [vp, c_uint32]),
        "ufsecp_gpu_ext_stream_create": (c_int, [vp, POINTER(vp)]),
        "ufsecp_gpu_ext_stream_synchronize": (c_int, [vp]),
        "ufsecp_gpu_ext_stream_destroy": (None, [vp]),
        "ufsecp_gpu_ext_selftest": (c_int, [vp, c_int, vp, vp, vp, c_size_t]),
        "ufsecp_gpu_ext_host_alloc": (c_int, [vp, c_size_t, POINTER(vp)]),
        "ufsecp_gpu_ext_host_free": (c_int, [vp, vp]),
        "ufsecp_gpu_ext_server_create": (c_int, [vp, vp, POINTER(vp)]),
        "ufsecp_gpu_ext_server_destroy": (None, [vp]),
        "ufsecp_gpu_ext_server_submit": (c_int, [vp, c_uint32, vp, vp, vp, c_size_t, vp,
                                                 POINTER(ctypes.c_uint64)]),
        "ufsecp_gpu_ext_server_submit_mapped": (c_int, [vp, c_uint32, vp, vp, vp, c_size_t, vp,
                                                        POINTER(ctypes.c_uint64)]),
        "ufsecp_gpu_ext_server_poll": (c_int, [vp, ctypes.c_uint64]),
        "ufsecp_gpu_ext_server_wait": (c_int, [vp, ctypes.c_uint64]),
        "ufsecp_gpu_ext_server_drain": (c_int, [vp]),
        "ufsecp_gpu_ext_server_stats": (c_int, [vp, vp, c_int]),
        "ufsecp_gpu_ext_engine_columns_hook": (c_int, [c_int, vp, vp, vp, c_size_t, vp]),
        "ufsecp_gpu_ext_engine_hook_install": (c_int, []),
        "ufsecp_gpu_ext_engine_hook_stats": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _ = u8p


def _nitems(b, stride: int) -> int:
    """Number of fixed-stride items in a buffer (byte size // stride)."""
    nbytes = getattr(b, "nbytes", None)
    return (nbytes if nbytes is not None else len(b)) // stride


def _buf(b) -> c_void_p:
    """Pointer to the start of a bytes-like / numpy buffer (None -> NULL)."""
    if b is None:
        return c_void_p(None)
    if isinstance(b, int):
        return c_void_p(b)
    try:
        import numpy as np
        if isinstance(b, np.ndarray):
            return c_void_p(b.ctypes.data)
    except ImportError:  # pragma: no cover
        pass
    if isinstance(b, bytes):
        return ctypes.cast(ctypes.c_char_p(b), c_void_p)
    if isinstance(b, (bytearray, memoryview)):
        arr = (ctypes.c_char * len(b)).from_buffer(b)
        return ctypes.cast(arr, c_void_p)
    raise TypeError(f"unsupported buffer type {type(b)}")


class UfsecpGpu:
    """One GPU context (backend 4 = ROCm) with the batch operations of the ABI.

    Usage::

        with UfsecpGpu() as gpu:
            verdicts = gpu.schnorr_verify_batch(msgs, xonly, sigs)
    """

    def __init__(self, device_index: int = 0, backend_id: int = BACKEND_ROCM,
                 lib_path: Optional[str] = None):
        self._lib = load_library(lib_path)
        ctx = c_void_p()
        rc = self._lib.ufsecp_gpu_ctx_create(byref(ctx), c_uint32(backend_id), c_uint32(device_index))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ctx_create", rc)
        self._ctx = ctx

    # -- lifecycle -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.ufsecp_gpu_ctx_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self) -> c_void_p:
        return self._ctx

    @property
    def lib(self) -> ctypes.CDLL:
        return self._lib

    def last_error_msg(self) -> str:
        return self._lib.ufsecp_gpu_last_error_msg(self._ctx).decode()

    def _check(self, op: str, rc: int) -> None:
        if rc != UFSECP_OK:
            raise UfsecpError(op, rc, self.last_error_msg())

    # -- batch verification ---------------------------------------------------
    def ecdsa_verify_batch(self, msgs, pubkeys33, sigs64, count: Optional[int] = None):
        import numpy as np
        n = _nitems(msgs, 32) if count is None else count
        out = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_ecdsa_verify_batch(self._ctx, _buf(msgs), _buf(pubkeys33), _buf(sigs64), n,
                                                     _buf(out))
        self._check("ufsecp_gpu_ecdsa_verify_batch", rc)
        return out

    def schnorr_verify_batch(self, msgs, pubkeys_x32, sigs64, count: Optional[int] = None):
        import numpy as np
        n = _nitems(msgs, 32) if count is None else count
        out = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_schnorr_verify_batch(self._ctx, _buf(msgs), _buf(pubkeys_x32), _buf(sigs64), n,
                                                       _buf(out))
        self._check("ufsecp_gpu_schnorr_verify_batch", rc)
        return out

    def ecdsa_verify_opaque_rows(self, rows, stride: int, count: Optional[int] = None, alias: bool = False):
        """libbitcoin rows: 32 digest | 33 pubkey | 64 opaque sig | tail, `stride` bytes each
        (ufsecp_gpu_ecdsa_verify_opaque_rows, or its ufsecp_gpu_ecdsa_verify_lbtc_rows alias)."""
        import numpy as np
        if count is None:
            nbytes = getattr(rows, "nbytes", None)
            nbytes = nbytes if nbytes is not None else len(rows)
            count = (nbytes - 129) // stride + 1 if nbytes >= 129 and stride else 0
        out = np.zeros(max(count, 1), dtype=np.uint8)
        name = "ufsecp_gpu_ecdsa_verify_lbtc_rows" if alias else "ufsecp_gpu_ecdsa_verify_opaque_rows"
        rc = getattr(self._lib, name)(self._ctx, _buf(rows), stride, count, _buf(out))
        self._check(name, rc)
        return out[:count]

    def ecdsa_verify_lbtc_columns(self, digests32, pubkeys33, opaque_sigs64, count: Optional[int] = None):
        import numpy as np
        n = _nitems(digests32, 32) if count is None else count
        out = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_ecdsa_verify_lbtc_columns(self._ctx, _buf(digests32), _buf(pubkeys33),
                                                            _buf(opaque_sigs64), n, _buf(out))
        self._check("ufsecp_gpu_ecdsa_verify_lbtc_columns", rc)
        return out

    def schnorr_verify_lbtc_columns(self, digests32, xonly32, sigs64, count: Optional[int] = None):
        import numpy as np
        n = _nitems(digests32, 32) if count is None else count
        out = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_schnorr_verify_lbtc_columns(self._ctx, _buf(digests32), _buf(xonly32),
                                                              _buf(sigs64), n, _buf(out))
        self._check("ufsecp_gpu_schnorr_verify_lbtc_columns", rc)
        return out

    # -- public-data key operations -------------------------------------------
    def pubkey_validate(self, pubkeys33, n: Optional[int] = None):
        import numpy as np
        n = _nitems(pubkeys33, 33) if n is None else n
        out = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_pubkey_validate(self._ctx, _buf(pubkeys33), n, _buf(out))
        self._check("ufsecp_gpu_pubkey_validate", rc)
        return out

    def xonly_validate(self, keys32, n: Optional[int] = None):
        import numpy as np
        n = _nitems(keys32, 32) if n is None else n
        out = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_xonly_validate(self._ctx, _buf(keys32), n, _buf(out))
        self._check("ufsecp_gpu_xonly_validate", rc)
        return out

    def commitment_verify(self, internal_x32, tweak32, tweaked_x32, parity, n: Optional[int] = None):
        import numpy as np
        n = _nitems(internal_x32, 32) if n is None else n
        out = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_commitment_verify(self._ctx, _buf(internal_x32), _buf(tweak32),
                                                    _buf(tweaked_x32), _buf(parity), n, _buf(out))
        self._check("ufsecp_gpu_commitment_verify", rc)
        return out

    def schnorr_batch_check(self, msgs, pubkeys_x32, sigs64, count: Optional[int] = None) -> bool:
        """True when every BIP-340 signature verifies (random linear combination, one MSM)."""
        n = _nitems(msgs, 32) if count is None else count
        ok = c_int(-1)
        rc = self._lib.ufsecp_gpu_ext_schnorr_batch_check(self._ctx, _buf(msgs), _buf(pubkeys_x32), _buf(sigs64), n,
                                                          byref(ok))
        self._check("ufsecp_gpu_ext_schnorr_batch_check", rc)
        return ok.value == 1

    def ecrecover_batch(self, msgs, sigs64, recids, count: Optional[int] = None):
        """(keys (count, 33) uint8, valid (count,) uint8); recids int32 0..3."""
        import numpy as np
        n = _nitems(msgs, 32) if count is None else count
        rec = np.ascontiguousarray(np.asarray(recids, dtype=np.int32))
        keys = np.zeros((n, 33), dtype=np.uint8)
        valid = np.zeros(n, dtype=np.uint8)
        rc = self._lib.ufsecp_gpu_ecrecover_batch(self._ctx, _buf(msgs), _buf(sigs64), _buf(rec), n, _buf(keys),
                                                  _buf(valid))
        self._check("ufsecp_gpu_ecrecover_batch", rc)
        return keys, valid

    def msm(self, scalars32, points33, n: Optional[int] = None) -> bytes:
        n = _nitems(scalars32, 32) if n is None else n
        out = ctypes.create_string_buffer(33)
        rc = self._lib.ufsecp_gpu_msm(self._ctx, _buf(scalars32), _buf(points33), n, out)
        self._check("ufsecp_gpu_msm", rc)
        return out.raw

    def msm_rc(self, scalars32, points33, n: Optional[int] = None, window_bits: Optional[int] = None):
        """(return code, 33-byte output) without raising; window_bits forces c."""
        n = _nitems(scalars32, 32) if n is None else n
        out = ctypes.create_string_buffer(33)
        if window_bits is None:
            rc = self._lib.ufsecp_gpu_msm(self._ctx, _buf(scalars32), _buf(points33), n, out)
        else:
            rc = self._lib.ufsecp_gpu_ext_msm(self._ctx, _buf(scalars32), _buf(points33), n, out, window_bits)
        return rc, out.raw

    def msm_device(self, d_scalars: int, d_points: int, n: int, d_out33: int, d_status: int, window_bits: int = 0,
                   stream: Optional[int] = None) -> None:
        rc = self._lib.ufsecp_gpu_ext_msm_device(self._ctx, c_void_p(d_scalars), c_void_p(d_points), n,
                                                 c_void_p(d_out33), c_void_p(d_status), window_bits,
                                                 c_void_p(stream))
        self._check("ufsecp_gpu_ext_msm_device", rc)

    def msm_affine_device(self, d_scalars: int, d_points64: int, n: int, d_out33: int, d_status: int,
                          window_bits: int = 0, stream: Optional[int] = None) -> None:
        """msm_device over pre-decompressed bases (64-byte big-endian x || y)."""
        rc = self._lib.ufsecp_gpu_ext_msm_affine_device(self._ctx, c_void_p(d_scalars), c_void_p(d_points64), n,
                                                        c_void_p(d_out33), c_void_p(d_status), window_bits,
                                                        c_void_p(stream))
        self._check("ufsecp_gpu_ext_msm_affine_device", rc)

    # -- device-resident variants (pointers are device addresses) --------------
    def schnorr_verify_device(self, d_msgs: int, d_pubx: int, d_sigs: int, count: int, d_out: int,
                              stream: Optional[int] = None) -> None:
        rc = self._lib.ufsecp_gpu_ext_schnorr_verify_device(self._ctx, c_void_p(d_msgs), c_void_p(d_pubx),
                                                            c_void_p(d_sigs), count, c_void_p(d_out),
                                                            c_void_p(stream))
        self._check("ufsecp_gpu_ext_schnorr_verify_device", rc)

    def ecdsa_verify_device(self, d_msgs: int, d_pubs: int, d_sigs: int, count: int, d_out: int,
                            stream: Optional[int] = None) -> None:
        rc = self._lib.ufsecp_gpu_ext_ecdsa_verify_device(self._ctx, c_void_p(d_msgs), c_void_p(d_pubs),
                                                          c_void_p(d_sigs), count, c_void_p(d_out),
                                                          c_void_p(stream))
        self._check("ufsecp_gpu_ext_ecdsa_verify_device", rc)

    @staticmethod
    def batch_table(batches) -> ctypes.Array:
        """ctypes array of VerifyBatchDesc from (kind, count, d_msgs, d_pubs, d_sigs, d_out[, row_stride])
        tuples."""
        arr = (VerifyBatchDesc * len(batches))()
        for i, b in enumerate(batches):
            kind, count, m, p, s, o = b[:6]
            arr[i] = VerifyBatchDesc(kind, count, m, p, s, o, b[6] if len(b) > 6 else 0, 0)
        return arr

    def verify_multi_device(self, table, nbatches: int, use_graph: bool = True,
                            stream: Optional[int] = None) -> None:
        rc = self._lib.ufsecp_gpu_ext_verify_multi_device(self._ctx, ctypes.cast(table, c_void_p), nbatches,
                                                          1 if use_graph else 0, c_void_p(stream))
        self._check("ufsecp_gpu_ext_verify_multi_device", rc)

    def verify_each_device(self, table, nbatches: int, stream: Optional[int] = None) -> None:
        rc = self._lib.ufsecp_gpu_ext_verify_each_device(self._ctx, ctypes.cast(table, c_void_p), nbatches,
                                                         c_void_p(stream))
        self._check("ufsecp_gpu_ext_verify_each_device", rc)

    # -- device memory / events (library-owned HIP runtime) -------------------
    def to_device(self, arr) -> "DeviceArray":
        return DeviceArray.from_numpy(self, arr)

    def empty(self, nbytes: int) -> "DeviceArray":
        return DeviceArray(self, nbytes)

    def event(self) -> "Event":
        return Event(self)

    def host_empty(self, shape, dtype="uint8"):
        """numpy array in pinned host memory that the GPU reads in place
        (ufsecp_gpu_ext_host_alloc); freed with the array's last reference."""
        import numpy as np
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        p = c_void_p()
        rc = self._lib.ufsecp_gpu_ext_host_alloc(self._ctx, max(1, n), byref(p))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_host_alloc", rc)
        raw = (ctypes.c_uint8 * max(1, n)).from_address(p.value)
        arr = np.frombuffer(raw, dtype=np.uint8, count=n).view(dt).reshape(shape)
        lib, ctx, addr = self._lib, self._ctx, p.value
        import weakref
        weakref.finalize(raw, lambda: lib.ufsecp_gpu_ext_host_free(ctx, c_void_p(addr)))
        return arr

    def synchronize(self) -> None:
        self._check("ufsecp_gpu_ext_synchronize", self._lib.ufsecp_gpu_ext_synchronize(self._ctx))

    def new_stream(self) -> int:
        """A new HIP stream on the context's device (default priority); release
        it with ``destroy_stream``."""
        st = c_void_p()
        self._check("ufsecp_gpu_ext_stream_create", self.lib.ufsecp_gpu_ext_stream_create(self.ctx, ctypes.byref(st)))
        return st.value

    def destroy_stream(self, stream: int) -> None:
        self.lib.ufsecp_gpu_ext_stream_destroy(c_void_p(stream))

    def stream(self) -> int:
        return self._lib.ufsecp_gpu_ext_stream(self._ctx) or 0

    def selftest(self, op: int, a, b):
        """Run device arithmetic op on uint32 arrays a, b of shape (N, 16)."""
        import numpy as np
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        n = a.shape[0]
        out = np.zeros((n, 32), dtype=np.uint32)
        rc = self._lib.ufsecp_gpu_ext_selftest(self._ctx, op, _buf(a), _buf(b), _buf(out), n)
        self._check("ufsecp_gpu_ext_selftest", rc)
        return out


class VerifyServer:
    """Continuous-batching verification server over a context
    (ufsecp_gpu_ext_server_*).  Submitted numpy buffers are kept referenced
    until the request is observed complete."""

    def __init__(self, gpu: UfsecpGpu, hbm_budget_bytes: int = 0, inflight_groups: int = 0,
                 key_cache_log2: int = 21, max_group_items: int = 0):
        self.gpu = gpu
        self._lib = gpu.lib
        cfg = ServerConfig(hbm_budget_bytes, inflight_groups, key_cache_log2, max_group_items, 0)
        h = c_void_p()
        rc = self._lib.ufsecp_gpu_ext_server_create(gpu.ctx, byref(cfg), byref(h))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_server_create", rc)
        self._h = h
        self._keep = {}

    def submit(self, kind: int, msgs, pubkeys, sigs, out, mapped: bool = False) -> int:
        """mapped=True: the inputs live in memory from UfsecpGpu.host_empty and
        are read in place by the GPU (no staging copy)."""
        n = _nitems(msgs, 32)
        rid = ctypes.c_uint64()
        fn = self._lib.ufsecp_gpu_ext_server_submit_mapped if mapped else self._lib.ufsecp_gpu_ext_server_submit
        rc = fn(self._h, kind, _buf(msgs), _buf(pubkeys), _buf(sigs), n, _buf(out), byref(rid))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_server_submit", rc)
        self._keep[rid.value] = (msgs, pubkeys, sigs, out)
        return rid.value

    def poll(self, rid: int) -> int:
        r = self._lib.ufsecp_gpu_ext_server_poll(self._h, rid)
        if r == 1:
            self._keep.pop(rid, None)
        return r

    def wait(self, rid: int) -> None:
        rc = self._lib.ufsecp_gpu_ext_server_wait(self._h, rid)
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_server_wait", rc)
        self._keep.pop(rid, None)

    def drain(self) -> None:
        rc = self._lib.ufsecp_gpu_ext_server_drain(self._h)
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_server_drain", rc)
        self._keep.clear()

    def stats(self, reset: bool = False) -> dict:
        st = ServerStats()
        rc = self._lib.ufsecp_gpu_ext_server_stats(self._h, byref(st), 1 if reset else 0)
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_server_stats", rc)
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ufsecp_gpu_ext_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """Device buffer allocated through the library (ufsecp_gpu_ext_alloc)."""

    def __init__(self, gpu: UfsecpGpu, nbytes: int):
        self.gpu = gpu
        self.nbytes = int(nbytes)
        p = c_void_p()
        rc = gpu.lib.ufsecp_gpu_ext_alloc(gpu.ctx, max(1, self.nbytes), byref(p))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_alloc", rc)
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, gpu: UfsecpGpu, arr) -> "DeviceArray":
        import numpy as np
        arr = np.ascontiguousarray(arr)
        d = cls(gpu, arr.nbytes)
        d.upload(arr)
        return d

    def upload(self, arr) -> None:
        import numpy as np
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes:
            rc = self.gpu.lib.ufsecp_gpu_ext_copy_to_device(self.gpu.ctx, c_void_p(self.ptr), _buf(arr), arr.nbytes)
            if rc != UFSECP_OK:
                raise UfsecpError("ufsecp_gpu_ext_copy_to_device", rc)

    def download(self, dtype="uint8", count: Optional[int] = None):
        import numpy as np
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dt)
        if out.nbytes:
            rc = self.gpu.lib.ufsecp_gpu_ext_copy_to_host(self.gpu.ctx, _buf(out), c_void_p(self.ptr), out.nbytes)
            if rc != UFSECP_OK:
                raise UfsecpError("ufsecp_gpu_ext_copy_to_host", rc)
        return out

    def at(self, offset: int) -> int:
        """Device address `offset` bytes into the buffer."""
        return self.ptr + offset

    def free(self) -> None:
        if self.ptr:
            self.gpu.lib.ufsecp_gpu_ext_free(self.gpu.ctx, c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):  # pragma: no cover
        try:
            if self.ptr and self.gpu.ctx:
                self.free()
        except Exception:
            pass


class Event:
    """hipEvent recorded on a stream (default: the context stream)."""

    def __init__(self, gpu: UfsecpGpu):
        self.gpu = gpu
        p = c_void_p()
        rc = gpu.lib.ufsecp_gpu_ext_event_create(gpu.ctx, byref(p))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_event_create", rc)
        self.ptr = p.value

    def record(self, stream: Optional[int] = None) -> None:
        rc = self.gpu.lib.ufsecp_gpu_ext_event_record(self.gpu.ctx, c_void_p(self.ptr), c_void_p(stream))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_event_record", rc)

    def elapsed_ms(self, stop: "Event") -> float:
        ms = ctypes.c_float()
        rc = self.gpu.lib.ufsecp_gpu_ext_event_elapsed(c_void_p(self.ptr), c_void_p(stop.ptr), byref(ms))
        if rc != UFSECP_OK:
            raise UfsecpError("ufsecp_gpu_ext_event_elapsed", rc)
        return ms.value

    def __del__(self):  # pragma: no cover
        try:
            self.gpu.lib.ufsecp_gpu_ext_event_destroy(c_void_p(self.ptr))
        except Exception:
            pass
