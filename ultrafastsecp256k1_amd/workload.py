"""Synthetic verify / MSM workloads (public test data only).

Input specification shared by the C++ generator (``csrc/workload.cpp``), the
fixture script and the tests.  Every item has its own SplitMix64 stream so
items can be generated in parallel:

    state_i = seed * 0x9E3779B97F4A7C15 + (i + 1) * 0xD1B54A32D192ED03  (mod 2^64)
    next():  state += 0x9E3779B97F4A7C15; z = state;
             z = (z ^ z>>30) * 0xBF58476D1CE4E5B9; z = (z ^ z>>27) * 0x94D049BB133111EB;
             return z ^ z>>31
    next256(): four next() outputs, each big-endian, concatenated (32 bytes)
    key():     next256() repeated until 1 <= value < n

ECDSA item i:    d = key(); msg = next256(); c = next()
Schnorr item i:  d = key(); msg = next256(); aux = next256(); c = next()
  Signatures: RFC6979 ECDSA with low-S (reference ufsecp_ecdsa_sign) and
  BIP-340 with aux_rand = aux (reference ufsecp_schnorr_sign).
  Corruption (1% of items): c % 100 == 0; field = (c >> 8) % 3 selects r,
  s or the digest; bit = (c >> 16) % 256 flips byte bit>>3, mask 1 << (bit&7).
MSM term i:      k = key(); s = next256() reduced once mod n; base = k*G.

The closed form of the MSM workload is (sum s_i k_i mod n) * G.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional

import numpy as np

N_ORDER = 0xFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141
P_FIELD = 2 ** 256 - 2 ** 32 - 977
MASK64 = (1 << 64) - 1

SEED_ECDSA = 0xEC05A
SEED_SCHNORR = 0x5C4A0
SEED_MSM = 0x3535
SEED_MIX = 0x30C3

GAMMA = 0x9E3779B97F4A7C15
ITEM_MUL = 0xD1B54A32D192ED03


class SplitMix64:
    def __init__(self, seed: int, item: int):
        self.state = (seed * GAMMA + (item + 1) * ITEM_MUL) & MASK64

    def next(self) -> int:
        self.state = (self.state + GAMMA) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def next256(self) -> bytes:
        return b"".join(self.next().to_bytes(8, "big") for _ in range(4))

    def key(self) -> bytes:
        while True:
            v = self.next256()
            if 1 <= int.from_bytes(v, "big") < N_ORDER:
                return v


def _corruption(c: int) -> dict:
    return {"corrupt": int(c % 100 == 0), "field": (c >> 8) % 3, "bit": (c >> 16) % 256}


def ecdsa_item_spec(seed: int, i: int) -> dict:
    r = SplitMix64(seed, i)
    d = r.key()
    msg = r.next256()
    return {"d": d, "msg": msg, **_corruption(r.next())}


def schnorr_item_spec(seed: int, i: int) -> dict:
    r = SplitMix64(seed, i)
    d = r.key()
    msg = r.next256()
    aux = r.next256()
    return {"d": d, "msg": msg, "aux": aux, **_corruption(r.next())}


def msm_term_spec(seed: int, i: int) -> dict:
    r = SplitMix64(seed, i)
    k = r.key()
    s = int.from_bytes(r.next256(), "big")
    if s >= N_ORDER:
        s -= N_ORDER
    return {"k": k, "s": s.to_bytes(32, "big")}


def apply_corruption(spec: dict, msg: bytes, sig: bytes):
    if not spec["corrupt"]:
        return msg, sig
    bit = spec["bit"]
    if spec["field"] == 2:
        m = bytearray(msg)
        m[bit >> 3] ^= 1 << (bit & 7)
        return bytes(m), sig
    s = bytearray(sig)
    off = 0 if spec["field"] == 0 else 32
    s[off + (bit >> 3)] ^= 1 << (bit & 7)
    return msg, bytes(s)


# ---------------------------------------------------------------------------
# Fast generation through libufsecp_workload.so (multi-threaded C++)
# ---------------------------------------------------------------------------
_WL: Optional[ctypes.CDLL] = None


def _lib() -> ctypes.CDLL:
    global _WL
    if _WL is None:
        p = Path(__file__).resolve().parent / "libufsecp_workload.so"
        if not p.is_file():
            raise FileNotFoundError(f"{p} not built: python -m ultrafastsecp256k1_amd.build workload")
        lib = ctypes.CDLL(str(p))
        vp = ctypes.c_void_p
        for name in ("ufsecp_wl_ecdsa", "ufsecp_wl_schnorr"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp, vp, ctypes.c_int]
        for name in ("ufsecp_wl_ecdsa_pool", "ufsecp_wl_schnorr_pool"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                           ctypes.c_uint32, vp, vp, vp, vp, ctypes.c_int]
        lib.ufsecp_wl_msm.restype = ctypes.c_int
        lib.ufsecp_wl_msm.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, vp, vp, vp, vp,
                                      ctypes.c_int]
        _WL = lib
    return _WL


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class VerifyBatch:
    """Arrays for one verify batch: msgs (N,32), pubkeys (N,33 or N,32), sigs (N,64)."""

    def __init__(self, kind: str, msgs, pubkeys, sigs, corrupted):
        self.kind = kind
        self.msgs = msgs
        self.pubkeys = pubkeys
        self.sigs = sigs
        self.corrupted = corrupted

    @property
    def n(self) -> int:
        return self.msgs.shape[0]


def generate(kind: str, n: int, seed: Optional[int] = None, first: int = 0, threads: int = 0) -> VerifyBatch:
    """Items first..first+n-1 of the ECDSA or Schnorr workload with `seed`."""
    lib = _lib()
    if seed is None:
        seed = SEED_ECDSA if kind == "ecdsa" else SEED_SCHNORR
    msgs = np.zeros((n, 32), np.uint8)
    pubs = np.zeros((n, 33 if kind == "ecdsa" else 32), np.uint8)
    sigs = np.zeros((n, 64), np.uint8)
    corr = np.zeros(n, np.uint8)
    fn = lib.ufsecp_wl_ecdsa if kind == "ecdsa" else lib.ufsecp_wl_schnorr
    rc = fn(seed, first, n, _ptr(msgs), _ptr(pubs), _ptr(sigs), _ptr(corr), threads)
    if rc != 0:
        raise RuntimeError(f"workload generation failed ({rc})")
    return VerifyBatch(kind, msgs, pubs, sigs, corr)


def generate_pool(kind: str, n: int, seed: int, pool_seed: int, pool_size: int = 1 << 16,
                  pool_permille: int = 200, first: int = 0, threads: int = 0) -> VerifyBatch:
    """Serving workload items: as generate(), except that each item's key is,
    with probability pool_permille/1000, one of pool_size keys of a seeded pool
    (first draw u of the item stream: u % 1000 < pool_permille selects pool key
    (u >> 20) % pool_size, pool key j being stream (pool_seed, j)'s first key()
    draw); otherwise the item stream's next key() draw."""
    lib = _lib()
    msgs = np.zeros((n, 32), np.uint8)
    pubs = np.zeros((n, 33 if kind == "ecdsa" else 32), np.uint8)
    sigs = np.zeros((n, 64), np.uint8)
    corr = np.zeros(n, np.uint8)
    fn = lib.ufsecp_wl_ecdsa_pool if kind == "ecdsa" else lib.ufsecp_wl_schnorr_pool
    rc = fn(seed, first, n, pool_seed, pool_size, pool_permille, _ptr(msgs), _ptr(pubs), _ptr(sigs), _ptr(corr),
            threads)
    if rc != 0:
        raise RuntimeError(f"workload generation failed ({rc})")
    return VerifyBatch(kind, msgs, pubs, sigs, corr)


def generate_msm(n: int, seed: int = SEED_MSM, first: int = 0, threads: int = 0, affine: bool = False):
    """(scalars (N,32), points33 (N,33), ks (N,32): base i is k_i G).  With
    affine=True a fourth array holds the same bases as (N,64) big-endian
    affine x || y."""
    lib = _lib()
    sc = np.zeros((n, 32), np.uint8)
    pts = np.zeros((n, 33), np.uint8)
    ks = np.zeros((n, 32), np.uint8)
    p64 = np.zeros((n, 64), np.uint8) if affine else None
    rc = lib.ufsecp_wl_msm(seed, first, n, _ptr(sc), _ptr(pts), _ptr(p64) if affine else None, _ptr(ks), threads)
    if rc != 0:
        raise RuntimeError(f"MSM workload generation failed ({rc})")
    return (sc, pts, ks, p64) if affine else (sc, pts, ks)


def mix_plan(nbatches: int, max_size: int, seed: int = SEED_MIX, ecdsa_share: float = 0.7):
    """Uneven batch mix: sizes log-uniform in [1, max_size], kind ECDSA with
    probability ecdsa_share (else BIP-340).  Returns [(kind, size)] with kind
    0 = ECDSA, 1 = BIP-340."""
    import math
    r = SplitMix64(seed, 0xFFFFFFFF)
    plan = []
    for _ in range(nbatches):
        u = r.next() / 2.0 ** 64
        size = int(math.exp(u * math.log(max_size + 1)))
        size = min(max(size, 1), max_size)
        kind = 0 if (r.next() % 1000) < int(ecdsa_share * 1000) else 1
        plan.append((kind, size))
    return plan


def generate_mix(plan, seed: int = SEED_MIX, threads: int = 0):
    """Inputs for a mix plan: one ECDSA and one BIP-340 pool (items of the
    ECDSA / BIP-340 workloads with seeds seed+1 / seed+2) and, per batch, its
    (kind, first item in the pool, size)."""
    ne = sum(s for k, s in plan if k == 0)
    ns = sum(s for k, s in plan if k == 1)
    pools = (generate("ecdsa", ne, seed + 1, threads=threads), generate("schnorr", ns, seed + 2, threads=threads))
    offs = [0, 0]
    layout = []
    for k, s in plan:
        layout.append((k, offs[k], s))
        offs[k] += s
    return pools, layout


def msm_closed_form_scalar(scalars: np.ndarray, ks: np.ndarray) -> int:
    """sum s_i * k_i mod n (the MSM result is this scalar times G)."""
    acc = 0
    s_int = [int.from_bytes(bytes(r), "big") for r in scalars]
    k_int = [int.from_bytes(bytes(r), "big") for r in ks]
    for a, b in zip(s_int, k_int):
        acc += a * b
    return acc % N_ORDER
