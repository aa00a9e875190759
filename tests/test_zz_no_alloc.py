"""ufsecp_gpu_ext_msm_device_noalloc: the device-pointer MSM without any
allocation (buffers carved from the library's static per-device scratch).

Two calls in a row through the C ABI, the second with fewer terms, each equal
to ufsecp_gpu_ext_msm_device on the same inputs and to the closed form
(sum s_i k_i) G, with the library's buffer-allocation count unchanged across
them; requests beyond the scratch's limits are BAD_INPUT with nothing
written."""
import ctypes

import numpy as np
import pytest

from ec_math import compress, mul
from ultrafastsecp256k1_amd import UFSECP_ERR_BAD_INPUT, UFSECP_OK
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu


def _noalloc(gpu, d_s, d_p, n, d_o, c=0):
    rc = gpu.lib.ufsecp_gpu_ext_msm_device_noalloc(gpu.ctx, ctypes.c_void_p(d_s.ptr), ctypes.c_void_p(d_p.ptr), n,
                                                   ctypes.c_void_p(d_o.ptr), ctypes.c_void_p(d_o.at(48)), c, None)
    gpu.synchronize()
    raw = d_o.download()
    return rc, int(np.frombuffer(raw[48:52].tobytes(), np.int32)[0]), raw[:33].tobytes()


def _regular(gpu, d_s, d_p, n, c=0):
    d_o = gpu.empty(64)
    gpu.msm_device(d_s.ptr, d_p.ptr, n, d_o.ptr, d_o.at(48), c)
    gpu.synchronize()
    raw = d_o.download()
    return int(np.frombuffer(raw[48:52].tobytes(), np.int32)[0]), raw[:33].tobytes()


def test_msm_noalloc_matches_msm_device_two_calls(gpu):
    jobs = []
    for n, seed, c in ((1 << 14, 0xA110C, 0), (3001, 0xA110D, 0), (4096, 0xA110E, 9)):
        sc, pts, ks = wl.generate_msm(n, seed=seed, threads=8)
        jobs.append((n, c, gpu.to_device(sc), gpu.to_device(pts), compress(mul(wl.msm_closed_form_scalar(sc, ks)))))
    d_o = gpu.empty(64)
    # the test's own device buffers are allocated before the count is taken
    bad = gpu.to_device(np.frombuffer(b"\x05" + bytes(32), np.uint8).reshape(1, 33).repeat(3001, 0))
    for n, c, d_s, d_p, want in jobs:
        assert _regular(gpu, d_s, d_p, n, c) == (0, want)      # warms the regular path's own buffers
    before = gpu.lib.ufsecp_gpu_ext_buffer_alloc_count()
    for n, c, d_s, d_p, want in jobs:                         # 16384 terms, then fewer: shared scratch reused
        d_o.upload(np.full(64, 0xCC, np.uint8))
        assert _noalloc(gpu, d_s, d_p, n, d_o, c) == (UFSECP_OK, 0, want), n
    assert gpu.lib.ufsecp_gpu_ext_buffer_alloc_count() == before
    # an invalid term in a call after valid ones: the error word starts over
    n, c, d_s, d_p, want = jobs[1]
    d_o.upload(np.full(64, 0xCC, np.uint8))
    assert _noalloc(gpu, d_s, bad, n, d_o)[:2] == (UFSECP_OK, 3)
    d_o.upload(np.full(64, 0xCC, np.uint8))
    assert _noalloc(gpu, d_s, d_p, n, d_o) == (UFSECP_OK, 0, want)
    assert gpu.lib.ufsecp_gpu_ext_buffer_alloc_count() == before


def test_msm_noalloc_limits(gpu):
    """More than 16384 terms, or a window whose buckets do not fit the
    scratch, is BAD_INPUT before anything is enqueued; n = 0 is OK."""
    n = (1 << 14) + 1
    sc, pts, _ = wl.generate_msm(n, seed=5, threads=8)
    d_s, d_p, d_o = gpu.to_device(sc), gpu.to_device(pts), gpu.empty(64)
    d_o.upload(np.full(64, 0xCC, np.uint8))
    assert _noalloc(gpu, d_s, d_p, n, d_o)[0] == UFSECP_ERR_BAD_INPUT
    assert _noalloc(gpu, d_s, d_p, 1 << 14, d_o, 18)[0] == UFSECP_ERR_BAD_INPUT
    assert bytes(d_o.download()) == b"\xcc" * 64
    assert _noalloc(gpu, d_s, d_p, 0, d_o)[0] == UFSECP_OK
