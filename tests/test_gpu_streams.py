"""Calls on different caller streams share one context's scratch (verify
descriptor / tile tables and lane slots, MSM buffers): each call is ordered
after the previous one, so back-to-back device-pointer calls on two streams,
with no host synchronisation between them, give the results of the same calls
made one at a time."""
import numpy as np
import pytest

from ec_math import compress, mul
from oracle import oracle as orc
from test_gpu_multibatch import _device_mix, _fill
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu


def test_multibatch_on_two_streams(gpu):
    s1, s2 = gpu.new_stream(), gpu.new_stream()
    try:
        mixes = [_device_mix(gpu, nb, 400, seed=seed) for seed, nb in ((61, 150), (62, 60))]
        exp = [[orc.verify_batch("ecdsa", pe.msgs, pe.pubkeys, pe.sigs),
                orc.verify_batch("schnorr", ps.msgs, ps.pubkeys, ps.sigs)] for _, (pe, ps), _, _, _ in mixes]
        for use_graph in (False, True):
            for _, _, _, outs, _ in mixes:
                _fill(gpu, outs, 7)
            # large list on stream 1, then immediately a different list on stream 2
            for (_, _, _, _, batches), st in zip(mixes, (s1, s2)):
                gpu.verify_multi_device(gpu.batch_table(batches), len(batches), use_graph=use_graph, stream=st)
            for st in (s1, s2):
                gpu.lib.ufsecp_gpu_ext_stream_synchronize(st)
            for (_, _, _, outs, _), e in zip(mixes, exp):
                for k in (0, 1):
                    assert np.array_equal(outs[k].download()[:len(e[k])], e[k]), (use_graph, k)
    finally:
        gpu.destroy_stream(s1)
        gpu.destroy_stream(s2)


def test_msm_device_on_two_streams(gpu):
    """Two MSMs of different sizes (the second grows the engine's buffers)
    enqueued on two streams back to back, then a host-buffer MSM on the
    context stream: each equals its closed form."""
    s1, s2 = gpu.new_stream(), gpu.new_stream()
    try:
        jobs = []
        for n, seed in ((1 << 14, 71), (3 << 14, 72)):
            sc, pts, ks = wl.generate_msm(n, seed=seed)
            jobs.append((n, gpu.to_device(sc), gpu.to_device(pts), gpu.empty(64),
                         compress(mul(wl.msm_closed_form_scalar(sc, ks)))))
        for (n, dsc, dpt, dout, _), st in zip(jobs, (s1, s2)):
            gpu.msm_device(dsc.at(0), dpt.at(0), n, dout.at(0), dout.at(48), stream=st)
        sc, pts, ks = wl.generate_msm(5000, seed=73)
        assert gpu.msm(sc, pts) == compress(mul(wl.msm_closed_form_scalar(sc, ks)))
        for st in (s1, s2):
            gpu.lib.ufsecp_gpu_ext_stream_synchronize(st)
        for n, _, _, dout, want in jobs:
            raw = dout.download()
            assert int.from_bytes(bytes(raw[48:52]), "little") == 0, n
            assert bytes(raw[:33]) == want, n
    finally:
        gpu.destroy_stream(s1)
        gpu.destroy_stream(s2)
