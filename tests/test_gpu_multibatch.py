"""Persistent multi-batch verification (one launch over many uneven batches,
optionally replayed from a captured hipGraph) gives exactly the per-batch
verdicts, which in turn equal the oracle's."""
import numpy as np
import pytest

from oracle import oracle as orc
from ultrafastsecp256k1_amd import workload as wl

pytestmark = pytest.mark.gpu


def _device_mix(gpu, nbatches, max_size, seed):
    plan = wl.mix_plan(nbatches, max_size, seed=seed)
    (pe, ps), layout = wl.generate_mix(plan, seed=seed)
    pools = [tuple(gpu.to_device(a) for a in (p.msgs, p.pubkeys, p.sigs)) for p in (pe, ps)]
    outs = [gpu.empty(max(1, p.n)) for p in (pe, ps)]
    pstride = (33, 32)
    batches = []
    for kind, first, size in layout:
        m, pk, sg = pools[kind]
        batches.append((kind, size, m.at(32 * first), pk.at(pstride[kind] * first), sg.at(64 * first),
                        outs[kind].at(first)))
    return plan, (pe, ps), pools, outs, batches


def _fill(gpu, outs, value):
    for o in outs:
        o.upload(np.full(o.nbytes, value, np.uint8))


def test_multibatch_matches_per_batch_and_oracle(gpu):
    plan, (pe, ps), pools, outs, batches = _device_mix(gpu, 120, 300, seed=41)
    table = gpu.batch_table(batches)
    exp = [orc.verify_batch("ecdsa", pe.msgs, pe.pubkeys, pe.sigs),
           orc.verify_batch("schnorr", ps.msgs, ps.pubkeys, ps.sigs)]
    # plain per-batch launches
    gpu.verify_each_device(table, len(batches))
    gpu.synchronize()
    per_batch = [o.download() for o in outs]
    for k in (0, 1):
        assert np.array_equal(per_batch[k][:len(exp[k])], exp[k])
    # persistent launch, direct and graph replays (with eager work in between)
    for use_graph in (False, True, True, True):
        _fill(gpu, outs, 7)
        gpu.verify_multi_device(table, len(batches), use_graph=use_graph)
        gpu.synchronize()
        got = [o.download() for o in outs]
        for k in (0, 1):
            assert np.array_equal(got[k][:len(exp[k])], exp[k]), (use_graph, k)
        b = wl.generate("schnorr", 70, first=9000)
        assert gpu.schnorr_verify_batch(b.msgs, b.pubkeys, b.sigs).sum() == 70 - b.corrupted.sum()


def test_multibatch_graph_replay_with_new_sizes(gpu):
    """A replayed graph picks up a different batch list (tile count in device memory)."""
    for seed, nb in ((43, 40), (44, 90), (45, 7)):
        plan, (pe, ps), pools, outs, batches = _device_mix(gpu, nb, 200, seed=seed)
        gpu.verify_multi_device(gpu.batch_table(batches), len(batches), use_graph=True)
        gpu.synchronize()
        got = [o.download() for o in outs]
        assert np.array_equal(got[0][:pe.n], orc.verify_batch("ecdsa", pe.msgs, pe.pubkeys, pe.sigs))
        assert np.array_equal(got[1][:ps.n], orc.verify_batch("schnorr", ps.msgs, ps.pubkeys, ps.sigs))
