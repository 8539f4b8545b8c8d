"""The candidate bound of the move evaluation.

A candidate is dropped before its module statistics are read only when a
lower bound of its dL, from the unit's flow to it alone, lies above dL = 0.
A dropped candidate can therefore be neither the best move nor a move at
all, and the bounded run must reproduce the unbounded run bit for bit:
the same partition and codelength after the whole schedule (sweeps at every
level, contractions, coarse tunes), in both precisions, with the units in
every kernel (group, block, hub partitions).  The bound must also drop rows:
the profiling mode's algorithmic bytes of the evaluation kernels (32 B per
module row read) fall.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graph():
    from relaxmap_amd import synth
    # heavy-tailed degrees: units in every bucket, hubs included
    return synth.chunglu(40_000, 2.1, 4, 6000, 16.0, 1.2, 50, 4000, 0.4, 0.01, 11)


def _run(g, precision, mode=None, bucket=-1, cuts=None, profile=False):
    from relaxmap_amd import Network
    with Network(precision=precision) as net:
        g.upload_to(net)
        if profile:
            net.set_profiling(True)
        net.initiate()
        if cuts is not None:
            net.set_bucket_cutoffs(cuts)
        if mode is not None:
            net.set_candidate_bound(mode, bucket)
        r = net.run(seed=3)
        stats = net.kernel_stats() if profile else None
        return net.partition(), r.computed_codelength, stats


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bound_modes_give_identical_runs(precision):
    g = _graph()
    part0, L0, _ = _run(g, precision, mode=0)
    for mode, bucket in ((1, -1), (None, -1), (1, 7), (1, 3)):
        part, L, _ = _run(g, precision, mode=mode, bucket=bucket)
        print(f"[gpu] {precision} bound {mode} bucket {bucket}: L {L:.12f} vs {L0:.12f}", flush=True)
        assert L == L0
        assert np.array_equal(part, part0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bound_in_every_kernel(precision):
    """Every unit through the hub partitions, then every unit of 17-2048 arcs
    through the workgroup tables: the bound gives the unbounded run's result."""
    g = _graph()
    for cuts in ([0] * 7, [16, 16, 16, 16, 16, 16, 2048]):
        part0, L0, _ = _run(g, precision, mode=0, cuts=cuts)
        part, L, _ = _run(g, precision, mode=1, cuts=cuts)
        print(f"[gpu] {precision} cuts {cuts}: L {L:.12f} vs {L0:.12f}", flush=True)
        assert L == L0
        assert np.array_equal(part, part0)


def test_bound_drops_module_rows():
    g = _graph()
    _, L0, s0 = _run(g, "mixed", mode=0, profile=True)
    _, L2, s2 = _run(g, "mixed", profile=True)
    assert L0 == L2
    evals = [k for k in s0 if k.startswith("k_eval_") or k.startswith("k_hubp_eval")]
    assert evals
    b0 = sum(s0[k][2] for k in evals)
    b2 = sum(s2[k][2] for k in evals if k in s2)
    print(f"[gpu] evaluation kernels' algorithmic bytes: {b0 / 1e6:.1f} MB unbounded, {b2 / 1e6:.1f} MB bounded",
          flush=True)
    assert b2 < 0.9 * b0


def test_bound_on_a_batch_of_instances():
    from relaxmap_amd import Network, synth
    from relaxmap_amd.batch import graphs_union
    graphs = [synth.sbm(n, b, 12.0, mu, 40 + k)
              for k, (n, b, mu) in enumerate([(600, 12, 0.2), (3000, 40, 0.3), (5000, 100, 0.25)])]
    goff, ptr, idx, w = graphs_union(graphs)
    seeds = np.array([1, 2, 1, 2, 1, 2], dtype=np.uint64)
    out = []
    for mode in (0, 1):
        with Network(precision="mixed") as net:
            net.batch_upload_seeds(goff, 2, seeds, ptr, idx, w)
            net.initiate()
            net.set_candidate_bound(mode)
            net.run(seed=1)
            L, nmod, _ = net.batch_results()
            out.append((np.asarray(L).copy(), np.asarray(nmod).copy(), net.partition()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_bound_setter_rejects():
    from relaxmap_amd import Network
    with Network() as net:
        with pytest.raises(RuntimeError):
            net.set_candidate_bound(2)
        with pytest.raises(RuntimeError):
            net.set_candidate_bound(1, bucket=8)
