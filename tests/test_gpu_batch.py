"""Many networks x seeds (relaxmap_amd.batch) and the CLI's attempts.

Concurrent contexts must give exactly the results of one context, and a run
with several attempts must keep the best of the same runs made one by one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs():
    from relaxmap_amd import synth
    return [synth.sbm(n, b, 12.0, mu, 40 + k)
            for k, (n, b, mu) in enumerate([(600, 12, 0.2), (1500, 30, 0.35), (3000, 40, 0.3),
                                            (800, 10, 0.45), (2500, 60, 0.15), (5000, 100, 0.25)])]


def test_instances_independent_of_workers():
    from relaxmap_amd.batch import run_instances
    graphs = _graphs()
    seeds = [1, 2, 3]
    one, t1 = run_instances(graphs, seeds, workers=1)
    many, t3 = run_instances(graphs, seeds, workers=3)
    print(f"[gpu] {len(graphs) * len(seeds)} instances: 1 worker {t1:.2f} s, 3 workers {t3:.2f} s", flush=True)
    assert set(one) == set(many) == {(g, s) for g in range(len(graphs)) for s in seeds}
    for k in one:
        assert one[k] == many[k], k


def test_instances_match_single_runs():
    """A seed run after set_partition(identity) equals a run after initiate."""
    from relaxmap_amd import Network
    from relaxmap_amd.batch import run_instances
    g = _graphs()[2]
    res, _ = run_instances([g], [5, 9], workers=1)
    for s in (5, 9):
        with Network(precision="mixed") as net:
            g.upload_to(net)
            net.initiate()
            r = net.run(seed=s)
            assert res[(0, s)][2] == r.computed_codelength


def test_attempts_keep_the_best():
    """Attempts one after another: the best of the same runs made one by one."""
    from relaxmap_amd import Network
    g = _graphs()[3]
    single = []
    for s in (7, 8, 9):
        with Network() as net:
            g.upload_to(net)
            net.initiate()
            single.append(net.run(seed=s).computed_codelength)
    with Network() as net:
        g.upload_to(net)
        net.set_attempt_mode(True)
        net.initiate()
        r = net.run(seed=7, attempts=3)
        best = int(np.argmin(single))
        assert r.best_attempt == best
        assert r.computed_codelength == single[best]
        assert abs(net.codelength - r.computed_codelength) <= 1e-9 * abs(r.computed_codelength)


def test_attempts_as_one_batch():
    """Attempts as one batch of copies of the network (the default where they
    fit): the kept partition is the best copy's of the same batch (copies with
    seeds seed, seed + 1, ..., partitioned with seed), and the context's
    codelength is that copy's."""
    from relaxmap_amd import Network
    from relaxmap_amd.batch import union_of
    g = _graphs()[3]
    seeds = [7, 8, 9, 10]
    _, node_off, iseeds, ptr, idx, w = union_of([g], seeds)
    with Network() as bnet:
        bnet.batch_upload(node_off, iseeds, ptr, idx, w)
        bnet.initiate()
        bnet.run(seed=7)
        Ls = bnet.batch_results()[0]
        bpart = bnet.partition()
    best = int(np.argmin(Ls))
    with Network() as net:
        g.upload_to(net)
        net.initiate()
        r = net.run(seed=7, attempts=len(seeds))
        part = net.partition()
        assert abs(net.codelength - r.computed_codelength) <= 1e-9 * abs(r.computed_codelength)
    print(f"[gpu] batched attempts: {[round(float(x), 6) for x in Ls]}, kept {r.best_attempt}", flush=True)
    assert r.best_attempt == best
    assert abs(r.computed_codelength - Ls[best]) <= 1e-9 * abs(Ls[best])
    assert np.array_equal(part, bpart[best * g.n:(best + 1) * g.n] - best * g.n)


def test_seeded_copies_equal_the_host_union():
    """rmx_batch_upload_seeds (each network sent once and copied on the
    device, one copy per seed) against rmx_batch_upload of the same union
    built on the host: identical flows, per-instance codelengths, module
    counts, iterations and partitions, in both precisions.  One network has
    no arcs (copies of an arc-less node range)."""
    from types import SimpleNamespace
    from relaxmap_amd import Network
    from relaxmap_amd.batch import graphs_union, union_of
    g = _graphs()
    empty = SimpleNamespace(n=5, m=0, ptr=np.zeros(6, np.int64), idx=np.zeros(0, np.int32), w=np.zeros(0))
    graphs = [g[0], empty, g[2], g[1]]
    seeds = [3, 1, 4]
    _, node_off, iseeds, ptr, idx, w = union_of(graphs, seeds)
    goff, gptr, gidx, gw = graphs_union(graphs)
    for prec in ("mixed", "fp64"):
        out = []
        for seeded in (False, True):
            with Network(precision=prec) as net:
                if seeded:
                    net.batch_upload_seeds(goff, len(seeds), iseeds, gptr, gidx, gw)
                else:
                    net.batch_upload(node_off, iseeds, ptr, idx, w)
                net.initiate()
                p = net.flows()
                net.run(seed=1)
                out.append((p, net.batch_results(), net.partition()))
        (p0, r0, q0), (p1, r1, q1) = out
        print(f"[gpu] {prec}: {len(r0[0])} instances, L[0] {r0[0][0]:.9f} / {r1[0][0]:.9f}", flush=True)
        assert np.array_equal(p0, p1)
        for a, b in zip(r0, r1):
            assert np.array_equal(a, b)
        assert np.array_equal(q0, q1)


# ---------------------------------------------------------------- one batch
def _oracle_codelength(g, part):
    from oracle.pyoracle import OracleNet
    src = np.repeat(np.arange(g.n), np.diff(np.asarray(g.ptr)))
    o = OracleNet.from_arcs(g.n, src, np.asarray(g.idx), np.asarray(g.w)).initiate()
    o.set_partition(np.asarray(part, dtype=np.int32))
    o.update_members()
    return o.calculate_codelength()


def test_batch_partition_valid_and_exact():
    """Every instance's modules stay inside it, and its codelength equals the
    oracle's codelength of its own partition (FP64 flows)."""
    from relaxmap_amd.batch import run_batched, union_of
    graphs = _graphs()[:4]
    seeds = [1, 2]
    res, dt, net = run_batched(graphs, seeds, precision="fp64", return_net=True)
    with net:
        part = net.partition()
    pairs, node_off, _, _, _, _ = union_of(graphs, seeds)
    print(f"[gpu] batch of {len(pairs)} instances in {dt:.2f} s", flush=True)
    for k, (gi, s) in enumerate(pairs):
        a, b = int(node_off[k]), int(node_off[k + 1])
        p = part[a:b]
        assert p.min() >= a and p.max() < b, (gi, s)
        ref = _oracle_codelength(graphs[gi], p - a)
        assert abs(res[(gi, s)][2] - ref) <= 1e-9 * abs(ref), (gi, s)
        assert res[(gi, s)][1] == len(np.unique(p))


def test_batch_flows_match_single_networks():
    """The steady state of every instance is that of its network alone."""
    from relaxmap_amd import Network
    from relaxmap_amd.batch import union_of
    graphs = _graphs()[:3]
    pairs, node_off, iseeds, ptr, idx, w = union_of(graphs, [1])
    # FP64: the reference's rule per instance (flows 1e-12 L1); mixed: its
    # 1e-8 residual stop (flows 1e-6 L1 of FP64)
    for prec, tol, tol_l in (("fp64", 1e-12, 1e-9), ("mixed", 1e-6, 1e-6)):
        with Network(precision=prec) as net:
            net.batch_upload(node_off, iseeds, ptr, idx, w)
            net.initiate()
            p = net.flows()
            L0 = net.batch_results()[0]
        for k, (gi, _) in enumerate(pairs):
            a, b = int(node_off[k]), int(node_off[k + 1])
            with Network(precision="fp64") as solo:
                graphs[gi].upload_to(solo)
                _, l0 = solo.initiate()
                q = solo.flows()
            assert abs(p[a:b].sum() - 1.0) <= 1e-12
            assert np.abs(p[a:b] - q).sum() <= tol, (prec, gi)
            assert abs(L0[k] - l0) <= tol_l * abs(l0), (prec, gi)   # singleton modules


def test_batch_close_to_single_runs():
    """Same networks and seeds, one batch vs one context each: the batch
    draws its own random sweeps, so codelengths agree to within 1 %."""
    from relaxmap_amd.batch import run_batched, run_instances
    graphs = _graphs()
    seeds = [1, 2]
    rb, tb = run_batched(graphs, seeds)
    ri, ti = run_instances(graphs, seeds, workers=4)
    print(f"[gpu] {len(rb)} instances: batch {tb:.2f} s, contexts {ti:.2f} s", flush=True)
    assert set(rb) == set(ri)
    rel = [abs(rb[k][0] - ri[k][0]) / ri[k][0] for k in rb]
    print(f"[gpu] codelength difference: median {np.median(rel):.2e}, max {max(rel):.2e}", flush=True)
    assert max(rel) <= 0.01


def test_instance_seeds_change_the_draws():
    """Copies of one network in one batch, each with its own seed: after one
    sweep their partitions differ (every instance draws its own sub-sweeps,
    the seed mixed into each unit's draw), and over whole runs their
    codelengths vary as much as runs of the network alone with those seeds
    do: the median absolute deviation over the seeds (max - min is ruled by
    single outlier runs, which either mode can draw)."""
    from relaxmap_amd import Network
    from relaxmap_amd.batch import run_batched, run_instances, union_of
    from relaxmap_amd import synth
    g = synth.lfr(20_000, 2.0, 5, 200, 1.0, 20, 500, 0.45, 0.0, 3)   # weakly modular: seeds matter
    seeds = list(range(1, 9))
    pairs, node_off, iseeds, ptr, idx, w = union_of([g], seeds)
    with Network(batches=4) as net:
        net.batch_upload(node_off, iseeds, ptr, idx, w)
        net.initiate()
        net.move(prioritized=False, seed=1)
        part = net.partition()
    local = [part[int(node_off[k]):int(node_off[k + 1])] - int(node_off[k]) for k in range(len(pairs))]
    same = sum(np.array_equal(local[0], local[k]) for k in range(1, len(local)))
    print(f"[gpu] {len(pairs)} copies after one sweep: {same} equal to the first", flush=True)
    assert same == 0
    rb, _ = run_batched([g], seeds)
    ri, _ = run_instances([g], seeds, workers=1)
    Lb = np.array([rb[(0, s)][0] for s in seeds])
    Li = np.array([ri[(0, s)][0] for s in seeds])
    mad = lambda x: float(np.median(np.abs(x - np.median(x))) / np.median(x))
    sb, si = mad(Lb), mad(Li)
    print(f"[gpu] codelength over {len(seeds)} seeds: batch {np.round(Lb, 6).tolist()}, alone "
          f"{np.round(Li, 6).tolist()}; median absolute deviation {sb:.2e} vs {si:.2e}", flush=True)
    assert len(np.unique(Lb)) >= 3
    assert 0.2 * si <= sb <= 5.0 * si


def test_batch_deterministic():
    from relaxmap_amd.batch import run_batched
    graphs = _graphs()[3:5]
    r1, _ = run_batched(graphs, [3, 4])
    r2, _ = run_batched(graphs, [3, 4])
    assert r1 == r2


class _Net:
    """A network given by its out-CSR (the attributes union_of reads)."""

    def __init__(self, n, ptr, idx, w):
        self.n, self.ptr, self.idx, self.w = n, np.asarray(ptr), np.asarray(idx), np.asarray(w, dtype=np.float64)
        self.m = len(self.idx)


def test_batch_degenerate_instances():
    """One node, one arc, a 3-cycle and a dangling pair beside an SBM: every
    instance is partitioned on its own and matches the oracle."""
    from relaxmap_amd import synth
    from relaxmap_amd.batch import run_batched, union_of
    graphs = [_Net(1, [0, 0], [], []), _Net(2, [0, 1, 1], [1], [1.0]), _Net(3, [0, 1, 2, 3], [1, 2, 0], [1, 1, 1]),
              _Net(4, [0, 1, 1, 2, 2], [1, 3], [2.0, 0.5]), synth.sbm(300, 6, 8.0, 0.2, 5)]
    for prec in ("fp64", "mixed"):
        res, _, net = run_batched(graphs, [1, 2], precision=prec, return_net=True)
        with net:
            part = net.partition()
            p = net.flows()
        pairs, node_off, _, _, _, _ = union_of(graphs, [1, 2])
        for k, (gi, s) in enumerate(pairs):
            a, b = int(node_off[k]), int(node_off[k + 1])
            assert abs(p[a:b].sum() - 1.0) <= 1e-12, (prec, gi)
            assert part[a:b].min() >= a and part[a:b].max() < b, (prec, gi)
            if graphs[gi].m == 0:
                assert res[(gi, s)][2] == 0.0 and res[(gi, s)][1] == 1
                continue
            ref = _oracle_codelength(graphs[gi], part[a:b] - a)
            tol = 1e-9 if prec == "fp64" else 1e-6
            assert abs(res[(gi, s)][2] - ref) <= tol * max(abs(ref), 1e-12), (prec, gi, res[(gi, s)][2], ref)


# ------------------------------------------------------------ graph capture
def _run(setup, graph_on, hub_mode=0, precision="mixed"):
    from relaxmap_amd import Network
    with Network(precision=precision) as net:
        net.set_graph_capture(graph_on)
        net.set_hub_path(hub_mode)
        setup(net)
        net.initiate()
        r = net.run(seed=3)
        return r.computed_codelength, r.sweeps, net.partition().tolist()


def test_graph_capture_same_results_batch():
    """Sweeps of a batch of instances as one captured graph each give the
    results of the same launches made one by one with host waits."""
    from relaxmap_amd.batch import union_of
    graphs = _graphs()[:4]
    _, node_off, iseeds, ptr, idx, w = union_of(graphs, [1, 2])
    setup = lambda net: net.batch_upload(node_off, iseeds, ptr, idx, w)
    for prec in ("mixed", "fp64"):
        on, off = _run(setup, True, precision=prec), _run(setup, False, precision=prec)
        print(f"[gpu] batch graph on/off {prec}: L {on[0]:.12f} / {off[0]:.12f}, sweeps {on[1]} / {off[1]}", flush=True)
        assert on == off


def test_graph_capture_same_results_coarse_tune_with_hubs():
    """One network: the coarse tunes' sub-module sweeps run captured; with
    units of more than 2048 arcs, also with every hub sent through the
    overflow fallback (launched in every captured batch, working only on
    flagged hubs)."""
    from relaxmap_amd import synth
    g = synth.chunglu(60000, 2.1, 4, 20000, 16.0, 1.2, 50, 5000, 0.4, 0.01, 11)
    setup = lambda net: g.upload_to(net)
    for hub_mode in (0, 2):
        on, off = _run(setup, True, hub_mode), _run(setup, False, hub_mode)
        print(f"[gpu] coarse graph on/off hub mode {hub_mode}: L {on[0]:.12f} / {off[0]:.12f}", flush=True)
        assert on == off
