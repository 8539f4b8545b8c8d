"""GPU path (librelaxmap.so through the C ABI) against the oracle and the
reference's golden fixtures.

Tolerances (written per check): steady-state flows 1e-12 in L1; codelength
of a fixed partition 1e-10 relative; contraction bit-exact in its integer
structure, arc weights 1e-13 relative (the weights are sums of flows, which
carry the flows' last-bit differences); sweeps are relaxed (batched moves),
so partitions are judged by codelength consistency and determinism here and
by final quality in test_gpu_quality.py.
"""
import math

import numpy as np
import pytest

from conftest import INPUTS, data_path, golden

pytestmark = pytest.mark.gpu

ALL = list(INPUTS)
PART_KEYS = [k for k in ALL if k != "lfr5k_mu05"]


def _gpu(key, **kw):
    from relaxmap_amd import Network
    net = Network(**kw)
    net.load(data_path(INPUTS[key]))
    net.initiate()
    return net


@pytest.mark.parametrize("key", ALL)
def test_flows(key):
    g = golden(f"{key}.flows")
    with _gpu(key) as net:
        print(f"[gpu] flows {key} n={net.n} m={net.m}", flush=True)
        assert net.n == g["n"]
        p = net.flows()
        assert np.abs(p - np.array(g["p"])).sum() <= 1e-12            # L1, FP64
        assert abs(net.iterations - g["iterations"]) <= 1
        q = net.node_exit()
        assert np.allclose(q, np.array(g["q"]), rtol=1e-12, atol=1e-17)
        assert net.n_dangling == g["n_dangling"]
        assert abs(net.initial_codelength - g["codelength"]) <= 1e-10 * abs(g["codelength"])
        # tracked (fixed-point) statistics agree with the literal calibrate
        L, S, nm = net.tracked()
        assert nm == g["n"]
        assert abs(L - g["codelength"]) <= 1e-10 * abs(g["codelength"])
        assert abs(S - g["sum_exit"]) <= 1e-12


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k"])
def test_flows_mixed_precision(key):
    g = golden(f"{key}.flows")
    with _gpu(key, precision="mixed") as net:
        p = net.flows()
        err = np.abs(p - np.array(g["p"])).sum()
        print(f"[gpu] mixed flows {key} L1 error {err:.3e}", flush=True)
        assert err <= 1e-6                                            # L1 vs FP64


@pytest.mark.parametrize("key", PART_KEYS)
@pytest.mark.parametrize("pname", ["swept", "random"])
def test_partition_codelength(key, pname):
    g = golden(f"{key}.part_{pname}")
    with _gpu(key) as net:
        net.set_partition(np.array(g["partition"], dtype=np.int32))
        calc = net.calculate_codelength()
        ref = g["codelength"]["calculated"]
        assert abs(calc - ref) <= 1e-10 * abs(ref)
        L, S, nm = net.tracked()
        assert abs(L - g["codelength"]["codelength"]) <= 1e-10 * abs(ref)
        assert abs(S - g["codelength"]["sum_exit"]) <= 1e-12
        assert nm == g["codelength"]["n_module"]
        assert np.array_equal(net.partition(), np.array(g["partition"]))


@pytest.mark.parametrize("key", PART_KEYS)
@pytest.mark.parametrize("pname", ["swept", "random"])
def test_contraction(key, pname):
    g = golden(f"{key}.part_{pname}")
    c = g["contract"]
    with _gpu(key) as net:
        net.set_partition(np.array(g["partition"], dtype=np.int32))
        ns, ms = net.convert_modules_to_super_nodes()
        assert ns == c["n_super"]
        s = net.super_nodes()
        ref_out_idx = [e[0] for sp in c["super"] for e in sp["out"]]
        ref_out_w = [e[1] for sp in c["super"] for e in sp["out"]]
        ref_in_idx = [e[0] for sp in c["super"] for e in sp["in"]]
        ref_in_w = [e[1] for sp in c["super"] for e in sp["in"]]
        ref_out_ptr = np.cumsum([0] + [len(sp["out"]) for sp in c["super"]])
        ref_in_ptr = np.cumsum([0] + [len(sp["in"]) for sp in c["super"]])
        # integer structure: bit-exact
        assert ms == len(ref_out_idx)
        assert np.array_equal(s["out_ptr"], ref_out_ptr)
        assert np.array_equal(s["in_ptr"], ref_in_ptr)
        assert s["out_idx"].tolist() == ref_out_idx
        assert s["in_idx"].tolist() == ref_in_idx
        assert s["mod"].tolist() == [sp["mod"] for sp in c["super"]]
        assert s["n_members"].tolist() == [sp["n_members"] for sp in c["super"]]
        # weights: sums of flows, 1e-13 relative
        assert np.allclose(s["out_w"], ref_out_w, rtol=1e-13, atol=0)
        assert np.allclose(s["in_w"], ref_in_w, rtol=1e-13, atol=0)
        assert np.allclose(s["size"], [sp["size"] for sp in c["super"]], rtol=1e-12, atol=1e-18)
        assert np.allclose(s["tp"], [sp["tp"] for sp in c["super"]], rtol=1e-12, atol=1e-18)
        assert np.allclose(s["dangling"], [sp["dangling"] for sp in c["super"]], rtol=1e-12, atol=1e-18)
        assert np.allclose(s["exit"], [sp["exit"] for sp in c["super"]], rtol=1e-10, atol=1e-17)


@pytest.mark.parametrize("key", PART_KEYS)
def test_contraction_mixed_precision(key):
    """Mixed precision contracts the FP32 arc flows (carried through the key
    sort in place of the arc index): the integer structure is the reference's
    bit for bit, the super-arc weights and statistics are FP64 sums of FP32
    flows of the mixed steady state, within 1e-5 relative of the reference."""
    g = golden(f"{key}.part_swept")
    c = g["contract"]
    with _gpu(key, precision="mixed") as net:
        net.set_partition(np.array(g["partition"], dtype=np.int32))
        ns, ms = net.convert_modules_to_super_nodes()
        s = net.super_nodes()
    assert ns == c["n_super"]
    assert ms == sum(len(sp["out"]) for sp in c["super"])
    assert np.array_equal(s["out_ptr"], np.cumsum([0] + [len(sp["out"]) for sp in c["super"]]))
    assert np.array_equal(s["in_ptr"], np.cumsum([0] + [len(sp["in"]) for sp in c["super"]]))
    assert s["out_idx"].tolist() == [e[0] for sp in c["super"] for e in sp["out"]]
    assert s["in_idx"].tolist() == [e[0] for sp in c["super"] for e in sp["in"]]
    ref_out_w = [e[1] for sp in c["super"] for e in sp["out"]]
    ref_in_w = [e[1] for sp in c["super"] for e in sp["in"]]
    assert np.allclose(s["out_w"], ref_out_w, rtol=1e-5, atol=0)
    assert np.allclose(s["in_w"], ref_in_w, rtol=1e-5, atol=0)
    assert np.allclose(s["exit"], [sp["exit"] for sp in c["super"]], rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k", "lfr5k_mu05"])
def test_contraction_from_super_level(key):
    """Contracting an already contracted level (its super arcs summed again)
    gives the structure of contracting the original arcs under the same node
    partition (the oracle, which follows the reference); integer structure
    bit-exact, weights 1e-13 relative (sums of partial sums round
    differently), super-node statistics as for a fresh contraction."""
    from oracle.pyoracle import OracleNet
    with _gpu(key) as net:
        for _ in range(2):
            net.move(prioritized=True, seed=5)
        net.update_members_in_module()
        net.convert_modules_to_super_nodes()
        for _ in range(3):
            net.move(prioritized=True, seed=6)
        net.update_members_in_module()
        ns, ms = net.convert_modules_to_super_nodes()     # from the super level
        s = net.super_nodes()
        part = net.partition()
        print(f"[gpu] contraction of contraction {key}: {ns} super nodes, {ms} super arcs", flush=True)
    o = OracleNet.load(data_path(INPUTS[key])).initiate()
    o.set_partition(part)
    o.update_members()
    o.contract()
    r = o.supernodes()
    assert ns == len(r["mod"]) and ms == len(r["out_idx"])
    for k in ("out_ptr", "in_ptr", "out_idx", "in_idx", "mod", "n_members"):
        assert np.array_equal(np.asarray(s[k], dtype=np.int64), np.asarray(r[k], dtype=np.int64)), k
    assert np.allclose(s["out_w"], r["out_w"], rtol=1e-13, atol=0)
    assert np.allclose(s["in_w"], r["in_w"], rtol=1e-13, atol=0)
    assert np.allclose(s["size"], r["size"], rtol=1e-12, atol=1e-18)
    assert np.allclose(s["tp"], r["tp"], rtol=1e-12, atol=1e-18)
    assert np.allclose(s["dangling"], r["dangling"], rtol=1e-12, atol=1e-18)
    assert np.allclose(s["exit"], r["exit"], rtol=1e-10, atol=1e-17)


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k", "lfr5k_mu05"])
@pytest.mark.parametrize("prior", [True, False])
def test_sweeps_consistent(key, prior):
    """The tracked codelength after every sweep equals the from-scratch
    codelength of the partition it describes."""
    with _gpu(key) as net:
        L_prev = net.codelength
        for it in range(4):
            moved = net.move(prioritized=prior, seed=3)
            L = net.codelength
            calc = net.calculate_codelength()
            print(f"[gpu] sweep {key} prior={prior} it={it} moved={moved} L={L / math.log(2):.6f}", flush=True)
            assert abs(L - calc) <= 1e-9 * abs(calc)
            assert L <= L_prev + 1e-9
            L_prev = L
        part = net.partition()
        assert len(np.unique(part)) == net.n_module


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k"])
def test_sweeps_deterministic(key):
    parts = []
    for _ in range(2):
        with _gpu(key) as net:
            for _ in range(3):
                net.move(prioritized=True, seed=11)
            parts.append((net.partition(), net.codelength))
    assert np.array_equal(parts[0][0], parts[1][0])
    assert parts[0][1] == parts[1][1]


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k"])
def test_supernode_sweeps_consistent(key):
    with _gpu(key) as net:
        for _ in range(2):
            net.move(prioritized=True, seed=5)
        net.update_members_in_module()
        net.convert_modules_to_super_nodes()
        for _ in range(3):
            net.move(prioritized=True, seed=6)
            L = net.codelength
            calc = net.calculate_codelength()
            assert abs(L - calc) <= 1e-9 * abs(calc)
        net.to_node_level()
        net.move(prioritized=True, seed=7)
        assert abs(net.codelength - net.calculate_codelength()) <= 1e-9 * abs(net.codelength)


def _activity_graph():
    """60k SBM nodes with teleport weight 1 and 20k nodes with weight 1e-6
    on a cycle of their own.  5k of the light nodes have an arc into the SBM
    and 5k an arc from it; the other 10k see the SBM only through them."""
    from relaxmap_amd import synth
    g = synth.sbm(60000, 600, 10.0, 0.2, 3)
    na, nb = 60000, 20000
    src = [np.repeat(np.arange(na), np.diff(g.ptr)), na + np.arange(nb), na + np.arange(5000),
           (np.arange(5000, 10000) * 13) % na]
    dst = [g.idx.astype(np.int64), na + (np.arange(nb) + 1) % nb, (np.arange(5000) * 7) % na,
           na + np.arange(5000, 10000)]
    w = [g.w, np.ones(nb), np.ones(5000), np.ones(5000)]
    src, dst, w = np.concatenate(src), np.concatenate(dst), np.concatenate(w)
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    n = na + nb
    ptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=ptr[1:])
    node_w = np.concatenate([np.ones(na), np.full(nb, 1e-6)])
    return n, ptr, src, dst.astype(np.int32), w, node_w


@pytest.mark.parametrize("batches", [1, 4])
def test_activity_after_a_sweep(batches):
    """A prioritized sweep makes every out- and in-neighbour of its movers
    active for the next one (Module.cpp:1661-1668), so the next sweep visits
    exactly those units' arcs.  With vthresh 1e-8 the SBM nodes move and the
    light cycle does not (its gains are about 1e-6 times smaller), so part
    of the cycle stays inactive, part is reached through out-arcs and part
    through in-arcs.  One batch of 80k units with more than half moving
    takes the receiving-side pass for its marks; four batches of 20k push
    them from the movers.  From singletons every mover has a neighbour
    module to go to, so the movers are the nodes whose module changed."""
    from relaxmap_amd import Network
    n, ptr, src, dst, w, node_w = _activity_graph()
    deg = np.diff(ptr) + np.bincount(dst, minlength=n)
    with Network(batches=batches) as net:
        net.upload(n, ptr, dst, w, node_w)
        net.initiate()
        before = net.partition()
        net.move(prioritized=True, vthresh=1e-8, seed=3)
        mover = net.partition() != before
        active = np.zeros(n, bool)
        active[dst[mover[src]]] = True
        active[src[mover[dst]]] = True
        net.move(prioritized=True, vthresh=1e-8, seed=4)
        print(f"[gpu] activity batches={batches}: movers {mover.sum()}, active {active.sum()} "
              f"(light {active[60000:].sum()}), visits {net.last_edge_visits}", flush=True)
        assert mover.sum() >= n // 2
        assert 0 < active[60000:].sum() < 20000
        assert net.last_edge_visits == int(deg[active].sum())


def _hub_graph():
    """60k nodes, 753k arcs, 23 units with more than 2048 arcs (up to 4786):
    several partitions and arc chunks per hub."""
    from relaxmap_amd import synth
    return synth.chunglu(60000, 2.1, 4, 20000, 16.0, 1.2, 50, 5000, 0.4, 0.01, 11)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hub_paths_agree(precision):
    """Units with more than 2048 arcs: evaluation by LDS partitions (default),
    by one global hash table per unit, and by partitions with every hub sent
    through the overflow fallback make identical moves (all compute
    bit-identical dL from the same exact integer flows), at the node level and
    after a contraction; the tracked codelength stays consistent."""
    from relaxmap_amd import Network
    g = _hub_graph()
    out = []
    for mode in (0, 1, 2):
        with Network(precision=precision) as net:
            g.upload_to(net)
            net.set_hub_path(mode)
            net.initiate()
            trace = []
            for _ in range(3):
                trace.append(net.move(prioritized=True, seed=5))
                L = net.codelength
                assert abs(L - net.calculate_codelength()) <= 1e-9 * abs(L)
            net.update_members_in_module()
            net.convert_modules_to_super_nodes()
            for _ in range(2):
                trace.append(net.move(prioritized=True, seed=6))
            L = net.codelength
            assert abs(L - net.calculate_codelength()) <= 1e-9 * abs(L)
            print(f"[gpu] hub path {mode} {precision}: moved {trace} L={L / math.log(2):.6f}", flush=True)
            out.append((net.partition(), trace, L))
    for k in (1, 2):
        assert out[0][1] == out[k][1]
        assert np.array_equal(out[0][0], out[k][0])
        assert out[0][2] == out[k][2]


def test_hub_paths_agree_full_run():
    from relaxmap_amd import Network
    g = _hub_graph()
    out = []
    for mode in (0, 1, 2):
        with Network() as net:
            g.upload_to(net)
            net.set_hub_path(mode)
            net.initiate()
            r = net.run(seed=3)
            out.append((net.partition(), r.codelength))
    for k in (1, 2):
        assert np.array_equal(out[0][0], out[k][0]) and out[0][1] == out[k][1]
