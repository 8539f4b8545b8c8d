"""Coarse tune on the GPU: sub-module partitions of every module
(generate_sub_modules, OmpRelaxmap.cpp:500-632) and super nodes from
sub-modules (Network::generateSuperNodesFromSubModules, Module.cpp:3083-3158).

The reference's sub-partitions use an unseeded RNG (OmpRelaxmap.cpp:529), so
no trajectory can be compared. What is checked:
- every sub-module lies inside its parent module;
- no instance's own codelength ends above its singletons' (the reference's
  sequential moves guarantee it; relaxed sub-sweeps reset such an instance)
  and some improve, with the map equation of the induced subgraph computed
  independently in numpy;
- the sub-module partitions are as good as the reference's sequential ones:
  over all instances, the codelength total of the GPU's relaxed sub-sweeps
  is within 0.5 % of partition_module_network's move() sweeps on the same
  induced networks (the oracle, mean of three seeds);
- the super graph has the structure of the oracle's contraction of the
  sub-module partition (integer structure bit-exact, weights 1e-13 relative),
  with the parent module as each super node's module;
- the tracked codelength stays consistent through the coarse sweeps.
"""
import math

import numpy as np
import pytest

from conftest import INPUTS, data_path
from submodule_eval import instance_codelengths

pytestmark = pytest.mark.gpu

def _prepared(key, sweeps=2):
    """A network after a few node sweeps (modules with several members)."""
    from relaxmap_amd import Network
    net = Network()
    net.load(data_path(INPUTS[key]))
    net.initiate()
    for _ in range(sweeps):
        net.move(prioritized=True, seed=5)
    net.update_members_in_module()
    return net


_instance_codelengths = instance_codelengths


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k", "lfr5k_mu05"])
def test_sub_modules_within_parents_and_improve(key):
    from oracle.pyoracle import OracleNet
    with _prepared(key) as net:
        part = net.partition()
        L_before = net.codelength
        k = net.generate_sub_modules(threshold=1e-3, max_iter=10, seed=3)
        sub = net.sub_modules()
        print(f"[gpu] {key}: {len(np.unique(part))} modules -> {k} sub-modules", flush=True)
        assert net.codelength == L_before                 # the partition itself is untouched
        assert np.array_equal(net.partition(), part)
        assert np.array_equal(part[sub], part)            # a sub-module id is a node of the same parent
        assert len(np.unique(sub)) == k
        p = net.flows()
    o = OracleNet.load(data_path(INPUTS[key])).initiate()
    ptr, idx, f = o.out_csr()
    tau = o.teleport()
    single = _instance_codelengths(p, tau, ptr, idx, f, part, np.arange(len(p)))
    after = _instance_codelengths(p, tau, ptr, idx, f, part, sub)
    gains = [single[I] - after[I] for I in single]
    assert min(gains) >= -1e-12
    assert sum(g > 1e-9 for g in gains) > 0


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k"])
def test_super_nodes_from_sub_modules(key):
    from oracle.pyoracle import OracleNet
    with _prepared(key) as net:
        part = net.partition()
        k = net.generate_sub_modules(threshold=1e-3, max_iter=10, seed=3)
        sub = net.sub_modules()
        ns, ms = net.generate_super_nodes_from_sub_modules()
        s = net.super_nodes()
        L = net.codelength
        assert abs(L - net.calculate_codelength()) <= 1e-9 * abs(L)
    assert ns == k
    o = OracleNet.load(data_path(INPUTS[key])).initiate()
    o.set_partition(sub)
    o.update_members()
    o.contract()
    r = o.supernodes()
    assert ms == len(r["out_idx"])
    for name in ("out_ptr", "in_ptr", "out_idx", "in_idx", "n_members"):
        assert np.array_equal(np.asarray(s[name], dtype=np.int64), np.asarray(r[name], dtype=np.int64)), name
    assert np.array_equal(s["mod"], part[r["mod"]])       # super node in its parent module
    assert np.allclose(s["out_w"], r["out_w"], rtol=1e-13, atol=0)
    assert np.allclose(s["in_w"], r["in_w"], rtol=1e-13, atol=0)
    assert np.allclose(s["size"], r["size"], rtol=1e-12, atol=1e-18)
    assert np.allclose(s["tp"], r["tp"], rtol=1e-12, atol=1e-18)
    assert np.allclose(s["dangling"], r["dangling"], rtol=1e-12, atol=1e-18)
    assert np.allclose(s["exit"], r["exit"], rtol=1e-10, atol=1e-17)   # the consistent exit of the sub-module


@pytest.mark.parametrize("key", ["lfr1k_mu03", "lfr5k_mu05"])
def test_coarse_sweeps_consistent(key):
    with _prepared(key) as net:
        net.generate_sub_modules(threshold=1e-3, max_iter=10, seed=3)
        net.generate_super_nodes_from_sub_modules()
        L_prev = net.codelength
        for it in range(4):
            net.move(prioritized=True, seed=8)
            L = net.codelength
            assert abs(L - net.calculate_codelength()) <= 1e-9 * abs(L)
            assert L <= L_prev + 1e-12
            L_prev = L
        net.to_node_level()
        net.move(prioritized=True, seed=9)
        assert abs(net.codelength - net.calculate_codelength()) <= 1e-9 * abs(net.codelength)


def test_coarse_tune_deterministic():
    out = []
    for _ in range(2):
        with _prepared("lfr5k_mu05") as net:
            net.generate_sub_modules(threshold=1e-3, max_iter=10, seed=4)
            out.append(net.sub_modules())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k", "lfr5k_mu05"])
def test_sub_module_quality_against_sequential_partitions(key):
    """generate_sub_modules partitions each small module with the sequential
    move() (OmpRelaxmap.cpp:534-552, partition_module_network at numTh 1,
    436-460).  The oracle runs exactly that on every module's induced network
    (orc_subnetwork); both partitions are scored with the same numpy map
    equation of the induced subgraph."""
    from oracle.pyoracle import OracleNet
    with _prepared(key) as net:
        part = net.partition()
        net.generate_sub_modules(threshold=1e-3, max_iter=10, seed=3)
        sub = net.sub_modules()
    o = OracleNet.load(data_path(INPUTS[key])).initiate()
    ptr, idx, f = o.out_csr()
    p, tau = o.flows(), o.teleport()
    gpu = instance_codelengths(p, tau, ptr, idx, f, part, sub)
    seq_tot = []
    for seed in (1, 2, 3):
        labels = np.arange(o.n)
        for I in np.unique(part):
            members = np.flatnonzero(part == I)
            if len(members) > 1:
                lab = o.subnetwork(members).partition_module_network(threshold=1e-3, max_iter=10,
                                                                     seed=seed * 7919 + int(I))
                labels[members] = members[lab]
        seq = instance_codelengths(p, tau, ptr, idx, f, part, labels)
        seq_tot.append(sum(seq.values()))
    g_tot, s_tot = sum(gpu.values()), float(np.mean(seq_tot))
    print(f"[gpu] {key}: sub-module codelength total GPU {g_tot:.6f}, sequential {s_tot:.6f} "
          f"(seeds {['%.6f' % x for x in seq_tot]}) nats over {len(gpu)} instances", flush=True)
    assert g_tot <= s_tot + 0.005 * abs(s_tot)
