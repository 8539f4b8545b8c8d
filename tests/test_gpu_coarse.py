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
- the super graph has the structure of the oracle's contraction of the
  sub-module partition (integer structure bit-exact, weights 1e-13 relative),
  with the parent module as each super node's module;
- the tracked codelength stays consistent through the coarse sweeps.
"""
import math

import numpy as np
import pytest

from conftest import INPUTS, data_path

pytestmark = pytest.mark.gpu

ALPHA, BETA = 0.15, 1.0 - 0.15


def plogp(x):
    x = np.asarray(x, dtype=np.float64)
    return np.where(x > 0, x * np.log(np.where(x > 0, x, 1.0)), 0.0)


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


def _instance_codelengths(p, tau, ptr, idx, f, part, labels):
    """Codelength of every parent module's induced subgraph under `labels`
    (map equation with the members' own p, tau and dangling flow; exits count
    only arcs inside the parent)."""
    n = len(p)
    src = np.repeat(np.arange(n), np.diff(ptr))
    dangling = np.diff(ptr) == 0
    d = np.where(dangling, p, 0.0)
    inside = part[src] == part[idx]
    cross = inside & (labels[src] != labels[idx])
    X = np.zeros(n)
    np.add.at(X, labels[src[cross]], f[cross])
    P, T, D = np.zeros(n), np.zeros(n), np.zeros(n)
    np.add.at(P, labels, p)
    np.add.at(T, labels, tau)
    np.add.at(D, labels, d)
    used = np.zeros(n, dtype=bool)
    used[labels] = True
    q = (ALPHA * P + BETA * D) * (1.0 - T) + BETA * X
    out = {}
    for I in np.unique(part):
        mods = np.unique(labels[part == I])
        S = q[mods].sum()
        out[I] = float(plogp(S) - 2 * plogp(q[mods]).sum() + plogp(q[mods] + P[mods]).sum()
                       - plogp(p[part == I]).sum())
    return out


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
