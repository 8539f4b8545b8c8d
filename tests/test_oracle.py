"""The oracle (oracle/liboracle.so) against the reference's own outputs.

Golden fixtures were written by tests/golden/make_golden.py from the
reference sources (numTh == 1).  At one thread the oracle restates the
reference operation by operation, so every comparison here is bit-exact.
The closed-form tests pin what the fixtures leave open.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import INPUTS, ROOT, data_path, golden

LN2 = math.log(2.0)


@pytest.fixture(scope="session", autouse=True)
def _built():
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def _net(key):
    from oracle.pyoracle import OracleNet
    return OracleNet.load(data_path(INPUTS[key])).initiate()


ALL = list(INPUTS)


# ---------------------------------------------------------------- goldens
@pytest.mark.parametrize("key", ALL)
def test_flows_bitexact(key):
    g = golden(f"{key}.flows")
    o = _net(key)
    assert o.n == g["n"]
    assert o.iterations == g["iterations"]
    assert o.n_dangling == g["n_dangling"]
    assert np.array_equal(o.flows(), np.array(g["p"]))
    assert np.array_equal(o.node_exit(), np.array(g["q"]))
    assert o.codelength == g["codelength"]
    assert o.sum_exit == g["sum_exit"]
    assert o.all_log_all == g["all_log_all"]


@pytest.mark.parametrize("key", ALL)
@pytest.mark.parametrize("variant", ["move", "prior", "pmove", "ppmove"])
def test_sweeps_bitexact(key, variant):
    name = f"{key}.sweeps_{variant}_s7"
    if not os.path.exists(os.path.join(ROOT, "tests", "golden", name + ".json.gz")):
        pytest.skip("no fixture for this variant")
    g = golden(name)
    o = _net(key)
    o.seed(7)
    o.reset_active()
    for s in g["sweeps"]:
        moved = o.sweep(variant)
        assert moved == s["moved"]
        assert o.codelength == s["codelength"]
        assert o.sum_exit == s["sum_exit"]
        assert o.n_module == s["n_module"]
        assert o.n_empty == s["n_empty"]
        assert np.array_equal(o.partition(), np.array(s["partition"]))
        if variant in ("prior", "ppmove"):
            assert o.n_active == s["n_active_next"]


PART_KEYS = [k for k in ALL if k != "lfr5k_mu05"]


@pytest.mark.parametrize("key", PART_KEYS)
@pytest.mark.parametrize("pname", ["swept", "random"])
def test_partition_codelength_bitexact(key, pname):
    g = golden(f"{key}.part_{pname}")
    o = _net(key)
    o.set_partition(g["partition"])
    assert o.codelength == g["codelength"]["codelength"]
    assert o.sum_exit == g["codelength"]["sum_exit"]
    assert o.calculate_codelength() == g["codelength"]["calculated"]
    assert o.n_module == g["codelength"]["n_module"]


@pytest.mark.parametrize("key", PART_KEYS)
@pytest.mark.parametrize("pname", ["swept", "random"])
def test_contract_bitexact(key, pname):
    g = golden(f"{key}.part_{pname}")
    o = _net(key)
    o.set_partition(g["partition"])
    ns = o.contract()
    c = g["contract"]
    assert ns == c["n_super"]
    s = o.supernodes()
    for i, sp in enumerate(c["super"]):
        assert s["size"][i] == sp["size"] and s["exit"][i] == sp["exit"]
        assert s["tp"][i] == sp["tp"] and s["dangling"][i] == sp["dangling"]
        assert s["mod"][i] == sp["mod"] and s["n_members"][i] == sp["n_members"]
        a, b = s["out_ptr"][i], s["out_ptr"][i + 1]
        assert s["out_idx"][a:b].tolist() == [e[0] for e in sp["out"]]
        assert s["out_w"][a:b].tolist() == [e[1] for e in sp["out"]]
        a, b = s["in_ptr"][i], s["in_ptr"][i + 1]
        assert s["in_idx"][a:b].tolist() == [e[0] for e in sp["in"]]
        assert s["in_w"][a:b].tolist() == [e[1] for e in sp["in"]]


@pytest.mark.parametrize("key", PART_KEYS)
@pytest.mark.parametrize("pname", ["swept", "random"])
@pytest.mark.parametrize("variant", ["move", "ppmove"])
def test_supernode_sweeps_bitexact(key, pname, variant):
    g = golden(f"{key}.part_{pname}")
    o = _net(key)
    o.set_partition(g["partition"])
    o.seed(11)
    o.contract()
    o.reset_active(sp=True)
    for s in g[f"spsweeps_{variant}"]["sweeps"]:
        assert o.sweep(variant, sp=True) == s["moved"]
        assert o.codelength == s["codelength"]
        assert np.array_equal(o.partition(), np.array(s["partition"]))


@pytest.mark.parametrize("key", ALL)
@pytest.mark.parametrize("mode", ["prior", "normal"])
def test_driver_bitexact(key, mode):
    g = golden(f"{key}.run_{mode}_s1")
    o = _net(key)
    o.seed(1)
    r = o.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=(mode == "prior"))
    assert r["steps"] == len(g["steps"])
    assert r["coarse_next"] == g["stopped_before_coarse"]
    assert r["moved"] == [m for st in g["steps"] for m in st["moved"]]
    assert o.codelength == g["steps"][-1]["codelength"]
    assert o.calculate_codelength() == g["calculated"]
    assert np.array_equal(o.partition(), np.array(g["partition"]))


# ---------------------------------------------------------- closed forms
def _arcs_net(n, src, dst, w=None):
    from oracle.pyoracle import OracleNet
    return OracleNet.from_arcs(n, src, dst, w).initiate()


def _fixed_point_residual(o):
    """|| p - ((alpha + beta D) tau + beta W^T p) ||_1 with W row-normalised."""
    p = o.flows()
    tau = o.teleport()
    ptr, idx, f = o.out_csr()
    deg = np.diff(ptr)
    src = np.repeat(np.arange(o.n), deg)
    what = f / p[src]                      # flows back to row-normalised weights
    dang = deg == 0
    D = p[dang].sum()
    y = (0.15 + 0.85 * D) * tau
    np.add.at(y, idx, 0.85 * what * p[src])
    return np.abs(y - p).sum()


def test_cycle_uniform_flow():
    n = 37
    o = _arcs_net(n, np.arange(n), (np.arange(n) + 1) % n)
    assert np.allclose(o.flows(), 1.0 / n, rtol=0, atol=1e-15)
    # one module holding every node has no exit: L = H(p)
    o.set_partition(np.zeros(n, dtype=np.int32))
    assert abs(o.calculate_codelength() - math.log(n)) < 1e-12


def test_flow_net_symmetry_and_partition():
    o = _net("flow_net")
    p = o.flows()
    # rotating the four 4-cycles by 4 maps the graph onto itself
    assert np.allclose(p, np.roll(p, 4), rtol=0, atol=1e-15)
    assert abs(p.sum() - 1.0) < 1e-14
    o.seed(1)
    o.run(threshold=1e-3, prior=True)
    part = o.partition()
    blocks = [set(part[4 * k:4 * k + 4]) for k in range(4)]
    assert all(len(b) == 1 for b in blocks)
    assert len(set.union(*blocks)) == 4


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k", "lfr5k_mu05"])
def test_stationarity(key):
    o = _net(key)
    # the reference stops at 200 iterations or at an L1 change <= 1e-15;
    # either way the fixed-point residual is tiny
    assert _fixed_point_residual(o) < 1e-12
    assert abs(o.flows().sum() - 1.0) < 1e-13


def test_dangling_stationarity():
    rng = np.random.default_rng(5)
    n = 300
    src = rng.integers(0, n - 30, 2000)        # nodes n-30..n-1 have no out-arcs
    dst = rng.integers(0, n, 2000)
    w = rng.random(2000) + 0.1
    o = _arcs_net(n, src, dst, w)
    assert o.n_dangling >= 30
    assert _fixed_point_residual(o) < 1e-12


def _codelength_from_scratch(o, part):
    """Map equation of a partition, computed independently of the oracle:
    q_M = (alpha P_M + beta D_M)(1 - T_M) + beta * (flow leaving M on arcs)."""
    p, tau = o.flows(), o.teleport()
    ptr, idx, f = o.out_csr()
    deg = np.diff(ptr)
    src = np.repeat(np.arange(o.n), deg)
    d = np.where(deg == 0, p, 0.0)
    mods = np.unique(part)
    L = 0.0
    qs, qp = [], []
    for m in mods:
        mem = part == m
        P, T, D = p[mem].sum(), tau[mem].sum(), d[mem].sum()
        ext = f[(part[src] == m) & (part[idx] != m)].sum()
        q = (0.15 * P + 0.85 * D) * (1 - T) + 0.85 * ext
        qs.append(q)
        qp.append(q + P)
    qs, qp = np.array(qs), np.array(qp)
    plogp = lambda x: np.where(x > 0, x * np.log(np.where(x > 0, x, 1)), 0.0)
    L = plogp(qs.sum()) - 2 * plogp(qs).sum() + plogp(qp).sum() - plogp(p).sum()
    return float(L)


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k"])
def test_codelength_matches_independent_formula(key):
    o = _net(key)
    rng = np.random.default_rng(2)
    for part in (rng.integers(0, 30, o.n), np.arange(o.n), np.zeros(o.n, dtype=int)):
        o.set_partition(part)
        ref = _codelength_from_scratch(o, np.asarray(part))
        assert abs(o.calculate_codelength() - ref) <= 1e-10 * abs(ref)


@pytest.mark.parametrize("variant", ["move", "pmove"])
def test_move_delta_equals_codelength_difference(variant):
    """Each committed single-node move changes the tracked codelength by
    exactly the recomputed difference (the map equation is exact)."""
    o = _net("lfr1k_mu03")
    o.seed(3)
    for _ in range(3):
        o.sweep(variant)
        o.update_members()
        assert abs(o.codelength - o.calculate_codelength()) < 1e-11


def test_contraction_invariants():
    o = _net("sbm1k")
    rng = np.random.default_rng(9)
    part = rng.integers(0, 40, o.n).astype(np.int32)
    o.set_partition(part)
    o.contract()
    s = o.supernodes()
    ptr, idx, f = o.out_csr()
    src = np.repeat(np.arange(o.n), np.diff(ptr))
    inter = part[src] != part[idx]
    # total inter-module flow is preserved
    assert abs(s["out_w"].sum() - f[inter].sum()) < 1e-12
    # out-lists are sorted and the in-lists are their transpose
    mods = np.unique(part)
    assert np.array_equal(s["mod"], mods)
    rank = {m: i for i, m in enumerate(mods)}
    dense = np.zeros((len(mods), len(mods)))
    np.add.at(dense, ([rank[m] for m in part[src[inter]]], [rank[m] for m in part[idx[inter]]]), f[inter])
    got = np.zeros_like(dense)
    so = np.repeat(np.arange(len(mods)), np.diff(s["out_ptr"]))
    got[so, s["out_idx"]] = s["out_w"]
    assert np.allclose(got, dense, rtol=1e-13, atol=0)
    si = np.repeat(np.arange(len(mods)), np.diff(s["in_ptr"]))
    got_in = np.zeros_like(dense)
    got_in[s["in_idx"], si] = s["in_w"]
    assert np.array_equal(got, got_in)


# ---------------------------------------------------------------- loaders
def test_loader_quirks(tmp_path):
    from oracle.pyoracle import OracleNet
    t = tmp_path / "q.txt"
    # comment; duplicate arc (summed); self-loop (dropped, counted); w <= 0
    # (dropped); a one-token line (second id re-reads the first: a
    # self-loop); a blank line (both ids parse from the cleared token "" as
    # 0: a phantom node 0 with a self-loop).  The reference reports 5 nodes,
    # 7 unique arcs, 1 aggregated link and 3 self-links on this file.
    t.write_text("# c\n10 20\n10 20 2\n20 30\n30 30\n30 10 -1\n40\n\n20 10\n")
    o = OracleNet.load(str(t))
    assert o.n == 5
    assert [o.name(i) for i in range(5)] == ["0", "10", "20", "30", "40"]
    assert o.n_double_links == 1
    assert o.n_selflinks == 3
    o.initiate()
    assert o.m == 3                              # 10->20 (w 3), 20->30, 20->10
    assert sorted(o.raw_out_w().tolist()) == [1.0, 1.0, 3.0]
    assert o.n_dangling == 3                     # 0, 30 (its arc has w < 0), 40


def test_loader_blank_line_phantom_node(tmp_path):
    from oracle.pyoracle import OracleNet
    t = tmp_path / "b.txt"
    t.write_text("1 2\n\n2 1\n")
    o = OracleNet.load(str(t))
    # the blank line re-reads the cleared token: atol("") == 0, a self-loop
    # on a phantom node 0 (FileIO.cpp:202-218)
    assert o.n == 3 and o.name(0) == "0" and o.n_selflinks == 1


def test_net_loader_errors(tmp_path):
    from oracle.pyoracle import OracleNet
    bad = tmp_path / "bad.net"
    bad.write_text("not pajek\n")
    with pytest.raises(ValueError):
        OracleNet.load(str(bad))
    bad.write_text("*Vertices 2\n1 \"a\"\n2 \"b\"\n*Nodes\n1 2\n")
    with pytest.raises(ValueError):
        OracleNet.load(str(bad))
    ok = tmp_path / "ok.net"
    ok.write_text("*Vertices 3\n1 \"a b\" 2.0\n2 x -1\n3 \"c\"\n*Arcs 2\n1 2\n2 3 0.5\n3 1 2\n")
    o = OracleNet.load(str(ok))
    assert [o.name(i) for i in range(3)] == ["a b", "x", "c"]
    o.initiate()
    assert np.allclose(o.teleport(), [0.5, 0.25, 0.25])
