"""Rows, modules and instances too large for one small lane group: the paths
that hand them to whole workgroups, chunk by chunk, against the oracle.

- Rows of more than 1024 arcs are keyed in chunks by the contraction
  (Network::convertModulesToSuperNodes, Module.cpp:3008-3077): the super graph
  of a network with a hub of 2600 out- and in-arcs has the oracle's integer
  structure bit for bit and its weights within 1e-13 relative, in FP64 and
  (from the FP32 flows, 1e-5) in mixed precision.
- Modules of more than 2048 members are summed in chunks by the from-scratch
  codelength (Network::calculateCodeLength, Module.cpp:2966-3002): equal to
  the oracle's within 1e-10 relative, and to H(p) for one module of all nodes.
- Coarse-tune instances of more than 2048 nodes have their totals summed in
  chunks, and the intra-module arcs of the hub are flagged in chunks
  (generate_sub_modules, OmpRelaxmap.cpp:500-632): the sub-modules lie in
  their parents, no instance ends above its singletons (numpy map equation of
  the induced subgraphs), and the super nodes from the sub-modules have the
  oracle's structure (Module.cpp:3083-3158)."""
import numpy as np
import pytest

from moves_common import csr, hub_network
from submodule_eval import instance_codelengths

pytestmark = pytest.mark.gpu


def _pair(precision="fp64", seed=5):
    from oracle.pyoracle import OracleNet
    from relaxmap_amd import Network
    n, src, dst = hub_network(seed=seed)
    rng = np.random.default_rng(seed)
    w = rng.random(len(src)) + 0.05
    ptr, idx = csr(n, src, dst)
    order = np.lexsort((dst, src))   # csr's arc order, for the weights
    o = OracleNet.from_arcs(n, src[order], dst[order], w[order]).initiate()
    net = Network(precision=precision)
    net.upload(n, ptr, idx, w[order])
    net.initiate()
    assert int(np.diff(ptr).max()) > 1024
    return n, o, net


def _blocks(n, seed):
    """Modules of the 60-node planted blocks, with a tenth of the nodes moved
    to other blocks' modules (so the hub's arcs reach many super nodes)."""
    rng = np.random.default_rng(seed)
    part = (np.arange(n) // 60 * 60).astype(np.int32)
    idx = rng.choice(n, n // 10, replace=False)
    part[idx] = part[rng.integers(0, n, len(idx))]
    return part


def _compare_super(s, r, rtol_w):
    for k in ("out_ptr", "in_ptr", "out_idx", "in_idx", "mod", "n_members"):
        assert np.array_equal(np.asarray(s[k], dtype=np.int64), np.asarray(r[k], dtype=np.int64)), k
    assert np.allclose(s["out_w"], r["out_w"], rtol=rtol_w, atol=0)
    assert np.allclose(s["in_w"], r["in_w"], rtol=rtol_w, atol=0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_contraction_with_hub_rows(precision):
    n, o, net = _pair(precision)
    with net:
        part = _blocks(n, 3)
        net.set_partition(part)
        net.update_members_in_module()
        ns, ms = net.convert_modules_to_super_nodes()
        s = net.super_nodes()
    o.set_partition(part)
    o.update_members()
    o.contract()
    r = o.supernodes()
    print(f"[gpu] hub contraction {precision}: {ns} super nodes, {ms} super arcs", flush=True)
    assert ns == len(r["mod"]) and ms == len(r["out_idx"])
    _compare_super(s, r, 1e-13 if precision == "fp64" else 1e-5)


def test_codelength_of_large_modules():
    n, o, net = _pair()
    with net:
        p = net.flows()
        cases = {
            "two halves": (np.arange(n) >= n // 2).astype(np.int32) * (n // 2),
            "2500 + 3500": (np.arange(n) >= 2500).astype(np.int32) * 2500,
            "one 3000-node module, blocks elsewhere": np.where(np.arange(n) < 3000, 0,
                                                             np.arange(n) // 60 * 60).astype(np.int32),
        }
        for tag, part in cases.items():
            net.set_partition(part)
            got = net.calculate_codelength()
            o.set_partition(part)
            ref = o.calculate_codelength()
            print(f"[gpu] large modules, {tag}: {got:.15f} vs oracle {ref:.15f}", flush=True)
            assert abs(got - ref) <= 1e-10 * abs(ref), tag
        net.set_partition(np.zeros(n, np.int32))   # one module: q = 0, L = H(p)
        one = net.calculate_codelength()
    hp = -float(np.sum(np.where(p > 0, p * np.log(np.where(p > 0, p, 1.0)), 0.0)))
    assert abs(one - hp) <= 1e-10 * abs(hp)


def test_coarse_tune_of_large_instances():
    n, o, net = _pair()
    part = (np.arange(n) >= 3000).astype(np.int32) * 3000   # two instances of 3000 nodes, the hub in the first
    with net:
        net.set_partition(part)
        net.update_members_in_module()
        L_before = net.codelength
        k = net.generate_sub_modules(threshold=1e-3, max_iter=10, seed=3)
        sub = net.sub_modules()
        assert net.codelength == L_before
        ns, ms = net.generate_super_nodes_from_sub_modules()
        s = net.super_nodes()
        L = net.codelength
        assert abs(L - net.calculate_codelength()) <= 1e-9 * abs(L)
        p = net.flows()
    print(f"[gpu] coarse tune of two 3000-node instances: {k} sub-modules", flush=True)
    assert np.array_equal(part[sub], part)   # a sub-module id is a node of the same parent
    assert len(np.unique(sub)) == k == ns
    ptr, idx, f = o.out_csr()
    tau = o.teleport()
    single = instance_codelengths(p, tau, ptr, idx, f, part, np.arange(n))
    after = instance_codelengths(p, tau, ptr, idx, f, part, sub)
    gains = [single[I] - after[I] for I in single]
    assert min(gains) >= -1e-12
    assert max(gains) > 1e-9
    o.set_partition(sub)
    o.update_members()
    o.contract()
    r = o.supernodes()
    assert ms == len(r["out_idx"])
    for name in ("out_ptr", "in_ptr", "out_idx", "in_idx", "n_members"):
        assert np.array_equal(np.asarray(s[name], dtype=np.int64), np.asarray(r[name], dtype=np.int64)), name
    assert np.array_equal(s["mod"], part[r["mod"]])   # super node in its parent module
    assert np.allclose(s["out_w"], r["out_w"], rtol=1e-13, atol=0)
