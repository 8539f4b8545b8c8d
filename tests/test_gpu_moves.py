"""Per-unit move decisions on the GPU (rmx_evaluate_moves: the evaluation
kernels of the sweeps, one batch, nothing committed) against the oracle's
(orc_evaluate_all, pinned in test_oracle_moves.py) on the same partition:
the candidate dL of Module.cpp:580-649, the argmin with ties to the smaller
id (Module.cpp:651) and the empty-module option (Module.cpp:1435-1439), at
the node level and on super nodes, through every hub path.

Both apply the sweeps' batch rule (a unit alone in its module does not take
a larger-id module that also holds one unit).  Tolerances: FP64 chooses the
same module for every unit except exact near-ties, where both best dL agree
to the dL tolerance; dL within 1e-13 nats + 1e-10 relative (the GPU sums
module statistics in 2^-61 fixed point, the oracle in FP64).  Mixed
precision chooses on FP32 estimates: the exact dL of its choice (the
oracle's value for that candidate) is within 1e-9 nats + 1e-5 relative of
the exact best."""
import numpy as np
import pytest

from conftest import INPUTS, data_path
from moves_common import csr, empty_wins, exact_tie, hub_network, perturbed_partition

pytestmark = pytest.mark.gpu

ATOL, RTOL = 1e-13, 1e-10


def _pair(n, src, dst, precision="fp64", hub_path=0):
    from oracle.pyoracle import OracleNet
    from relaxmap_amd import Network
    o = OracleNet.from_arcs(n, src, dst).initiate()
    ptr, idx = csr(n, src, dst)
    net = Network(precision=precision)
    net.set_hub_path(hub_path)
    net.upload(n, ptr, idx, np.ones(len(idx)))
    net.initiate()
    return o, net


def _compare_fp64(bm, bd, om, od, tag):
    assert len(bm) == len(om)
    assert np.all(np.abs(bd - od) <= ATOL + RTOL * np.abs(od)), (tag, float(np.abs(bd - od).max()))
    diff = np.flatnonzero(bm != om)
    for u in diff:   # only near-ties may pick another module: their best dL agree
        assert bm[u] != -1 and om[u] != -1, (tag, u, bm[u], om[u], bd[u], od[u])
    assert len(diff) <= max(1, len(bm) // 1000), (tag, len(diff))
    print(f"[gpu] moves {tag}: same module for {len(bm) - len(diff)} of {len(bm)} units, "
          f"largest dL difference {float(np.abs(bd - od).max()):.2e} nats", flush=True)
    return len(diff)


def _check(o, net, part, tag, sp=False, mixed=False):
    if not sp:
        o.set_partition(part)
        net.set_partition(part)
    bm, bd = net.evaluate_moves()
    om, od = o.evaluate_all(sp=sp, no_singleton_swaps=True)
    moving = int((om != -1).sum())
    empty = int((om == o.n + 1).sum())
    print(f"[gpu] moves {tag}: {len(bm)} units, {moving} would move, {empty} to the empty module", flush=True)
    if not mixed:
        nd = _compare_fp64(bm, bd, om, od, tag)
        assert int((bm == o.n + 1).sum()) == empty or nd > 0
        return bm, om
    # mixed: the exact dL of the GPU's choice against the exact best
    for u in np.flatnonzero((bm != -1) | (om != -1)):
        if bm[u] == -1:
            assert od[u] > -1e-9 - 1e-5 * abs(od[u]), (tag, u, od[u])
            continue
        mods, diffs = o.candidates(u, sp=sp)
        k = np.flatnonzero(mods == bm[u])
        assert len(k) == 1, (tag, u, bm[u])
        exact = diffs[k[0]]
        assert exact <= od[u] + 1e-9 + 1e-5 * abs(od[u]), (tag, u, exact, od[u])
    return bm, om


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k", "lfr5k_mu05", "flow_net"])
def test_moves_match_oracle(key):
    from oracle.pyoracle import OracleNet
    from relaxmap_amd import Network
    path = data_path(INPUTS[key])
    o = OracleNet.load(path).initiate()
    with Network() as net:
        net.load(path)
        net.initiate()
        for seed in (3, 4):
            _check(o, net, perturbed_partition(o, seed), f"{key} seed {seed}")
        _check(o, net, np.arange(o.n, dtype=np.int32), f"{key} singletons")


def test_moves_on_super_nodes():
    from oracle.pyoracle import OracleNet
    from relaxmap_amd import Network
    path = data_path(INPUTS["lfr5k_mu05"])
    o = OracleNet.load(path).initiate()
    part = perturbed_partition(o, 7, frac=0.05, sweeps=1)
    with Network() as net:
        net.load(path)
        net.initiate()
        o.set_partition(part)
        net.set_partition(part)
        o.update_members()
        net.update_members_in_module()
        ns_o = o.contract()
        ns, _ = net.convert_modules_to_super_nodes()
        assert ns == ns_o
        _check(o, net, None, "lfr5k super nodes", sp=True)


def test_empty_module_and_tie_on_gpu():
    n, src, dst, part = empty_wins()
    o, net = _pair(n, src, dst)
    with net:
        bm, _ = _check(o, net, part, "empty module wins")
        assert bm[20] == n + 1 and bm[21] == 10
    n, src, dst, part = exact_tie()
    o, net = _pair(n, src, dst)
    with net:
        o.set_partition(part)
        net.set_partition(part)
        bm, bd = net.evaluate_moves()
        assert bm[0] == 1                      # the exact tie goes to the smaller module id
        mods, diffs = o.candidates(0)
        assert abs(bd[0] - diffs[0]) <= ATOL + RTOL * abs(diffs[0])


@pytest.mark.parametrize("hub_path", [0, 1, 2])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_moves_with_a_hub(hub_path, precision):
    n, src, dst = hub_network()
    o, net = _pair(n, src, dst, precision=precision, hub_path=hub_path)
    with net:
        part = perturbed_partition(o, 9, frac=0.1, sweeps=1)
        _check(o, net, part, f"hub path {hub_path} {precision}", mixed=precision == "mixed")
        _check(o, net, np.arange(n, dtype=np.int32), f"hub path {hub_path} {precision} singletons",
               mixed=precision == "mixed")
