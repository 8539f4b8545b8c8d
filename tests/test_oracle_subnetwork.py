"""orc_subnetwork (generate_network_from_module, OmpRelaxmap.cpp:640-695)
pinned by what fixes it independently:
- the network of all nodes is the network itself: the same singleton
  exits and codelength bit for bit, and four move() sweeps from seed 7 give
  the reference's own trajectory (tests/golden/*.sweeps_move_s7);
- the singleton codelength of a module's network equals the map equation of
  its induced subgraph computed in numpy (exits over arcs inside it only),
  and so does its codelength after partition_module_network's sweeps.
"""
import numpy as np
import pytest

from conftest import INPUTS, data_path, golden
from submodule_eval import instance_codelengths


@pytest.mark.parametrize("key", ["flow_net", "lfr1k_mu03", "sbm1k"])
def test_whole_network_as_its_own_module(key):
    from oracle.pyoracle import OracleNet
    o = OracleNet.load(data_path(INPUTS[key])).initiate()
    sub = o.subnetwork(np.arange(o.n))
    assert sub.codelength == o.codelength
    assert np.array_equal(sub.node_exit(), o.node_exit())
    g = golden(f"{key}.sweeps_move_s7")
    sub.seed(7)
    sub.reset_active()
    for s in g["sweeps"]:
        assert sub.sweep("move") == s["moved"]
        assert sub.codelength == s["codelength"]
        assert np.array_equal(sub.partition(), np.array(s["partition"]))


@pytest.mark.parametrize("key", ["lfr1k_mu03", "sbm1k"])
def test_module_networks_match_the_induced_map_equation(key):
    from oracle.pyoracle import OracleNet
    o = OracleNet.load(data_path(INPUTS[key])).initiate()
    part = np.array(golden(f"{key}.part_swept")["partition"], dtype=np.int64)
    ptr, idx, f = o.out_csr()
    p, tau = o.flows(), o.teleport()
    single = instance_codelengths(p, tau, ptr, idx, f, part, np.arange(o.n))
    labels = np.arange(o.n)
    checked = 0
    for I in np.unique(part):
        members = np.flatnonzero(part == I)
        if len(members) < 2:
            continue
        sub = o.subnetwork(members)
        assert abs(sub.codelength - single[int(I)]) <= 1e-12 * max(1.0, abs(single[int(I)]))
        lab = sub.partition_module_network(threshold=1e-3, max_iter=10, seed=int(I) + 1)
        labels[members] = members[lab]
        checked += 1
        after = sub.codelength
        got = instance_codelengths(p, tau, ptr, idx, f, part, labels)[int(I)]
        assert abs(after - got) <= 1e-11 * max(1.0, abs(got))
    assert checked > 0
