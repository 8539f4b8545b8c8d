"""Final partitions of the GPU driver against the reference's.

Move order is relaxed on the GPU, so the final partition is judged, as the
specification of this path states, by its codelength (within 0.5% of the
reference's) and by NMI to the planted communities.  The reference numbers
come from tests/golden/*.run_*_s1 (the reference's own driver at one thread,
up to its first coarse tune).
"""
import math

import numpy as np
import pytest

from conftest import DATA, INPUTS, data_path, golden

pytestmark = pytest.mark.gpu


def nmi(a, b):
    """Normalised mutual information 2 I(A;B) / (H(A) + H(B)) (natural log)."""
    a = np.unique(np.asarray(a), return_inverse=True)[1]
    b = np.unique(np.asarray(b), return_inverse=True)[1]
    n = len(a)
    cont = np.zeros((a.max() + 1, b.max() + 1))
    np.add.at(cont, (a, b), 1.0)
    pij = cont / n
    pi, pj = pij.sum(1), pij.sum(0)
    nz = pij > 0
    mi = (pij[nz] * np.log(pij[nz] / np.outer(pi, pj)[nz])).sum()
    h = lambda p: -(p[p > 0] * np.log(p[p > 0])).sum()
    return 2 * mi / (h(pi) + h(pj))


def planted(key, names):
    files = {"lfr1k_mu03": "directNet_1k_sm_mu_03_community.dat", "lfr5k_mu05": "directNet_5k_sm_mu_05_community.dat",
             "sbm1k": "sbm_1k_c25_s1_community.dat"}
    lab = {}
    with open(f"{DATA}/{files[key]}") as f:
        for line in f:
            t = line.split()
            if len(t) >= 2:
                lab[t[0]] = int(t[1])
    return np.array([lab[x] for x in names])


def test_nmi_closed_forms():
    x = np.array([0, 0, 1, 1, 2, 2])
    assert abs(nmi(x, x) - 1.0) < 1e-12
    assert abs(nmi(x, [5, 5, 7, 7, 9, 9]) - 1.0) < 1e-12
    assert abs(nmi([0, 1, 0, 1], [0, 0, 1, 1])) < 1e-12


@pytest.mark.parametrize("batches", [0, 8])
@pytest.mark.parametrize("key", ["lfr1k_mu03", "lfr5k_mu05", "sbm1k"])
@pytest.mark.parametrize("mode", ["prior", "normal"])
def test_codelength_and_nmi(key, mode, batches):
    """batches=0 is the default, per-level sub-sweep count (4 at these sizes)."""
    from relaxmap_amd import Network
    g = golden(f"{key}.run_{mode}_s1")
    ref_L = g["calculated"]
    with Network(batches=batches) as net:
        net.load(data_path(INPUTS[key]))
        net.initiate()
        r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=(mode == "prior"), seed=1)
        L = r.computed_codelength
        print(f"[gpu] {key} {mode} batches={batches}: L={L / math.log(2):.5f} ref={ref_L / math.log(2):.5f} bits, "
              f"modules={r.n_module}, steps={r.super_steps}", flush=True)
        assert abs(r.codelength - L) <= 1e-9 * L
        assert L <= ref_L * 1.005
        part = net.partition()
        # the names of .txt nodes are their original ids
        from oracle.pyoracle import OracleNet
        o = OracleNet.load(data_path(INPUTS[key]))
        names = [o.name(i) for i in range(o.n)]
        truth = planted(key, names)
        ref_nmi = nmi(np.array(g["partition"]), truth)
        got = nmi(part, truth)
        print(f"[gpu] {key} {mode}: NMI={got:.4f} ref NMI={ref_nmi:.4f}", flush=True)
        assert got >= ref_nmi - 0.02


def test_flow_net_four_modules(tmp_path):
    from relaxmap_amd import Network
    with Network() as net:
        net.load(data_path("flow.net"))
        net.initiate()
        net.run(threshold=1e-3, prior=True, seed=1)
        part = net.partition()
        blocks = [set(part[4 * k:4 * k + 4]) for k in range(4)]
        assert all(len(b) == 1 for b in blocks) and len(set.union(*blocks)) == 4
        net.write_tree(tmp_path / "flow.tree")
        net.write_clu(tmp_path / "flow.clu")
    tree = (tmp_path / "flow.tree").read_text().splitlines()
    assert tree[0].startswith("# Code length ") and tree[0].endswith(" in 4 modules.")
    assert len(tree) == 17
    clu = (tmp_path / "flow.clu").read_bytes()
    assert clu.startswith(b"*Vertices 16\r\n") and clu.count(b"\r\n") == 17


def test_determinism_of_run():
    from relaxmap_amd import Network
    out = []
    for _ in range(2):
        with Network() as net:
            net.load(data_path(INPUTS["lfr1k_mu03"]))
            net.initiate()
            r = net.run(seed=9)
            out.append((net.partition(), r.codelength))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


def test_rerun_on_same_network_repeats():
    """initiate + run twice on one network gives the same partition."""
    from relaxmap_amd import Network
    with Network() as net:
        net.load(data_path(INPUTS["lfr5k_mu05"]))
        out = []
        for _ in range(2):
            net.initiate()
            r = net.run(seed=3)
            out.append((net.partition(), r.codelength))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_per_level_batches_small_network(prec):
    """Per-level sub-sweeps (batches=0) give 4 sub-sweeps at every level of a
    network of fewer than 163840 nodes: the same run as batches=4, bit for bit."""
    from relaxmap_amd import Network
    out = []
    for batches in (0, 4):
        with Network(precision=prec, batches=batches) as net:
            net.load(data_path(INPUTS["lfr5k_mu05"]))
            net.initiate()
            r = net.run(seed=4)
            out.append((net.partition(), r.codelength, r.sweeps))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1] and out[0][2] == out[1][2]


def test_per_level_batches_quality_100k_nodes():
    """On the 100k-node LFR-style network (4 sub-sweeps at every level) the
    per-level count reaches the codelength and NMI of 8 sub-sweeps per sweep."""
    from relaxmap_amd import Network, synth
    g = synth.config2()
    res = {}
    for batches in (0, 8):
        with Network(precision="mixed", batches=batches) as net:
            g.upload_to(net)
            net.initiate()
            r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
            res[batches] = (r.computed_codelength, nmi(net.partition(), g.community), r.sweeps)
    print(f"[gpu] 100k per-level batches: {res}", flush=True)
    assert res[0][0] <= res[8][0] * 1.001
    assert res[0][1] >= res[8][1] - 0.005


@pytest.mark.parametrize("key", ["flow_net", "lfr1k_mu03", "lfr5k_mu05", "sbm1k", "lfr100k"])
def test_codelength_against_reference_cli_runs(key):
    """The reference command line's whole schedule (coarse tunes included) on
    every thread of the machine that wrote tests/golden/reference_runs, five
    seeds: the GPU's final codelength, seeds 1-3, stays within 0.5 % of the
    median of those runs (6 printed digits)."""
    from relaxmap_amd import Network, synth
    runs = golden("reference_runs")
    ref = [r["computed_codelength_bits"] for r in runs["runs"][key] if r["returncode"] == 0]
    assert len(ref) >= 3
    med = float(np.median(ref))
    g = synth.config2() if key == "lfr100k" else None
    for seed in (1, 2, 3):
        with Network() as net:
            if g is not None:
                g.upload_to(net)
            else:
                net.load(data_path(INPUTS[key]))
            net.initiate()
            r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=seed)
        bits = r.computed_codelength / math.log(2)
        print(f"[gpu] {key} seed {seed}: L={bits:.5f} bits, reference CLI median {med:.5f} "
              f"({runs['threads']} threads, {len(ref)} seeds)", flush=True)
        assert bits <= med * 1.005
