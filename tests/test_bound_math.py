"""The candidate bound's inequalities against the oracle's own dL.

The move evaluation drops a candidate module when a lower bound of its dL,
computed from the unit's flow to it alone (and upper bounds of every module's
teleport weight and flow), lies above dL = 0.  Here that bound, restated in
FP64 (the kernels evaluate its terms in FP32 with a margin of 1e-3 of their
scale, added here too), is checked against the dL the oracle computes for
every candidate of every unit (Module.cpp:580-649 via orc_candidates): it may
never exceed it.  A dropped sign, term or index in the bound would push it
above some candidate's dL on these inputs.  The partitions cover singletons
(a level's first sweep), a partition after oracle sweeps and the planted
communities, on a golden network and on a heavy-tailed one with hubs.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

ALPHA, BETA = 0.15, 1.0 - 0.15


def _phi(x):
    x = np.asarray(x, dtype=np.float64)
    return np.where(x > 0, x * np.log(np.where(x > 0, x, 1.0)), 0.0)


def _bound_vs_oracle(net, rng_units=None):
    """(max of bound - dL over all candidates, share of candidates bounded above 0)."""
    n = net.n
    p = net.flows()
    tau = net.teleport()
    qu = net.node_exit()
    ptr, idx, af = net.out_csr()          # arc flows p_a w_ab / sum_b w_ab
    src = np.repeat(np.arange(n), np.diff(ptr))
    part = net.partition().astype(np.int64)
    st = net.module_stats()
    q, P, T, D = st["exit"], st["sum_pr"], st["sum_tp"], st["sum_dangling"]
    S = net.sum_exit
    dang = np.diff(ptr) == 0
    du = np.where(dang, p, 0.0)
    tele_out = ALPHA * p + BETA * du
    # flows of every unit to / from every neighbouring module
    flow_to = {}
    for a, b, f in zip(src, idx, af):
        flow_to.setdefault((a, part[b]), [0.0, 0.0])[0] += f
        flow_to.setdefault((b, part[a]), [0.0, 0.0])[1] += f
    tmax, pmax = T.max(), P.max()
    worst, dropped, total = -np.inf, 0, 0
    units = range(n) if rng_units is None else rng_units
    for u in units:
        mods, diffs = net.candidates(u)
        if len(mods) == 0:
            continue
        o = part[u]
        fo_o, fi_o = flow_to.get((u, o), (0.0, 0.0))
        outo = BETA * fo_o + tele_out[u] * (T[o] - tau[u])
        ino = BETA * fi_o + (ALPHA * (P[o] - p[u]) + BETA * (D[o] - du[u])) * tau[u]
        q1 = q[o] - qu[u] + outo + ino
        Kr = -2 * _phi(q1) + 2 * _phi(q[o]) + _phi(q1 + P[o] - p[u]) - _phi(q[o] + P[o])
        d1 = q1 - q[o]
        assert qu[u] <= p[u] * (1 + 1e-12)
        for m, dl in zip(mods, diffs):
            fo, fi = flow_to.get((u, m), (0.0, 0.0))
            dhi = qu[u] - BETA * (fo + fi)
            dlo = dhi - (tele_out[u] * tmax + pmax * tau[u])
            if not (dlo + p[u] >= 0):
                continue
            e = d1 + dlo
            if dlo < 0 and not (4 * (S + e) >= p[u] - dlo):
                continue
            if dlo >= 0:
                F = _phi(dlo + p[u]) - 2 * _phi(dlo)
            else:
                F = (p[u] - dlo) * np.log(p[u] - dlo) + 2 * dlo * np.log(2.0)
            lb = _phi(S + e) - _phi(S) + F + Kr
            scale = abs(d1) + 4 * max(abs(dlo), abs(dhi)) + p[u]
            worst = max(worst, (lb - dl) / scale)
            total += 1
            dropped += int(lb - (1e-3 * scale + 1e-11) > 0.0)
    return worst, dropped / max(total, 1)


def _nets():
    import pyoracle as po
    from relaxmap_amd import synth
    gold = po.OracleNet.load(os.path.join(ROOT, "tests", "golden", "data", "directNet_1k_sm_mu_03.txt"))
    g = synth.chunglu(6000, 2.1, 4, 1500, 16.0, 1.2, 50, 800, 0.4, 0.01, 5)
    src = np.repeat(np.arange(g.n), np.diff(g.ptr)).astype(np.int32)
    heavy = po.OracleNet.from_arcs(g.n, src, g.idx.astype(np.int32), np.asarray(g.w, dtype=np.float64))
    return [("lfr1k_mu03", gold, None), ("chunglu6k", heavy, g.community.copy())]


@pytest.mark.parametrize("which", [0, 1])
def test_bound_never_exceeds_the_oracle_dl(which):
    name, net, planted = _nets()[which]
    net.initiate()
    worst, share = _bound_vs_oracle(net)
    print(f"{name} singletons: max (bound - dL) / scale {worst:.3e}, bounded above 0: {share:.3f}")
    assert worst <= 0.0
    net.seed(1)
    net.reset_active()
    net.sweep("ppmove")
    net.sweep("ppmove")
    worst, share = _bound_vs_oracle(net)
    print(f"{name} after two sweeps: max (bound - dL) / scale {worst:.3e}, bounded above 0: {share:.3f}")
    assert worst <= 0.0
    assert share > 0.3   # the bound does drop candidates once modules have formed
    if planted is not None:
        net.set_partition(planted.astype(np.int32))
        worst, share = _bound_vs_oracle(net)
        print(f"{name} planted: max (bound - dL) / scale {worst:.3e}, bounded above 0: {share:.3f}")
        assert worst <= 0.0
