"""The oracle's per-unit decision (orc_evaluate_all / orc_candidates: the
candidate dL of Module.cpp:580-649 as prioritize_parallelMove computes it at
Module.cpp:1355-1488, the strict-< argmin over ascending module ids of
Module.cpp:651, the empty-module option of Module.cpp:1435-1439) pinned by
what does not depend on that code:
- every candidate's dL equals calculateCodeLength after the single move minus
  before it (the exact identity of the map equation for one node's move),
  to 1e-12 nats;
- the candidate set is the modules of the unit's in- and out-neighbours other
  than its own, plus the empty module iff its module has more than one member
  and some module is empty;
- the argmin is the brute-force minimum over those codelength differences,
  ties to the smaller id, and none when no difference is negative;
- closed cases: a node without arcs in a foreign module leaves for the empty
  module; a mirror-image pair of modules ties exactly and the smaller id wins.
"""
import numpy as np
import pytest

from conftest import INPUTS, data_path
from moves_common import empty_wins, exact_tie, perturbed_partition

TOL = 1e-12   # nats, absolute: the dL terms are O(1) nats


def _neighbour_modules(o, part):
    ptr, idx, _ = o.out_csr()
    n = o.n
    src = np.repeat(np.arange(n), np.diff(ptr))
    nb = [set() for _ in range(n)]
    for a, b in zip(src, idx):
        nb[a].add(int(part[b]))
        nb[b].add(int(part[a]))
    return nb


def _p5_after(o, part, u, m):
    q = part.copy()
    if m == o.n + 1:   # the empty module: any module id without members
        m = int(np.setdiff1d(np.arange(o.n), part)[0])
    q[u] = m
    o.set_partition(q)
    return o.calculate_codelength()


def _check_units(o, part, units):
    o.set_partition(part)
    L0 = o.calculate_codelength()
    nb = _neighbour_modules(o, part)
    counts = np.bincount(part, minlength=o.n)
    has_empty = bool((counts == 0).any())
    cand = {u: o.candidates(u) for u in units}
    best_m, best_d = o.evaluate_all()
    worst = 0.0
    for u in units:
        mods, diffs = cand[u]
        want = sorted(nb[u] - {int(part[u])})
        if counts[part[u]] > 1 and has_empty:
            want.append(o.n + 1)
        assert list(mods) == want, (u, list(mods), want)
        brute = np.array([_p5_after(o, part, u, int(m)) - L0 for m in mods])
        if len(mods):
            worst = max(worst, float(np.abs(brute - diffs).max()))
            assert np.allclose(diffs, brute, rtol=0, atol=TOL), (u, diffs, brute)
        lo = brute.min() if len(mods) else 0.0
        if best_m[u] == -1:
            assert lo > -TOL, (u, lo)                  # no move although one lowers the codelength
            assert best_d[u] == 0.0
        else:
            k = int(np.flatnonzero(mods == best_m[u])[0])
            assert brute[k] <= lo + TOL, (u, brute[k], lo)   # the chosen move is the brute-force minimum
            assert abs(best_d[u] - brute[k]) <= TOL and best_d[u] < 0
    o.set_partition(part)
    return worst


@pytest.mark.parametrize("key", ["flow_net", "lfr1k_mu03", "sbm1k"])
def test_candidates_match_codelength_differences(key):
    from oracle.pyoracle import OracleNet
    o = OracleNet.load(data_path(INPUTS[key])).initiate()
    part = perturbed_partition(o, seed=3)
    rng = np.random.default_rng(1)
    units = rng.choice(o.n, min(o.n, 120), replace=False)
    worst = _check_units(o, part, units)
    print(f"[cpu] {key}: largest |dL - (P5 after - P5 before)| {worst:.2e}")


def test_empty_module_wins():
    from oracle.pyoracle import OracleNet
    n, src, dst, part = empty_wins()
    o = OracleNet.from_arcs(n, src, dst).initiate()
    _check_units(o, part, np.arange(n))
    bm, bd = o.evaluate_all()
    assert bm[20] == n + 1 and bd[20] < 0          # the only candidate is the empty module
    assert bm[21] == 10 and bd[21] < 0             # a neighbour module beats the empty one
    mods, diffs = o.candidates(21)
    assert list(mods) == [10, n + 1] and diffs[0] < diffs[1] < 0


def test_exact_tie_goes_to_the_smaller_id():
    from oracle.pyoracle import OracleNet
    n, src, dst, part = exact_tie()
    o = OracleNet.from_arcs(n, src, dst).initiate()
    p = o.flows()
    assert p[1] == p[2] and p[3] == p[4]           # mirror images, bit for bit
    o.set_partition(part)
    mods, diffs = o.candidates(0)
    assert list(mods) == [1, 2] and diffs[0] == diffs[1] and diffs[0] < 0
    bm, _ = o.evaluate_all()
    assert bm[0] == 1


def test_singleton_rule_only_removes_larger_singletons():
    from oracle.pyoracle import OracleNet
    o = OracleNet.load(data_path(INPUTS["lfr1k_mu03"])).initiate()
    o.set_partition(np.arange(o.n, dtype=np.int32))
    ref_m, ref_d = o.evaluate_all()
    rule_m, rule_d = o.evaluate_all(no_singleton_swaps=True)
    for u in range(o.n):
        mods, diffs = o.candidates(u)
        ok = (mods < u) | (mods > o.n)   # from singletons: only smaller ids (and the empty module) remain
        if ok.any() and diffs[ok].min() < 0:
            k = np.flatnonzero(ok)[np.argmin(diffs[ok])]
            assert rule_m[u] == mods[k] and rule_d[u] == diffs[k]
        else:
            assert rule_m[u] == -1
        if ref_m[u] != -1 and ref_m[u] < u:
            assert rule_m[u] == ref_m[u]
