"""The host-side fixed-point residual used at full sizes (tests/flowcheck.py)
checked against the oracle's converged flows (small residual) and against
flows perturbed by a known amount (a residual of that order)."""
import numpy as np

from conftest import INPUTS, data_path
from flowcheck import fixed_point_residual


def test_residual_of_converged_and_perturbed_flows():
    from oracle.pyoracle import OracleNet
    from relaxmap_amd import NetworkFile
    path = data_path(INPUTS["lfr5k_mu05"])
    o = OracleNet.load(path).initiate()
    ptr, idx, w, nw = NetworkFile(path).arrays()
    p = o.flows()
    assert fixed_point_residual(o.n, ptr, idx, w, p, nw, chunk_rows=777) <= 1e-14
    rng = np.random.default_rng(0)
    q = p * (1 + 1e-6 * rng.standard_normal(o.n))
    q /= q.sum()
    r = fixed_point_residual(o.n, ptr, idx, w, q, nw)
    assert 1e-8 < r < 1e-5
