"""The two large workloads at their full sizes, through properties that hold
at any size (the oracle's loops are too slow there), plus the R-MAT workload
at a reduced scale against the oracle itself.

- R-MAT scale 22 (4,194,304 ids, edge factor 16, a, b, c, d = .57, .19, .19,
  .05, FP32 U(0,1] weights, dangling nodes kept): FP64 flows satisfy the
  reference's fixed point to 1e-12 in L1 (residual computed on the host) and
  sum to 1 within 1e-12; mixed-precision flows are within 1e-6 L1 of FP64
  (the precision pattern's budget); the codelength of one fixed partition
  (FP64's converged one) differs between the precisions by at most 1e-6
  relative.  At scale 16 the FP64 flows match the oracle to 1e-12 L1.
- Chung-Lu, 32M nodes / 529M arcs (degree exponent 2.1 in [4, 100000],
  communities exponent 1.2 in [50, 50000], mu 0.4, FP32 lognormal weights,
  1 % dangling): mixed vs FP64 flows within 1e-6 L1, final codelengths
  within 0.5 % of each other, tracked codelength = recomputed codelength to
  1e-9 relative in both, sum p = 1 within 1e-12 and the FP64 residual within
  1e-12 L1.
"""
import math

import numpy as np
import pytest

from flowcheck import fixed_point_residual

pytestmark = pytest.mark.gpu


def _network(g, precision):
    from relaxmap_amd import Network
    net = Network(precision=precision)
    g.upload_to(net)
    net.initiate()
    return net


def test_rmat_scale16_flows_match_oracle():
    from oracle.pyoracle import OracleNet
    from relaxmap_amd import synth
    g = synth.config3(16)
    src = np.repeat(np.arange(g.n), np.diff(g.ptr)).astype(np.int32)
    o = OracleNet.from_arcs(g.n, src, np.asarray(g.idx, np.int32), np.asarray(g.w)).initiate()
    with _network(g, "fp64") as net:
        print(f"[gpu] R-MAT scale 16: {g.n} nodes, {g.m} arcs", flush=True)
        err = float(np.abs(net.flows() - o.flows()).sum())
        print(f"[gpu] R-MAT scale 16: FP64 flow L1 vs oracle {err:.2e}", flush=True)
        assert err <= 1e-12
        assert abs(net.iterations - o.iterations) <= 1


@pytest.fixture(scope="module")
def rmat22():
    from relaxmap_amd import synth
    return synth.config3()


def test_rmat_scale22_precision(rmat22):
    g = rmat22
    print(f"[gpu] R-MAT scale 22: {g.n} nodes, {g.m} arcs", flush=True)
    with _network(g, "fp64") as nd:
        pd = nd.flows()
        r = nd.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
        part = nd.partition()
        nd.set_partition(part)
        Ld = nd.calculate_codelength()
    res = fixed_point_residual(g.n, g.ptr, g.idx, g.w, pd)
    print(f"[gpu] R-MAT 22 FP64: residual {res:.2e}, sum p - 1 = {pd.sum() - 1:.1e}, "
          f"L = {r.computed_codelength / math.log(2):.5f} bits", flush=True)
    assert res <= 1e-12
    assert abs(pd.sum() - 1.0) <= 1e-12
    with _network(g, "mixed") as nm:
        pm = nm.flows()
        nm.set_partition(part)
        Lm = nm.calculate_codelength()
    l1 = float(np.abs(pm - pd).sum())
    rel = (Lm - Ld) / Ld
    print(f"[gpu] R-MAT 22 mixed vs FP64: flow L1 {l1:.2e}, codelength of the fixed partition rel {rel:.2e}",
          flush=True)
    assert l1 <= 1e-6
    assert abs(rel) <= 1e-6


@pytest.fixture(scope="module")
def chunglu32m():
    from relaxmap_amd import synth
    return synth.config5(mean_deg=18.5)


def test_chunglu_32m_precision(chunglu32m):
    g = chunglu32m
    print(f"[gpu] Chung-Lu: {g.n} nodes, {g.m} arcs", flush=True)
    out = {}
    for prec in ("mixed", "fp64"):
        with _network(g, prec) as net:
            p = net.flows()
            r = net.run(threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1)
            out[prec] = (p, r)
            print(f"[gpu] Chung-Lu {prec}: {net.iterations} iterations, L tracked {r.codelength:.10f} "
                  f"computed {r.computed_codelength:.10f} nats, {r.n_module} modules", flush=True)
            assert abs(r.codelength - r.computed_codelength) <= 1e-9 * abs(r.computed_codelength)
            assert abs(p.sum() - 1.0) <= 1e-12
    pm, rm = out["mixed"]
    pd, rd = out["fp64"]
    l1 = float(np.abs(pm - pd).sum())
    rel = (rm.computed_codelength - rd.computed_codelength) / rd.computed_codelength
    print(f"[gpu] Chung-Lu mixed vs FP64: flow L1 {l1:.2e}, final codelength rel {rel:.2e}", flush=True)
    assert l1 <= 1e-6
    assert abs(rel) <= 5e-3
    res = fixed_point_residual(g.n, g.ptr, g.idx, g.w, pd)
    print(f"[gpu] Chung-Lu FP64 residual {res:.2e}", flush=True)
    assert res <= 1e-12
