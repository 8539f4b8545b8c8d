"""The blocked push / pull steady state, mixed precision (FP32 products) and
FP64 (FP64 products in two-word fixed point), on shapes that reach every
branch of its kernels, against the oracle's FP64 flows, and the FP64
blocked path against the FP64 in-CSR pull.

- long runs of rows without out-arcs inside one push tile (row bounds read
  from global memory instead of LDS);
- fewer than 8 arcs in a tile (the scalar tail);
- arc counts that are whole multiples of the tile;
- many bins, and a hub that every node points to (one LDS accumulator
  takes a whole bin's arcs);
- a context that uploads another network, one without arcs, and the first
  again: each gets its own layout.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _csr(n, src, dst, w):
    src, dst, w = np.asarray(src, np.int64), np.asarray(dst, np.int64), np.asarray(w, np.float64)
    keep = src != dst
    src, dst, w = src[keep], dst[keep], w[keep]
    key = src * n + dst
    key, first = np.unique(key, return_index=True)
    src, dst, w = key // n, key % n, w[first]
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, src + 1, 1)
    return np.cumsum(ptr), dst, w, src


def _random_arcs(rng, nodes, n, deg_max):
    src, dst = [], []
    for i in nodes:
        d = int(rng.integers(1, deg_max + 1))
        src.append(np.full(d, i))
        dst.append(rng.integers(0, n, d))
    return np.concatenate(src), np.concatenate(dst)


def _graphs():
    rng = np.random.default_rng(17)
    out = {}
    n = 16000   # rows 100..13999 have no out-arcs: one tile spans ~14000 rows
    s, d = _random_arcs(rng, [i for i in range(n) if not 100 <= i < 14000], n, 20)
    out["dangling_run"] = (n, s, d, rng.lognormal(0, 1, len(s)))
    out["cycle3"] = (3, [0, 1, 2], [1, 2, 0], [1.0, 2.0, 3.0])
    n = 512     # 8 distinct targets each: 4096 arcs, two full tiles
    s = np.repeat(np.arange(n), 8)
    d = (s + 1 + np.tile(np.arange(8) * 37, n)) % n
    out["two_tiles"] = (n, s, d, rng.random(len(s)) + 0.1)
    n = 70000   # many bins; node 0 is a hub in both directions
    s, d = _random_arcs(rng, range(1, n), n, 6)
    s = np.concatenate([s, np.arange(1, n), np.zeros(n - 1, np.int64)])
    d = np.concatenate([d, np.zeros(n - 1, np.int64), np.arange(1, n)])
    out["hub_bins"] = (n, s, d, rng.lognormal(0, 0.5, len(s)))
    return out


@pytest.mark.parametrize("name", ["dangling_run", "cycle3", "two_tiles", "hub_bins"])
def test_mixed_flows_match_oracle(name):
    from oracle.pyoracle import OracleNet
    from relaxmap_amd import Network
    n, s, d, w = _graphs()[name]
    ptr, idx, ww, src = _csr(n, s, d, w)
    print(f"[gpu] push/pull {name}: n={n} m={len(idx)}", flush=True)
    ref = OracleNet.from_arcs(n, src, idx, ww).initiate().flows()
    for prec, blocked, tol in (("mixed", True, 1e-6), ("fp64", True, 1e-12), ("fp64", False, 1e-12)):
        with Network(precision=prec) as net:
            net.set_fp64_flow(blocked)
            net.upload(n, ptr, idx, ww)
            net.initiate()
            p = net.flows()
        err = np.abs(p - ref).sum()
        print(f"[gpu]   {prec} {'blocked' if blocked else 'pull'}: L1 {err:.3e}", flush=True)
        assert abs(p.sum() - 1.0) <= 1e-12
        assert err <= tol, (prec, err)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_blocked_flows_deterministic(prec):
    from relaxmap_amd import Network
    n, s, d, w = _graphs()["hub_bins"]
    ptr, idx, ww, _ = _csr(n, s, d, w)
    runs = []
    for _ in range(2):
        with Network(precision=prec) as net:
            net.upload(n, ptr, idx, ww)
            net.initiate()
            runs.append(net.flows())
    assert np.array_equal(runs[0], runs[1])


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_layout_follows_each_upload(prec):
    """A context keeps its push / pull layout from one steady state to the
    next, until the next upload.  A network uploaded after another, in the
    same context, gets its own layout: one without arcs (every node dangling,
    so p is the uniform teleport distribution) and then the first one again
    (the same flows as in a fresh context)."""
    from relaxmap_amd import Network
    n, s, d, w = _graphs()["hub_bins"]
    ptr, idx, ww, _ = _csr(n, s, d, w)
    with Network(precision=prec) as fresh:
        fresh.upload(n, ptr, idx, ww)
        fresh.initiate()
        p_ref = fresh.flows()
    with Network(precision=prec) as net:
        net.upload(n, ptr, idx, ww)
        net.initiate()
        n2 = 5000
        net.upload(n2, np.zeros(n2 + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))
        net.initiate()
        p0 = net.flows()
        net.upload(n, ptr, idx, ww)
        net.initiate()
        p1 = net.flows()
    print(f"[gpu] re-upload {prec}: arc-less L1 {np.abs(p0 - 1.0 / n2).sum():.2e}", flush=True)
    assert np.abs(p0 - 1.0 / n2).sum() <= 1e-12
    assert np.array_equal(p1, p_ref)


@pytest.mark.parametrize("key", ["lfr1k_mu03", "lfr5k_mu05", "sbm1k", "flow_net"])
def test_fp64_blocked_matches_fp64_pull(key):
    """Both FP64 iterations take the reference's stop rule and FP64 sums:
    same iteration count (within one), flows within 1e-13 L1 of each other."""
    from conftest import INPUTS, data_path
    from relaxmap_amd import Network
    out = []
    for blocked in (True, False):
        with Network() as net:
            net.set_fp64_flow(blocked)
            net.load(data_path(INPUTS[key]))
            it, _ = net.initiate()
            out.append((it, net.flows()))
    err = float(np.abs(out[0][1] - out[1][1]).sum())
    print(f"[gpu] {key}: blocked {out[0][0]} it, pull {out[1][0]} it, L1 {err:.2e}", flush=True)
    assert abs(out[0][0] - out[1][0]) <= 1
    assert err <= 1e-13
