"""Host-side checks of a steady state that hold at any size: the fixed-point
residual of the reference's iteration (Module.cpp:340-469),
||p - ((alpha + beta D) tau + beta W^T p)||_1 with W the row-normalised arc
weights and D the dangling flow, computed in FP64 numpy, chunk by chunk over
the out-CSR so that 500M-arc networks fit."""
import numpy as np

ALPHA, BETA = 0.15, 0.85


def fixed_point_residual(n, ptr, idx, w, p, node_w=None, chunk_rows=1 << 21):
    ptr = np.asarray(ptr, dtype=np.int64)
    tau = np.full(n, 1.0 / n) if node_w is None else node_w / node_w.sum()
    deg = np.diff(ptr)
    dangling = deg == 0
    D = float(p[dangling].sum())
    y = np.zeros(n)
    for lo in range(0, n, chunk_rows):
        hi = min(n, lo + chunk_rows)
        a0, a1 = int(ptr[lo]), int(ptr[hi])
        if a1 == a0:
            continue
        rows = np.repeat(np.arange(lo, hi), deg[lo:hi])
        wc = np.asarray(w[a0:a1], dtype=np.float64)
        rsum = np.bincount(rows - lo, weights=wc, minlength=hi - lo)
        contrib = p[rows] * (wc / rsum[rows - lo])
        y += np.bincount(np.asarray(idx[a0:a1]), weights=contrib, minlength=n)
    target = (ALPHA + BETA * D) * tau + BETA * y
    return float(np.abs(p - target).sum())
