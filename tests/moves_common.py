"""Shared inputs of the per-unit move tests: partly converged partitions with
misplaced nodes, and small networks where the empty module wins or two
candidates tie exactly."""
import numpy as np


def perturbed_partition(o, seed, frac=0.15, sweeps=2):
    """`sweeps` oracle prioritize_parallelMove sweeps from singletons, then a
    fraction of the nodes moved to modules drawn from the partition (so some
    sit in modules they have no arc to)."""
    o.set_partition(np.arange(o.n, dtype=np.int32))
    o.seed(seed)
    o.reset_active()
    for _ in range(sweeps):
        o.sweep("ppmove")
    part = o.partition()
    rng = np.random.default_rng(seed)
    idx = rng.choice(o.n, max(1, int(frac * o.n)), replace=False)
    part[idx] = part[rng.integers(0, o.n, len(idx))]
    return part.astype(np.int32)


def _arcs(pairs):
    a = np.array(sorted(set(pairs)), dtype=np.int64)
    return a[:, 0].astype(np.int32), a[:, 1].astype(np.int32)


def empty_wins():
    """Two 10-cliques, node 20 without arcs and node 21 linked only to the
    second clique, both placed in the first clique's module: node 20's best
    move is the empty module (its only candidate), node 21's the second
    clique's module."""
    pairs = [(i, j) for i in range(10) for j in range(10) if i != j]
    pairs += [(i, j) for i in range(10, 20) for j in range(10, 20) if i != j]
    pairs += [(21, 10), (11, 21)]
    src, dst = _arcs(pairs)
    part = np.zeros(22, np.int32)
    part[10:20] = 10
    return 22, src, dst, part


def exact_tie():
    """Node 0 linked both ways to node 1 and node 2; modules {1, 3} and
    {2, 4} are mirror images (1<->3 and 2<->4 both ways), so moving node 0
    into either gives the same dL bit for bit: the tie goes to module 1."""
    pairs = [(0, 1), (1, 0), (0, 2), (2, 0), (1, 3), (3, 1), (2, 4), (4, 2)]
    src, dst = _arcs(pairs)
    part = np.array([0, 1, 2, 1, 2], np.int32)
    return 5, src, dst, part


def hub_network(n=6000, hub_deg=2600, seed=5):
    """Planted blocks of 60 nodes with about 8 out-arcs per node, plus node 0
    with `hub_deg` out- and in-arcs to random nodes (a unit above 2048 arcs,
    evaluated by the hub partitions on the GPU)."""
    rng = np.random.default_rng(seed)
    blk = np.arange(n) // 60
    src = np.repeat(np.arange(n), 8)
    inside = rng.random(len(src)) < 0.8
    dst = np.where(inside, blk[src] * 60 + rng.integers(0, 60, len(src)), rng.integers(0, n, len(src)))
    dst = np.minimum(dst, n - 1)
    h = rng.choice(np.arange(1, n), hub_deg, replace=False)
    src = np.concatenate([src, np.zeros(hub_deg, np.int64), h])
    dst = np.concatenate([dst, h, np.zeros(hub_deg, np.int64)])
    keep = src != dst
    pairs = np.unique(np.stack([src[keep], dst[keep]], 1), axis=0)
    return n, pairs[:, 0].astype(np.int32), pairs[:, 1].astype(np.int32)


def csr(n, src, dst):
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, src.astype(np.int64) + 1, 1)
    return np.cumsum(ptr), dst.astype(np.int32)
