"""Many independent networks x seeds on one GPU.

The reference partitions one network per process (`ompRelaxmap`, one seed
each). Here a set of networks is partitioned with a set of seeds in one of
two ways:

* `run_batched`: every (network, seed) pair is an instance of one batch
  (`rmx_batch_upload`): the union of all instances is resident at once and
  every sweep, contraction and coarse tune covers all of them, with each
  instance's own codelength and stop decisions.
* `run_instances`: a few concurrent contexts, each with its own HIP stream
  and host thread, take the networks one after another (the C calls release
  the GIL, so small networks from different contexts overlap on the GPU).
  Each seed starts from singleton modules (`rmx_set_partition` with the
  identity, the state `rmx_initiate` leaves).

Results of either do not depend on the number of workers.
"""
import queue
import threading
import time

import numpy as np

from .core import Network


def union_of(graphs, seeds):
    """The batch of every (graph, seed) pair, graph-major: node offsets,
    per-instance seeds and the union out-CSR (node ids shifted per instance)."""
    pairs = [(gi, s) for gi in range(len(graphs)) for s in seeds]
    node_off, ptr, idx, w = graphs_union([graphs[gi] for gi, _ in pairs])
    iseeds = np.array([s for _, s in pairs], dtype=np.uint64)
    return pairs, node_off, iseeds, ptr, idx, w


def graphs_union(graphs):
    """The networks' union out-CSR (node ids shifted per network) and the
    first node of every network."""
    ns = np.array([g.n for g in graphs], dtype=np.int64)
    ms = np.array([g.m for g in graphs], dtype=np.int64)
    node_off = np.zeros(len(graphs) + 1, dtype=np.int64)
    node_off[1:] = np.cumsum(ns)
    arc_off = np.zeros(len(graphs) + 1, dtype=np.int64)
    arc_off[1:] = np.cumsum(ms)
    N, M = int(node_off[-1]), int(arc_off[-1])
    ptr = np.empty(N + 1, dtype=np.int64)
    idx = np.empty(M, dtype=np.int32)
    w = np.empty(M, dtype=np.float64)
    for k, g in enumerate(graphs):
        a, b = int(node_off[k]), int(node_off[k + 1])
        ptr[a:b] = np.asarray(g.ptr[:-1], dtype=np.int64) + arc_off[k]
        idx[arc_off[k]:arc_off[k + 1]] = np.asarray(g.idx, dtype=np.int64) + a
        w[arc_off[k]:arc_off[k + 1]] = g.w
    ptr[N] = M
    return node_off, ptr, idx, w


def run_batched(graphs, seeds, precision="mixed", batches=0, threshold=1e-3, vthresh=0.0, max_iter=10, prior=True,
                return_net=False):
    """Partition every graph with every seed as one batch of instances.

    Returns ``(results, seconds)`` as `run_instances` does (with
    ``return_net``, also the open `Network`, for inspection). The time covers
    building the union, the upload, the steady states and the partition.
    """
    t0 = time.perf_counter()
    pairs = [(gi, s) for gi in range(len(graphs)) for s in seeds]
    goff, ptr, idx, w = graphs_union(graphs)
    iseeds = np.array([s for _, s in pairs], dtype=np.uint64)
    net = Network(precision=precision, batches=batches)
    try:
        net.batch_upload_seeds(goff, len(seeds), iseeds, ptr, idx, w)
        net.initiate()
        net.run(threshold=threshold, vthresh=vthresh, max_iter=max_iter, prior=prior, seed=1)
        L, nmod, _ = net.batch_results()
    except Exception:
        net.close()
        raise
    dt = time.perf_counter() - t0
    results = {pair: (float(L[k]) / np.log(2.0), int(nmod[k]), float(L[k])) for k, pair in enumerate(pairs)}
    if return_net:
        return results, dt, net
    net.close()
    return results, dt


def run_instances(graphs, seeds, workers=8, precision="mixed", batches=0, threshold=1e-3, vthresh=0.0,
                  max_iter=10, prior=True):
    """Partition every graph (objects with ``upload_to(net)`` and ``n``) with
    every seed.

    Returns ``(results, seconds)``. ``results[(g, seed)]`` holds
    ``(codelength_bits, n_modules, computed_codelength_nats)``.
    """
    jobs = queue.Queue()
    for gi in range(len(graphs)):
        jobs.put(gi)
    results = {}
    errors = []
    lock = threading.Lock()

    def worker():
        net = None
        try:
            net = Network(precision=precision, batches=batches)
            while True:
                try:
                    gi = jobs.get_nowait()
                except queue.Empty:
                    return
                g = graphs[gi]
                g.upload_to(net)
                net.initiate()
                ident = np.arange(g.n, dtype=np.int32)
                for k, s in enumerate(seeds):
                    if k > 0:
                        net.set_partition(ident)
                    r = net.run(threshold=threshold, vthresh=vthresh, max_iter=max_iter, prior=prior, seed=s)
                    with lock:
                        results[(gi, s)] = (r.computed_codelength / np.log(2.0), r.n_module, r.computed_codelength)
        except Exception as e:  # reported to the caller
            with lock:
                errors.append(e)
        finally:
            if net is not None:
                net.close()

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(max(1, int(workers)))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    dt = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return results, dt
