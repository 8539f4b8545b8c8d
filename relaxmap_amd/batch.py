"""Many independent networks x seeds on one GPU.

The reference partitions one network per process (`ompRelaxmap`, one seed
each). Here a set of networks is partitioned with a set of seeds by a few
concurrent contexts. Each context has its own HIP stream and is driven from
its own host thread; the C calls release the GIL, so small networks from
different contexts overlap on the GPU.

Every network is uploaded and its steady state computed once per context
that takes it. Each seed then starts from singleton modules
(`rmx_set_partition` with the identity, the state `rmx_initiate` leaves).
Results do not depend on the number of workers.
"""
import queue
import threading
import time

import numpy as np

from .core import Network


def run_instances(graphs, seeds, workers=8, precision="mixed", batches=8, threshold=1e-3, vthresh=0.0,
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
