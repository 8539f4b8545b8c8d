"""Many networks x seeds (relaxmap_amd.batch) and the CLI's attempts.

Concurrent contexts must give exactly the results of one context, and a run
with several attempts must keep the best of the same runs made one by one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs():
    from relaxmap_amd import synth
    return [synth.sbm(n, b, 12.0, mu, 40 + k)
            for k, (n, b, mu) in enumerate([(600, 12, 0.2), (1500, 30, 0.35), (3000, 40, 0.3),
                                            (800, 10, 0.45), (2500, 60, 0.15), (5000, 100, 0.25)])]


def test_instances_independent_of_workers():
    from relaxmap_amd.batch import run_instances
    graphs = _graphs()
    seeds = [1, 2, 3]
    one, t1 = run_instances(graphs, seeds, workers=1)
    many, t3 = run_instances(graphs, seeds, workers=3)
    print(f"[gpu] {len(graphs) * len(seeds)} instances: 1 worker {t1:.2f} s, 3 workers {t3:.2f} s", flush=True)
    assert set(one) == set(many) == {(g, s) for g in range(len(graphs)) for s in seeds}
    for k in one:
        assert one[k] == many[k], k


def test_instances_match_single_runs():
    """A seed run after set_partition(identity) equals a run after initiate."""
    from relaxmap_amd import Network
    from relaxmap_amd.batch import run_instances
    g = _graphs()[2]
    res, _ = run_instances([g], [5, 9], workers=1)
    for s in (5, 9):
        with Network(precision="mixed") as net:
            g.upload_to(net)
            net.initiate()
            r = net.run(seed=s)
            assert res[(0, s)][2] == r.computed_codelength


def test_attempts_keep_the_best():
    from relaxmap_amd import Network
    g = _graphs()[3]
    single = []
    for s in (7, 8, 9):
        with Network() as net:
            g.upload_to(net)
            net.initiate()
            single.append(net.run(seed=s).computed_codelength)
    with Network() as net:
        g.upload_to(net)
        net.initiate()
        r = net.run(seed=7, attempts=3)
        best = int(np.argmin(single))
        assert r.best_attempt == best
        assert r.computed_codelength == single[best]
        assert abs(net.codelength - r.computed_codelength) <= 1e-9 * abs(r.computed_codelength)
