"""Degree-bucket cut-offs and the autotuner.

A unit's best move does not depend on which kernel evaluates it: every
kernel sums the same exact fixed-point flows and breaks ties the same way.
So any admissible cut-offs, including sending every unit with an arc
through the hub partitions, must reproduce the default run bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEFAULT = [16, 32, 128, 256, 512, 1023, 2048]


def _graph():
    from relaxmap_amd import synth
    # heavy-tailed degrees: units in every bucket, hubs included
    return synth.chunglu(40_000, 2.1, 4, 6000, 16.0, 1.2, 50, 4000, 0.4, 0.01, 11)


def _run(g, cuts=None, tune=False, precision="mixed"):
    from relaxmap_amd import Network
    with Network(precision=precision) as net:
        g.upload_to(net)
        net.initiate()
        if cuts is not None:
            net.set_bucket_cutoffs(cuts)
        tuned = net.autotune() if tune else None
        r = net.run(seed=3)
        return net.partition(), r.computed_codelength, net.bucket_cutoffs(), tuned


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cutoffs_do_not_change_results(precision):
    g = _graph()
    part0, L0, cuts0, _ = _run(g, precision=precision)
    assert list(cuts0) == DEFAULT
    for cuts in ([8, 16, 64, 128, 256, 512, 1024], [0, 32, 32, 256, 256, 1023, 1023], [0] * 7):
        part, L, got, _ = _run(g, cuts=cuts, precision=precision)
        print(f"[gpu] {precision} cuts {cuts}: L {L:.12f} vs {L0:.12f}", flush=True)
        assert list(got) == cuts
        assert L == L0
        assert np.array_equal(part, part0)


def test_autotune_installs_admissible_cutoffs():
    g = _graph()
    part0, L0, _, _ = _run(g)
    part, L, got, tuned = _run(g, tune=True)
    print(f"[gpu] tuned cut-offs {list(tuned)}", flush=True)
    assert list(got) == list(tuned)
    assert all(0 <= t <= d for t, d in zip(tuned, DEFAULT))
    assert all(a <= b for a, b in zip(tuned[:-1], tuned[1:]))
    assert L == L0 and np.array_equal(part, part0)


def test_cutoffs_rejected_when_out_of_range():
    from relaxmap_amd import Network
    with Network() as net:
        with pytest.raises(RuntimeError):
            net.set_bucket_cutoffs([17, 32, 128, 256, 512, 1023, 2048])   # above the table limit
        with pytest.raises(RuntimeError):
            net.set_bucket_cutoffs([16, 8, 128, 256, 512, 1023, 2048])    # not ascending
