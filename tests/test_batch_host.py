"""The host side of a batch: the union of many networks x seeds."""
import numpy as np


def test_union_of_shifts_every_instance():
    from relaxmap_amd import synth
    from relaxmap_amd.batch import union_of
    graphs = [synth.sbm(200, 5, 6.0, 0.2, 1), synth.sbm(350, 7, 5.0, 0.3, 2)]
    seeds = [4, 9, 11]
    pairs, off, iseeds, ptr, idx, w = union_of(graphs, seeds)
    assert pairs == [(g, s) for g in range(2) for s in seeds]
    assert list(iseeds) == seeds * 2
    assert off[-1] == len(ptr) - 1 and ptr[-1] == len(idx) == len(w)
    for k, (gi, _) in enumerate(pairs):
        g = graphs[gi]
        a, b = int(off[k]), int(off[k + 1])
        assert b - a == g.n
        assert np.array_equal(ptr[a:b + 1] - ptr[a], np.asarray(g.ptr))
        arcs = slice(int(ptr[a]), int(ptr[b]))
        assert np.array_equal(idx[arcs] - a, np.asarray(g.idx))      # heads inside the instance
        assert np.array_equal(w[arcs], np.asarray(g.w))
