"""Input checks of rmx_graph_upload / rmx_batch_upload / rmx_batch_upload_seeds
(api.cpp check_csr): each broken rule is reported, on a small network (one host thread) and on
one with more than 2^20 arcs (row ranges checked by several threads), with
the defect in a late row."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = 5


def _net(n):
    """n nodes, node i -> i+1..i+K (mod n), heads sorted, no self-loops."""
    heads = (np.arange(n)[:, None] + np.arange(1, K + 1)[None, :]) % n
    heads.sort(axis=1)
    return np.arange(n + 1, dtype=np.int64) * K, heads.reshape(-1).astype(np.int32), np.ones(n * K)


def _broken(n, what):
    ptr, idx, w = _net(n)
    r = n - 7                      # a late row
    a = int(ptr[r])
    if what == "range":
        idx[a + K - 1] = n
    elif what == "self":
        idx[a] = r
        idx[a:a + K].sort()
    elif what == "order":
        idx[a], idx[a + 1] = idx[a + 1], idx[a]
    elif what == "weight":
        w[a + 2] = 0.0
    elif what == "nan":
        w[a + 3] = np.nan
    return ptr, idx, w


MSG = {"range": "out of range", "self": "self-loop", "order": "not sorted", "weight": "not positive",
       "nan": "not positive"}


@pytest.mark.parametrize("n", [1000, 300_000])
@pytest.mark.parametrize("what", sorted(MSG))
def test_upload_rejects(n, what):
    from relaxmap_amd import Network
    ptr, idx, w = _broken(n, what)
    with Network() as net:
        with pytest.raises(RuntimeError, match=MSG[what]):
            net.upload(n, ptr, idx, w)
        ptr, idx, w = _net(n)           # the same context still takes a valid network
        net.upload(n, ptr, idx, w)
        net.initiate()
        assert abs(net.flows().sum() - 1.0) <= 1e-12


@pytest.mark.parametrize("n", [1000, 300_000])
def test_batch_upload_rejects_cross_instance_arcs(n):
    from relaxmap_amd import Network
    ptr, idx, w = _net(n)              # the ring joins the two halves
    half = n // 2
    node_off = np.array([0, half, n], dtype=np.int32)
    with Network() as net:
        with pytest.raises(RuntimeError, match="joins two instances"):
            net.batch_upload(node_off, [1, 2], ptr, idx, w)


def test_batch_upload_seeds_rejects():
    """rmx_batch_upload_seeds: an arc between two networks, offsets that do
    not span the nodes, and a seed count that does not match are refused; the
    context then still takes a valid batch, whose copies of one network have
    identical flows summing to 1 each."""
    from relaxmap_amd import Network
    n = 1000
    ptr, idx, w = _net(n)              # the ring joins the two halves
    with Network() as net:
        with pytest.raises(RuntimeError, match="joins two networks"):
            net.batch_upload_seeds([0, n // 2, n], 2, [1, 2, 3, 4], ptr, idx, w)
        with pytest.raises(RuntimeError, match="does not span"):
            net.batch_upload_seeds([0, n - 1], 2, [1, 2], ptr, idx, w)
        with pytest.raises(ValueError, match="one seed per"):
            net.batch_upload_seeds([0, n], 2, [1], ptr, idx, w)
        net.batch_upload_seeds([0, n], 2, [1, 2], ptr, idx, w)
        net.initiate()
        p = net.flows()
        assert np.array_equal(p[:n], p[n:])
        assert abs(p[:n].sum() - 1.0) <= 1e-12


def test_precision_change_needs_initiate():
    """rmx_initiate makes the arc weights of the precision set at the time;
    after a change of precision the sweeps refuse to run until it is called
    again, and the re-initiated run is that of a context created with the
    new precision."""
    from relaxmap_amd import Network, synth
    from relaxmap_amd.core import RMX_MIXED
    g = synth.sbm(2000, 20, 8.0, 0.2, 5)
    with Network(precision="fp64") as net:
        g.upload_to(net)
        net.initiate()
        net._check(net._L.rmx_set_precision(net._h, RMX_MIXED))
        with pytest.raises(RuntimeError, match="rmx_initiate first"):
            net.run(seed=2)
        net.initiate()
        net.run(seed=2)
        part, L = net.partition(), net.codelength
    with Network(precision="mixed") as ref:
        g.upload_to(ref)
        ref.initiate()
        ref.run(seed=2)
        assert np.array_equal(part, ref.partition()) and L == ref.codelength
