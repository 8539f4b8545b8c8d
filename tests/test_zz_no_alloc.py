"""rmx_calculate_codelength_noalloc (the entry point that allocates nothing)
against rmx_calculate_codelength (Network::calculateCodeLength,
Module.cpp:2966-3002) on the same network and partition, and against the
oracle.  Tolerance: 1e-10 relative (the codelength of a fixed partition).
Two calls in a row, the second on a smaller network, so that what the first
left in the static scratch array cannot leak into the second.

The device arrays are the test's own, made with hipMalloc from the HIP
runtime the library itself links (libamdhip64.so.7), so one runtime serves
the whole process."""
import ctypes as C

import numpy as np
import pytest

from conftest import INPUTS, data_path

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    from relaxmap_amd import load_library
    load_library()   # loads libamdhip64.so.7 as the library's dependency
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipSetDevice.argtypes = [C.c_int]
    return h


class DeviceArrays:
    """Named device copies of host arrays (freed on close)."""

    def __init__(self, hip, **arrays):
        self.hip = hip
        self.ptr = {}
        assert hip.hipSetDevice(0) == 0
        for k, a in arrays.items():
            self.put(k, a)

    def put(self, k, a):
        a = np.ascontiguousarray(a)
        if k not in self.ptr:
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
            self.ptr[k] = p
        assert self.hip.hipMemcpy(self.ptr[k], a.ctypes.data_as(C.c_void_p), a.nbytes, _H2D) == 0

    def __getitem__(self, k):
        return self.ptr[k]

    def close(self):
        for p in self.ptr.values():
            self.hip.hipFree(p)
        self.ptr = {}


def _host_inputs(path, part):
    from relaxmap_amd import Network, NetworkFile
    f = NetworkFile(path)
    optr, oidx, w, nw = f.arrays()
    n, m = f.n, f.m
    with Network() as net:
        f.upload_to(net)
        net.initiate()
        p = net.flows()
        net.set_partition(part)
        ref = net.calculate_codelength()
    src = np.repeat(np.arange(n), np.diff(optr))
    outdeg = np.diff(optr)
    rows = np.where(outdeg > 0, np.bincount(src, weights=w, minlength=n), 1.0)
    flow = p[src] * (w / rows[src])
    host = dict(out_ptr=optr.astype(np.uint32), out_idx=oidx.astype(np.uint32), flow=flow, p=p,
                tau=nw / nw.sum(), dangling=(outdeg == 0).astype(np.uint8), module=np.asarray(part, np.int32))
    return n, m, host, ref, p


_ORDER = ("out_ptr", "out_idx", "flow", "p", "tau", "dangling", "module")


def _call(n, m, dev, stream=None):
    from relaxmap_amd.core import calculate_codelength_noalloc
    return calculate_codelength_noalloc(0, stream, n, m, *(dev[k] for k in _ORDER))


def _partition(n, k, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, k, n).astype(np.int32) * max(1, n // k)


def test_noalloc_matches_calculate_codelength_twice():
    from oracle.pyoracle import OracleNet
    hip = _hip()
    cases = [("lfr5k_mu05", 37), ("flow_net", 3)]   # the second input is smaller
    for key, k in cases:
        path = data_path(INPUTS[key])
        o = OracleNet.load(path).initiate()
        part = _partition(o.n, k, 11)
        n, m, host, ref, p = _host_inputs(path, part)
        dev = DeviceArrays(hip, **host)
        try:
            print(f"[gpu] noalloc {key} n={n} m={m} modules<={k}", flush=True)
            got = _call(n, m, dev)
            o.set_partition(part)
            oc = o.calculate_codelength()
            assert abs(got - ref) <= 1e-10 * abs(ref), (key, got, ref)     # vs rmx_calculate_codelength
            assert abs(got - oc) <= 1e-10 * abs(oc), (key, got, oc)        # vs the oracle
            # one module: q = 0, so L = H(p) (closed form)
            dev.put("module", np.zeros(n, np.int32))
            one = _call(n, m, dev)
            hp = -float(np.sum(np.where(p > 0, p * np.log(np.where(p > 0, p, 1.0)), 0.0)))
            assert abs(one - hp) <= 1e-10 * abs(hp), (key, one, hp)
        finally:
            dev.close()


def test_noalloc_on_a_stream_and_refusals():
    from relaxmap_amd.core import calculate_codelength_noalloc
    hip = _hip()
    path = data_path(INPUTS["lfr1k_mu03"])
    n, m, host, ref, _ = _host_inputs(path, _partition(1000, 25, 3))
    dev = DeviceArrays(hip, **host)
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    try:
        print(f"[gpu] noalloc on a stream n={n} m={m}", flush=True)
        got = _call(n, m, dev, stream=s)
        assert abs(got - ref) <= 1e-10 * abs(ref)
        with pytest.raises(ValueError, match="static scratch"):   # refused on the host
            calculate_codelength_noalloc(0, None, (1 << 20) + 1, m, *(dev[k] for k in _ORDER))
        bad = host["module"].copy()
        bad[5] = n   # one module index out of range: refused by the device-side check
        dev.put("module", bad)
        with pytest.raises(ValueError, match="module index"):
            _call(n, m, dev)
        dev.put("module", host["module"])   # the array still serves a correct call after a refusal
        assert abs(_call(n, m, dev) - ref) <= 1e-10 * abs(ref)
    finally:
        hip.hipStreamDestroy(s)
        dev.close()
