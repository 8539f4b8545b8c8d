"""ctypes view of liboracle.so, the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py as the checker.  The relaxmap_amd product
never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_F64P = C.POINTER(C.c_double)


def _p(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        vp = C.c_void_p
        sig = {
            "orc_load": (vp, [C.c_char_p, C.c_char_p, C.c_int]),
            "orc_from_arcs": (vp, [C.c_int, C.c_int64, _I32P, _I32P, _F64P, _F64P]),
            "orc_free": (None, [vp]),
            "orc_n": (C.c_int, [vp]),
            "orc_m": (C.c_int64, [vp]),
            "orc_n_selflinks": (C.c_int, [vp]),
            "orc_n_double_links": (C.c_int, [vp]),
            "orc_name": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int]),
            "orc_initiate": (C.c_int, [vp]),
            "orc_iterations": (C.c_int, [vp]),
            "orc_n_dangling": (C.c_int, [vp]),
            "orc_get_flows": (None, [vp, _F64P]),
            "orc_get_node_exit": (None, [vp, _F64P]),
            "orc_get_teleport": (None, [vp, _F64P]),
            "orc_get_out_csr": (None, [vp, _I64P, _I32P, _F64P]),
            "orc_get_raw_out_w": (None, [vp, _F64P]),
            "orc_codelength": (C.c_double, [vp]),
            "orc_sum_exit": (C.c_double, [vp]),
            "orc_all_log_all": (C.c_double, [vp]),
            "orc_n_module": (C.c_int, [vp]),
            "orc_n_empty": (C.c_int, [vp]),
            "orc_set_partition": (C.c_int, [vp, _I32P]),
            "orc_get_partition": (None, [vp, _I32P]),
            "orc_get_module_stats": (None, [vp, _F64P, _F64P, _F64P, _F64P, _I32P]),
            "orc_calculate_codelength": (C.c_double, [vp]),
            "orc_update_members": (None, [vp]),
            "orc_contract": (C.c_int, [vp]),
            "orc_n_super": (C.c_int, [vp]),
            "orc_super_m": (C.c_int64, [vp]),
            "orc_get_super": (None, [vp, _F64P, _F64P, _F64P, _F64P, _I32P, _I32P]),
            "orc_get_super_out": (None, [vp, _I64P, _I32P, _F64P]),
            "orc_get_super_in": (None, [vp, _I64P, _I32P, _F64P]),
            "orc_seed": (None, [vp, C.c_uint32]),
            "orc_rand_int": (C.c_uint32, [vp, C.c_uint32]),
            "orc_sweep": (C.c_int, [vp, C.c_int, C.c_int, C.c_double]),
            "orc_reset_active": (None, [vp, C.c_int]),
            "orc_n_active": (C.c_int, [vp]),
            "orc_run": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "orc_subnetwork": (vp, [vp, C.c_int, _I32P]),
            "orc_evaluate_all": (None, [vp, C.c_int, C.c_int, _I32P, _F64P]),
            "orc_candidates": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _I32P, _F64P]),
            "orc_trace_len": (C.c_int, [vp]),
            "orc_get_trace": (None, [vp, _I32P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


VARIANTS = {"move": 0, "prior": 1, "pmove": 2, "ppmove": 3}


class OracleNet:
    """One reference ``Network`` as restated by the oracle."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle could not build the network")
        self._h = handle
        self._L = lib()

    @classmethod
    def load(cls, path):
        err = C.create_string_buffer(256)
        h = lib().orc_load(str(path).encode(), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    @classmethod
    def from_arcs(cls, n, src, dst, w=None, node_w=None):
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.float64)
            wp = _p(w, _F64P)
        nwp = None
        if node_w is not None:
            node_w = np.ascontiguousarray(node_w, dtype=np.float64)
            nwp = _p(node_w, _F64P)
        return cls(lib().orc_from_arcs(int(n), len(src), _p(src, _I32P), _p(dst, _I32P), wp, nwp))

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_free(self._h)
            self._h = None

    # --- scalars -------------------------------------------------------
    @property
    def n(self):
        return self._L.orc_n(self._h)

    @property
    def m(self):
        return self._L.orc_m(self._h)

    def initiate(self):
        self._L.orc_initiate(self._h)
        return self

    iterations = property(lambda s: s._L.orc_iterations(s._h))
    n_dangling = property(lambda s: s._L.orc_n_dangling(s._h))
    codelength = property(lambda s: s._L.orc_codelength(s._h))
    sum_exit = property(lambda s: s._L.orc_sum_exit(s._h))
    all_log_all = property(lambda s: s._L.orc_all_log_all(s._h))
    n_module = property(lambda s: s._L.orc_n_module(s._h))
    n_empty = property(lambda s: s._L.orc_n_empty(s._h))
    n_selflinks = property(lambda s: s._L.orc_n_selflinks(s._h))
    n_double_links = property(lambda s: s._L.orc_n_double_links(s._h))
    n_active = property(lambda s: s._L.orc_n_active(s._h))

    def name(self, i):
        b = C.create_string_buffer(512)
        self._L.orc_name(self._h, i, b, 512)
        return b.value.decode()

    # --- arrays --------------------------------------------------------
    def _darr(self, fn, k=None):
        a = np.empty(self.n if k is None else k, dtype=np.float64)
        fn(self._h, _p(a, _F64P))
        return a

    def flows(self):
        return self._darr(self._L.orc_get_flows)

    def node_exit(self):
        return self._darr(self._L.orc_get_node_exit)

    def teleport(self):
        return self._darr(self._L.orc_get_teleport)

    def out_csr(self):
        ptr = np.empty(self.n + 1, dtype=np.int64)
        idx = np.empty(self.m, dtype=np.int32)
        w = np.empty(self.m, dtype=np.float64)
        self._L.orc_get_out_csr(self._h, _p(ptr, _I64P), _p(idx, _I32P), _p(w, _F64P))
        return ptr, idx, w

    def raw_out_w(self):
        return self._darr(self._L.orc_get_raw_out_w, self.m)

    def set_partition(self, part):
        part = np.ascontiguousarray(part, dtype=np.int32)
        if self._L.orc_set_partition(self._h, _p(part, _I32P)) != 0:
            raise ValueError("bad partition")

    def partition(self):
        a = np.empty(self.n, dtype=np.int32)
        self._L.orc_get_partition(self._h, _p(a, _I32P))
        return a

    def module_stats(self):
        n = self.n
        ex, sp, tp, dg = (np.empty(n) for _ in range(4))
        nm = np.empty(n, dtype=np.int32)
        self._L.orc_get_module_stats(self._h, _p(ex, _F64P), _p(sp, _F64P), _p(tp, _F64P),
                                     _p(dg, _F64P), _p(nm, _I32P))
        return {"exit": ex, "sum_pr": sp, "sum_tp": tp, "sum_dangling": dg, "n_members": nm}

    def calculate_codelength(self):
        return self._L.orc_calculate_codelength(self._h)

    def update_members(self):
        self._L.orc_update_members(self._h)

    def contract(self):
        return self._L.orc_contract(self._h)

    def supernodes(self):
        s = self._L.orc_n_super(self._h)
        m = self._L.orc_super_m(self._h)
        size, ex, tp, dg = (np.empty(s) for _ in range(4))
        mod = np.empty(s, dtype=np.int32)
        nm = np.empty(s, dtype=np.int32)
        self._L.orc_get_super(self._h, _p(size, _F64P), _p(ex, _F64P), _p(tp, _F64P), _p(dg, _F64P),
                              _p(mod, _I32P), _p(nm, _I32P))
        optr = np.empty(s + 1, dtype=np.int64)
        oidx = np.empty(m, dtype=np.int32)
        ow = np.empty(m)
        self._L.orc_get_super_out(self._h, _p(optr, _I64P), _p(oidx, _I32P), _p(ow, _F64P))
        iptr = np.empty(s + 1, dtype=np.int64)
        iidx = np.empty(m, dtype=np.int32)
        iw = np.empty(m)
        self._L.orc_get_super_in(self._h, _p(iptr, _I64P), _p(iidx, _I32P), _p(iw, _F64P))
        return {"size": size, "exit": ex, "tp": tp, "dangling": dg, "mod": mod, "n_members": nm,
                "out_ptr": optr, "out_idx": oidx, "out_w": ow,
                "in_ptr": iptr, "in_idx": iidx, "in_w": iw}

    def seed(self, s):
        self._L.orc_seed(self._h, int(s) & 0xFFFFFFFF)

    def rand_int(self, n):
        return self._L.orc_rand_int(self._h, int(n))

    def reset_active(self, sp=False):
        self._L.orc_reset_active(self._h, 1 if sp else 0)

    def sweep(self, variant="move", sp=False, vthresh=0.0):
        return self._L.orc_sweep(self._h, VARIANTS[variant], 1 if sp else 0, float(vthresh))

    def subnetwork(self, members):
        """generate_network_from_module: the network of the given nodes (ascending
        ids) with this network's flows, in singleton modules."""
        members = np.ascontiguousarray(members, dtype=np.int32)
        h = self._L.orc_subnetwork(self._h, len(members), _p(members, _I32P))
        if not h:
            raise ValueError("bad member list")
        sub = OracleNet(h)
        sub._parent = self   # names and flows were copied; the parent may go
        return sub

    def partition_module_network(self, threshold=1e-3, max_iter=10, seed=1):
        """partition_module_network(fast) at one thread (OmpRelaxmap.cpp:436-460):
        move() sweeps until one gains less than `threshold` bits, at most
        max_iter; returns the partition."""
        import math
        self.seed(seed)
        for _ in range(max_iter):
            old = self.codelength
            self.sweep("move")
            if (old - self.codelength) / math.log(2.0) < threshold:
                break
        return self.partition()

    def evaluate_all(self, sp=False, no_singleton_swaps=False):
        """Every unit's best move against one module snapshot (nothing moves):
        (best module, -1 for none and n + 1 for the empty module; its dL)."""
        k = self._L.orc_n_super(self._h) if sp else self.n
        bm = np.empty(k, dtype=np.int32)
        bd = np.empty(k, dtype=np.float64)
        self._L.orc_evaluate_all(self._h, 1 if sp else 0, 1 if no_singleton_swaps else 0, _p(bm, _I32P),
                                 _p(bd, _F64P))
        return bm, bd

    def candidates(self, u, sp=False, cap=100000):
        """Unit u's candidate modules and their dL, in the order evaluated."""
        mods = np.empty(cap, dtype=np.int32)
        diffs = np.empty(cap, dtype=np.float64)
        k = self._L.orc_candidates(self._h, 1 if sp else 0, int(u), cap, _p(mods, _I32P), _p(diffs, _F64P))
        return mods[:k].copy(), diffs[:k].copy()

    def run(self, threshold=1e-3, vthresh=0.0, max_iter=10, prior=True):
        """main()'s super-step loop at one thread, until the first coarse tune."""
        coarse = C.c_int(0)
        steps = self._L.orc_run(self._h, threshold, -vthresh, max_iter, 1 if prior else 0, C.byref(coarse))
        k = self._L.orc_trace_len(self._h)
        tr = np.empty(k, dtype=np.int32)
        self._L.orc_get_trace(self._h, _p(tr, _I32P))
        return {"steps": steps, "coarse_next": bool(coarse.value), "moved": tr.tolist()}
