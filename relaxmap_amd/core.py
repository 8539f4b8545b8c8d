"""ctypes binding of librelaxmap.so (include/relaxmap.h).

``Network`` mirrors the reference's C++ ``Network`` (uwescience/RelaxMap
Module.h:99-200): ``initiate``, ``calculate_codelength``, ``move`` sweeps,
``update_members_in_module``, ``convert_modules_to_super_nodes`` and the
CLI's ``partition`` loop.  Every call goes to the GPU through the C ABI; a
missing library or GPU raises instead of falling back to anything else.
"""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_F64P = C.POINTER(C.c_double)

RMX_FP64 = 0
RMX_MIXED = 1
RMX_EIO = -5


def library_path():
    return os.path.join(_HERE, "librelaxmap.so")


class _Opts(C.Structure):
    _fields_ = [("threshold", C.c_double), ("vthresh", C.c_double), ("max_iter", C.c_int32),
                ("prior", C.c_int32), ("seed", C.c_uint64), ("verbose", C.c_int32), ("attempts", C.c_int32)]


class _Result(C.Structure):
    _fields_ = [("codelength", C.c_double), ("computed_codelength", C.c_double), ("n_module", C.c_int32),
                ("super_steps", C.c_int32), ("sweeps", C.c_int64), ("edge_visits", C.c_int64),
                ("time_partition_s", C.c_double), ("time_sweeps_s", C.c_double),
                ("time_contract_s", C.c_double), ("time_coarse_s", C.c_double), ("best_attempt", C.c_int32)]


@dataclass
class PartitionResult:
    codelength: float
    computed_codelength: float
    n_module: int
    super_steps: int
    sweeps: int
    edge_visits: int
    time_partition_s: float
    time_sweeps_s: float
    time_contract_s: float
    time_coarse_s: float
    best_attempt: int = 0


def load_library():
    """Loads librelaxmap.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() or make -C relaxmap_amd/csrc")
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "rmx_device_count": (C.c_int, []),
        "rmx_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "rmx_destroy": (None, [vp]),
        "rmx_last_error": (C.c_char_p, [vp]),
        "rmx_set_precision": (C.c_int, [vp, C.c_int]),
        "rmx_set_batches": (C.c_int, [vp, C.c_int]),
        "rmx_set_hub_path": (C.c_int, [vp, C.c_int]),
        "rmx_set_fp64_flow": (C.c_int, [vp, C.c_int]),
        "rmx_set_graph_capture": (C.c_int, [vp, C.c_int]),
        "rmx_set_candidate_bound": (C.c_int, [vp, C.c_int, C.c_int]),
        "rmx_set_attempt_mode": (C.c_int, [vp, C.c_int]),
        "rmx_net_read": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp), C.c_char_p, C.c_int]),
        "rmx_net_free": (None, [vp]),
        "rmx_net_info": (C.c_int, [vp, _I32P, _I64P, _I32P, _I32P]),
        "rmx_net_arrays": (C.c_int, [vp, C.POINTER(_I64P), C.POINTER(_I32P), C.POINTER(_F64P), C.POINTER(_F64P)]),
        "rmx_net_name": (C.c_char_p, [vp, C.c_int32]),
        "rmx_net_write": (C.c_int, [vp, C.c_char_p]),
        "rmx_net_upload": (C.c_int, [vp, vp]),
        "rmx_generate_sub_modules": (C.c_int, [vp, C.c_double, C.c_int32, C.c_uint64, _I32P]),
        "rmx_get_sub_modules": (C.c_int, [vp, _I32P]),
        "rmx_contract_sub_modules": (C.c_int, [vp, _I32P, _I64P]),
        "rmx_graph_upload": (C.c_int, [vp, C.c_int32, C.c_int64, _I64P, _I32P, _F64P, _F64P]),
        "rmx_batch_upload": (C.c_int, [vp, C.c_int32, _I32P, C.POINTER(C.c_uint64), C.c_int32, C.c_int64, _I64P,
                                       _I32P, _F64P, _F64P]),
        "rmx_batch_upload_seeds": (C.c_int, [vp, C.c_int32, _I32P, C.c_int32, C.POINTER(C.c_uint64), C.c_int32,
                                             C.c_int64, _I64P, _I32P, _F64P, _F64P]),
        "rmx_batch_results": (C.c_int, [vp, _F64P, _I32P, _I32P]),
        "rmx_set_bucket_cutoffs": (C.c_int, [vp, _I32P]),
        "rmx_get_bucket_cutoffs": (C.c_int, [vp, _I32P]),
        "rmx_autotune": (C.c_int, [vp, _I32P]),
        "rmx_load_network": (C.c_int, [vp, C.c_char_p, _I32P, _I64P, _I32P, _I32P]),
        "rmx_initiate": (C.c_int, [vp, _I32P, _F64P, _I64P]),
        "rmx_get_flows": (C.c_int, [vp, _F64P]),
        "rmx_get_node_exit": (C.c_int, [vp, _F64P]),
        "rmx_codelength": (C.c_int, [vp, _F64P, _F64P, _I32P]),
        "rmx_calculate_codelength": (C.c_int, [vp, _F64P]),
        "rmx_set_partition": (C.c_int, [vp, _I32P]),
        "rmx_get_partition": (C.c_int, [vp, _I32P]),
        "rmx_reset_active": (C.c_int, [vp]),
        "rmx_move_sweep": (C.c_int, [vp, C.c_int, C.c_double, C.c_uint64, _I64P, _F64P, _I64P]),
        "rmx_evaluate_moves": (C.c_int, [vp, C.c_int32, _I32P, _I32P, _F64P]),
        "rmx_update_members": (C.c_int, [vp]),
        "rmx_contract": (C.c_int, [vp, _I32P, _I64P]),
        "rmx_get_super": (C.c_int, [vp, _F64P, _F64P, _F64P, _F64P, _I32P, _I32P]),
        "rmx_get_super_arcs": (C.c_int, [vp, _I64P, _I32P, _F64P, _I64P, _I32P, _F64P]),
        "rmx_to_node_level": (C.c_int, [vp]),
        "rmx_partition": (C.c_int, [vp, C.POINTER(_Opts), C.POINTER(_Result)]),
        "rmx_write_tree": (C.c_int, [vp, C.c_char_p]),
        "rmx_write_clu": (C.c_int, [vp, C.c_char_p]),
        "rmx_timing": (C.c_int, [vp, _F64P, _F64P, _F64P, _I64P, _I64P]),
        "rmx_set_profiling": (C.c_int, [vp, C.c_int]),
        "rmx_kernel_stats": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int, _I64P, _F64P, _F64P]),
        "rmx_calculate_codelength_noalloc": (C.c_int, [C.c_int, vp, C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp,
                                                       vp, _F64P, C.c_char_p, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def device_count():
    return load_library().rmx_device_count()


def _p(a, t):
    return a.ctypes.data_as(t)


def calculate_codelength_noalloc(device, stream, n, m, out_ptr, out_idx, flow, p, tau, dangling, module):
    """rmx_calculate_codelength_noalloc on device addresses (ints): the
    codelength (nats) of partition ``module`` of a network held in device
    memory, computed with no device allocation.  Raises ValueError when the
    library refuses the input (status -22) and RuntimeError otherwise."""
    L = load_library()
    out = C.c_double(0.0)
    err = C.create_string_buffer(512)
    rc = L.rmx_calculate_codelength_noalloc(int(device), stream, int(n), int(m), out_ptr, out_idx, flow, p, tau,
                                            dangling, module, C.byref(out), err, 512)
    if rc != 0:
        raise (ValueError if rc == -22 else RuntimeError)(err.value.decode() or f"status {rc}")
    return out.value


class NetworkFile:
    """A network file parsed on the host (no GPU): .net / .txt with the
    reference's rules on several threads, or a .rmxcsr binary."""

    def __init__(self, path, threads=0):
        L = load_library()
        self._L = L
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = L.rmx_net_read(str(path).encode(), int(threads), C.byref(h), err, 512)
        if rc != 0:
            raise (IOError if rc == RMX_EIO else ValueError)(err.value.decode() or f"status {rc}")
        self._h = h
        n, m, ns, nd = C.c_int32(), C.c_int64(), C.c_int32(), C.c_int32()
        L.rmx_net_info(h, C.byref(n), C.byref(m), C.byref(ns), C.byref(nd))
        self.n, self.m, self.n_self_links, self.n_double_links = n.value, m.value, ns.value, nd.value

    def arrays(self):
        """(out_ptr, out_idx, w, node_w) as copies."""
        p, i, w, nw = _I64P(), _I32P(), _F64P(), _F64P()
        self._L.rmx_net_arrays(self._h, C.byref(p), C.byref(i), C.byref(w), C.byref(nw))
        n, m = self.n, self.m
        return (np.ctypeslib.as_array(p, shape=(n + 1,)).copy(),
                np.ctypeslib.as_array(i, shape=(max(m, 1),))[:m].copy(),
                np.ctypeslib.as_array(w, shape=(max(m, 1),))[:m].copy(),
                np.ctypeslib.as_array(nw, shape=(max(n, 1),))[:n].copy())

    def name(self, i):
        r = self._L.rmx_net_name(self._h, int(i))
        return None if r is None else r.decode()

    def write(self, path):
        if self._L.rmx_net_write(self._h, str(path).encode()) != 0:
            raise IOError(f"cannot write {path}")

    def upload_to(self, net):
        net._check(self._L.rmx_net_upload(net._h, self._h))
        net.n, net.m = self.n, self.m
        return net

    def close(self):
        if getattr(self, "_h", None):
            self._L.rmx_net_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Network:
    """One network resident on one GPU (the reference's ``Network``)."""

    def __init__(self, device=0, precision="fp64", batches=0):
        L = load_library()
        self._L = L
        h = C.c_void_p()
        rc = L.rmx_create(C.byref(h), int(device))
        if rc != 0:
            raise RuntimeError(f"rmx_create failed with status {rc} (no usable GPU?)")
        self._h = h
        self.n = 0
        self.m = 0
        self._check(L.rmx_set_precision(h, RMX_MIXED if precision == "mixed" else RMX_FP64))
        self._check(L.rmx_set_batches(h, int(batches)))

    def set_fp64_flow(self, blocked):
        """FP64 steady state: the blocked push / pull (True, the default) or
        the in-CSR pull (False).  Call before initiate."""
        self._check(self._L.rmx_set_fp64_flow(self._h, 1 if blocked else 0))

    def set_hub_path(self, mode):
        """Evaluate units with more than 2048 arcs by LDS partitions (0, the
        default), with one global hash table each (1), or by partitions with
        every hub sent through the overflow fallback (2); same moves."""
        self._check(self._L.rmx_set_hub_path(self._h, int(mode)))

    def set_attempt_mode(self, sequential):
        """run(attempts > 1) as one batch of copies of the network (False, the
        default, where the copies fit) or one attempt after another (True)."""
        self._check(self._L.rmx_set_attempt_mode(self._h, 1 if sequential else 0))

    def set_graph_capture(self, on):
        """Sweeps of a batch of instances and of sub-module partitions as one
        captured HIP graph each (True) or launch by launch with a host wait
        per batch (False, the default and the faster); same results."""
        self._check(self._L.rmx_set_graph_capture(self._h, 1 if on else 0))

    def set_candidate_bound(self, mode, bucket=-1):
        """Candidate bound of the move evaluation on (1, the default for
        units of more than 32 arcs) or off (0) for one degree bucket (0-6, 7
        the hub partitions) or all (-1); same moves either way."""
        self._check(self._L.rmx_set_candidate_bound(self._h, int(bucket), int(mode)))

    def close(self):
        if getattr(self, "_h", None):
            self._L.rmx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            msg = self._L.rmx_last_error(self._h)
            raise RuntimeError(f"relaxmap status {rc}: {msg.decode() if msg else ''}")

    # ----------------------------------------------------------- input
    def load(self, path):
        """load_pajek_format_network / load_linkList_format_network (FileIO.cpp)."""
        n, m, ns, nd = C.c_int32(), C.c_int64(), C.c_int32(), C.c_int32()
        self._check(self._L.rmx_load_network(self._h, str(path).encode(), C.byref(n), C.byref(m), C.byref(ns),
                                             C.byref(nd)))
        self.n, self.m = n.value, m.value
        self.n_self_links, self.n_double_links = ns.value, nd.value
        return self

    def batch_upload(self, node_off, seeds, out_ptr, out_idx, w, node_w=None):
        """A batch of independent networks: the union out-CSR, instance I owning
        nodes [node_off[I], node_off[I+1]); seeds[I] drives its sweeps."""
        node_off = np.ascontiguousarray(node_off, dtype=np.int32)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        out_ptr = np.ascontiguousarray(out_ptr, dtype=np.int64)
        out_idx = np.ascontiguousarray(out_idx, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        nwp = None
        if node_w is not None:
            node_w = np.ascontiguousarray(node_w, dtype=np.float64)
            nwp = _p(node_w, _F64P)
        n = len(out_ptr) - 1
        self._check(self._L.rmx_batch_upload(self._h, len(node_off) - 1, _p(node_off, _I32P),
                                             _p(seeds, C.POINTER(C.c_uint64)), int(n), len(out_idx),
                                             _p(out_ptr, _I64P), _p(out_idx, _I32P), _p(w, _F64P), nwp))
        self.n, self.m = int(n), len(out_idx)
        self.n_inst = len(node_off) - 1
        return self

    def batch_upload_seeds(self, graph_off, n_seed, seeds, out_ptr, out_idx, w, node_w=None):
        """A batch of networks under n_seed seeds each: the union out-CSR of the
        networks (network g owning nodes [graph_off[g], graph_off[g+1])), copied
        on the device; instance g * n_seed + r is network g with seeds[g * n_seed + r]."""
        graph_off = np.ascontiguousarray(graph_off, dtype=np.int32)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        out_ptr = np.ascontiguousarray(out_ptr, dtype=np.int64)
        out_idx = np.ascontiguousarray(out_idx, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        n_graph = len(graph_off) - 1
        if len(seeds) != n_graph * int(n_seed):
            raise ValueError("one seed per (network, seed) instance")
        nwp = None
        if node_w is not None:
            node_w = np.ascontiguousarray(node_w, dtype=np.float64)
            nwp = _p(node_w, _F64P)
        n = len(out_ptr) - 1
        self._check(self._L.rmx_batch_upload_seeds(self._h, n_graph, _p(graph_off, _I32P), int(n_seed),
                                                   _p(seeds, C.POINTER(C.c_uint64)), int(n), len(out_idx),
                                                   _p(out_ptr, _I64P), _p(out_idx, _I32P), _p(w, _F64P), nwp))
        self.n, self.m = int(n) * int(n_seed), len(out_idx) * int(n_seed)
        self.n_inst = n_graph * int(n_seed)
        return self

    def set_bucket_cutoffs(self, max_deg):
        """Upper degree bounds of the 7 evaluation buckets below the hubs."""
        a = np.ascontiguousarray(max_deg, dtype=np.int32)
        if a.shape != (7,):
            raise ValueError("7 cut-offs")
        self._check(self._L.rmx_set_bucket_cutoffs(self._h, _p(a, _I32P)))

    def bucket_cutoffs(self):
        a = np.empty(7, dtype=np.int32)
        self._check(self._L.rmx_get_bucket_cutoffs(self._h, _p(a, _I32P)))
        return a

    def autotune(self):
        """Times the evaluation kernels on the current level and installs the
        fastest cut-offs; returns them."""
        a = np.empty(7, dtype=np.int32)
        self._check(self._L.rmx_autotune(self._h, _p(a, _I32P)))
        return a

    def batch_results(self):
        """Per instance: codelength (nats), non-empty modules, steady-state iterations."""
        k = self.n_inst
        L = np.empty(k, dtype=np.float64)
        nm = np.empty(k, dtype=np.int32)
        it = np.empty(k, dtype=np.int32)
        self._check(self._L.rmx_batch_results(self._h, _p(L, _F64P), _p(nm, _I32P), _p(it, _I32P)))
        return L, nm, it

    def upload(self, n, out_ptr, out_idx, w, node_w=None):
        out_ptr = np.ascontiguousarray(out_ptr, dtype=np.int64)
        out_idx = np.ascontiguousarray(out_idx, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        nwp = None
        if node_w is not None:
            node_w = np.ascontiguousarray(node_w, dtype=np.float64)
            nwp = _p(node_w, _F64P)
        self._check(self._L.rmx_graph_upload(self._h, int(n), len(out_idx), _p(out_ptr, _I64P), _p(out_idx, _I32P),
                                             _p(w, _F64P), nwp))
        self.n, self.m = int(n), len(out_idx)
        return self

    # -------------------------------------------------- steady state
    def initiate(self):
        """Network::initiate (Module.cpp:210-333); returns (iterations, codelength)."""
        it, L0, nd = C.c_int32(), C.c_double(), C.c_int64()
        self._check(self._L.rmx_initiate(self._h, C.byref(it), C.byref(L0), C.byref(nd)))
        self.iterations, self.initial_codelength, self.n_dangling = it.value, L0.value, nd.value
        return self.iterations, self.initial_codelength

    def flows(self):
        a = np.empty(self.n, dtype=np.float64)
        self._check(self._L.rmx_get_flows(self._h, _p(a, _F64P)))
        return a

    def node_exit(self):
        a = np.empty(self.n, dtype=np.float64)
        self._check(self._L.rmx_get_node_exit(self._h, _p(a, _F64P)))
        return a

    # ------------------------------------------------------ codelength
    def tracked(self):
        L, S, nm = C.c_double(), C.c_double(), C.c_int32()
        self._check(self._L.rmx_codelength(self._h, C.byref(L), C.byref(S), C.byref(nm)))
        return L.value, S.value, nm.value

    @property
    def codelength(self):
        return self.tracked()[0]

    @property
    def sum_exit(self):
        return self.tracked()[1]

    @property
    def n_module(self):
        return self.tracked()[2]

    def calculate_codelength(self):
        """Network::calculateCodeLength (Module.cpp:2966-3002)."""
        L = C.c_double()
        self._check(self._L.rmx_calculate_codelength(self._h, C.byref(L)))
        return L.value

    # ------------------------------------------------------- partitions
    def set_partition(self, part):
        part = np.ascontiguousarray(part, dtype=np.int32)
        self._check(self._L.rmx_set_partition(self._h, _p(part, _I32P)))

    def partition(self):
        a = np.empty(self.n, dtype=np.int32)
        self._check(self._L.rmx_get_partition(self._h, _p(a, _I32P)))
        return a

    # ------------------------------------------------------------ sweeps
    def reset_active(self):
        self._check(self._L.rmx_reset_active(self._h))

    def move(self, prioritized=True, vthresh=0.0, seed=1):
        """One sweep (Network::prioritize_parallelMove / parallelMove and the
        super-node twins); returns the number of original nodes moved."""
        mv, L, ev = C.c_int64(), C.c_double(), C.c_int64()
        self._check(self._L.rmx_move_sweep(self._h, 1 if prioritized else 0, float(vthresh), int(seed) & (2**64 - 1),
                                           C.byref(mv), C.byref(L), C.byref(ev)))
        self.last_edge_visits = ev.value
        return mv.value

    def evaluate_moves(self):
        """The decision half of one sweep over every unit of the current level,
        all against the same module statistics, nothing committed: (best
        module per unit, -1 for none and n + 1 for the empty module; its
        delta codelength in nats)."""
        bm = np.empty(self.n, dtype=np.int32)
        bd = np.empty(self.n, dtype=np.float64)
        k = C.c_int32()
        self._check(self._L.rmx_evaluate_moves(self._h, self.n, C.byref(k), _p(bm, _I32P), _p(bd, _F64P)))
        return bm[:k.value].copy(), bd[:k.value].copy()

    def update_members_in_module(self):
        self._check(self._L.rmx_update_members(self._h))

    def convert_modules_to_super_nodes(self):
        ns, ms = C.c_int32(), C.c_int64()
        self._check(self._L.rmx_contract(self._h, C.byref(ns), C.byref(ms)))
        self.n_super, self.m_super = ns.value, ms.value
        return self.n_super, self.m_super

    def generate_sub_modules(self, threshold=1e-3, max_iter=10, seed=1):
        """generate_sub_modules (OmpRelaxmap.cpp:500-632): every module's induced
        subgraph partitioned as its own network; returns the sub-module count."""
        k = C.c_int32()
        self._check(self._L.rmx_generate_sub_modules(self._h, float(threshold), int(max_iter), int(seed),
                                                     C.byref(k)))
        return k.value

    def sub_modules(self):
        """Sub-module label per node after generate_sub_modules."""
        a = np.empty(self.n, dtype=np.int32)
        self._check(self._L.rmx_get_sub_modules(self._h, _p(a, _I32P)))
        return a

    def generate_super_nodes_from_sub_modules(self):
        """Network::generateSuperNodesFromSubModules (Module.cpp:3083-3158)."""
        ns, ms = C.c_int32(), C.c_int64()
        self._check(self._L.rmx_contract_sub_modules(self._h, C.byref(ns), C.byref(ms)))
        self.n_super, self.m_super = ns.value, ms.value
        return self.n_super, self.m_super

    def super_nodes(self):
        s, m = self.n_super, self.m_super
        size, ex, tp, dg = (np.empty(s) for _ in range(4))
        mod = np.empty(s, dtype=np.int32)
        nm = np.empty(s, dtype=np.int32)
        self._check(self._L.rmx_get_super(self._h, _p(size, _F64P), _p(ex, _F64P), _p(tp, _F64P), _p(dg, _F64P),
                                          _p(mod, _I32P), _p(nm, _I32P)))
        optr = np.empty(s + 1, dtype=np.int64)
        iptr = np.empty(s + 1, dtype=np.int64)
        oidx = np.empty(max(m, 1), dtype=np.int32)
        iidx = np.empty(max(m, 1), dtype=np.int32)
        ow = np.empty(max(m, 1))
        iw = np.empty(max(m, 1))
        self._check(self._L.rmx_get_super_arcs(self._h, _p(optr, _I64P), _p(oidx, _I32P), _p(ow, _F64P),
                                               _p(iptr, _I64P), _p(iidx, _I32P), _p(iw, _F64P)))
        return {"size": size, "exit": ex, "tp": tp, "dangling": dg, "mod": mod, "n_members": nm,
                "out_ptr": optr, "out_idx": oidx[:m], "out_w": ow[:m], "in_ptr": iptr, "in_idx": iidx[:m],
                "in_w": iw[:m]}

    def to_node_level(self):
        self._check(self._L.rmx_to_node_level(self._h))

    # ------------------------------------------------------------ driver
    def run(self, threshold=1e-3, vthresh=0.0, max_iter=10, prior=True, seed=1, verbose=False, attempts=1):
        """main()'s super-step loop (OmpRelaxmap.cpp:181-219); with attempts > 1,
        independent runs with seeds seed, seed + 1, ... keep the lowest codelength."""
        o = _Opts(float(threshold), float(vthresh), int(max_iter), 1 if prior else 0, int(seed), 1 if verbose else 0,
                  int(attempts))
        r = _Result()
        self._check(self._L.rmx_partition(self._h, C.byref(o), C.byref(r)))
        return PartitionResult(r.codelength, r.computed_codelength, r.n_module, r.super_steps, r.sweeps,
                               r.edge_visits, r.time_partition_s, r.time_sweeps_s, r.time_contract_s,
                               r.time_coarse_s, r.best_attempt)

    def write_tree(self, path):
        self._check(self._L.rmx_write_tree(self._h, str(path).encode()))

    def write_clu(self, path):
        self._check(self._L.rmx_write_clu(self._h, str(path).encode()))

    def set_profiling(self, on=True):
        self._check(self._L.rmx_set_profiling(self._h, 1 if on else 0))

    def kernel_stats(self):
        """{name: (launches, total ms, total algorithmic bytes)} of the profiling mode."""
        k = self._L.rmx_kernel_stats(self._h, -1, None, 0, None, None, None)
        out = {}
        for i in range(k):
            name = C.create_string_buffer(64)
            n, ms, b = C.c_int64(), C.c_double(), C.c_double()
            self._check(self._L.rmx_kernel_stats(self._h, i, name, 64, C.byref(n), C.byref(ms), C.byref(b)))
            out[name.value.decode()] = (n.value, ms.value, b.value)
        return out

    def timing(self):
        f, s, c, e, k = C.c_double(), C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
        self._check(self._L.rmx_timing(self._h, C.byref(f), C.byref(s), C.byref(c), C.byref(e), C.byref(k)))
        return {"flow_ms": f.value, "sweep_ms": s.value, "contract_ms": c.value, "edge_visits": e.value,
                "sweeps": k.value}
