"""ctypes binding of libg500bfs_hip.so (C ABI in include/g500bfs.h).

Thin plumbing for the test-suite and bench.py: every call goes straight to
the HIP library; there is no CPU fallback.  A missing library raises.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libg500bfs_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "g500bfs.h")

AT_NUM = 9
CONSTRUCT_COREBFS = 1
CONSTRUCT_SYMMETRIC_INPUT = 2
VERR_NUM = 8
VERR_NAMES = ["range", "root", "self_parent", "depth_map", "edge_tree",
              "edge_gap", "pred_depth", "no_parent_edge"]


class G500Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"g500 error {code}: {msg}")
        self.code = code


class GraphInfo(C.Structure):
    _fields_ = [("log_n", C.c_int), ("corebfs", C.c_int), ("n", C.c_int64),
                ("n_c", C.c_int64), ("m", C.c_int64), ("m_sym", C.c_int64),
                ("nedges_input", C.c_int64), ("max_used_vertex", C.c_int64),
                ("max_degree", C.c_int64), ("core_vertices", C.c_int64),
                ("tree_vertices", C.c_int64), ("device_bytes", C.c_int64),
                ("removed_parents", C.c_int64), ("hd_root", C.c_int64)]


class LevelTrace(C.Structure):
    _fields_ = [("level", C.c_int), ("top_down", C.c_int), ("nq_size", C.c_int64),
                ("nq_edges", C.c_int64), ("gpu_ms", C.c_float), ("exec_bottom_up", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("nlevels", C.c_int), ("gpu_ms", C.c_double), ("td_ms", C.c_double),
                ("bu_ms", C.c_double), ("td_edges", C.c_int64), ("root_level", C.c_int64),
                ("gather_ms", C.c_double), ("sweep_ms", C.c_double),
                ("sweeps", C.c_int32), ("counted", C.c_int32),
                ("alg_bytes", C.c_int64), ("sweep_alg_bytes", C.c_int64),
                ("gather_alg_bytes", C.c_int64), ("gap_ms", C.c_double),
                ("gaps", C.c_int32), ("mispredicted", C.c_int32)]


class MsbfsTuning(C.Structure):
    _fields_ = [("heavy_degree", C.c_int64), ("medium_degree", C.c_int64),
                ("medium_group", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("use_graph", C.c_int32), ("heavy_chunk", C.c_int32),
                ("td_alpha", C.c_double), ("stream_frac", C.c_double),
                ("hot_bytes", C.c_int64), ("dense_algo", C.c_int32), ("split_bp", C.c_int32)]


class MsbfsStats(C.Structure):
    _fields_ = [("nroots", C.c_int32), ("nlevels", C.c_int32), ("gpu_ms", C.c_double),
                ("discovered", C.c_int64), ("lane_discoveries", C.c_int64 * 64)]


MSBFS_MAX_ROOTS = 256
NOALLOC_MAX_PAIRS = 65536  # G500_NOALLOC_MAX_PAIRS

_lib = None


def load():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: build it (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp, i64, i64p = C.c_void_p, C.c_int64, C.POINTER(C.c_int64)
    u32p = C.POINTER(C.c_uint32)
    sig = {
        "g500_graph_generate": [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.POINTER(vp)],
        "g500_graph_from_edges": [C.c_int, i64p, i64p, i64, C.POINTER(vp)],
        "g500_graph_create": [C.c_int, i64, C.POINTER(vp)],
        "g500_graph_append_edges": [vp, i64p, i64p, i64],
        "g500_graph_construct": [vp, C.c_int],
        "g500_graph_get_info": [vp, C.POINTER(GraphInfo)],
        "g500_graph_has_edge": [vp, i64, C.POINTER(C.c_int)],
        "g500_graph_copy_edges": [vp, i64, i64, i64p, i64p],
        "g500_graph_copy_csr": [vp, i64p, u32p, i64p, u32p, u32p],
        "g500_graph_copy_forest": [vp, i64p, C.POINTER(C.c_uint8)],
        "g500_find_roots": [vp, C.c_int, C.c_int, C.c_int, i64p, C.POINTER(C.c_int)],
        "g500_graph_free": [vp],
        "g500_make_random_numbers": [i64, C.c_uint64, C.c_uint64, i64, C.POINTER(C.c_double)],
        "g500_bfs_create": [vp, C.POINTER(vp)],
        "g500_bfs_create_ex": [vp, C.c_int, C.c_double, C.c_double, C.POINTER(vp)],
        "g500_bfs_run": [vp, i64, C.c_int, C.c_double, C.c_double, i64p],
        "g500_bfs_get_pred": [vp, i64p],
        "g500_bfs_get_levels": [vp, C.POINTER(C.c_uint16)],
        "g500_bfs_validate": [vp, i64, i64p, i64p, i64p, C.POINTER(C.c_int)],
        "g500_bfs_set_timing": [vp, C.c_int],
        "g500_bfs_set_exec_alpha": [vp, C.c_double],
        "g500_bfs_set_counting": [vp, C.c_int],
        "g500_bfs_get_stats": [vp, C.POINTER(RunStats), C.POINTER(LevelTrace), C.c_int],
        "g500_bfs_destroy": [vp],
        "g500_msbfs_create": [vp, C.c_int, C.POINTER(vp)],
        "g500_msbfs_footprint": [vp, C.c_int, C.c_int, i64p],
        "g500_msbfs_create_ex": [vp, C.c_int, C.c_int, C.POINTER(vp)],
        "g500_device_memory": [C.c_int, i64p, i64p],
        "g500_msbfs_set_tuning": [vp, C.POINTER(MsbfsTuning)],
        "g500_msbfs_get_tuning": [vp, C.POINTER(MsbfsTuning)],
        "g500_msbfs_launch": [vp, i64p, C.c_int],
        "g500_msbfs_query": [vp, C.POINTER(C.c_int)],
        "g500_msbfs_wait": [vp],
        "g500_msbfs_run": [vp, i64p, C.c_int],
        "g500_msbfs_get_levels": [vp, C.c_int, C.POINTER(C.c_uint8)],
        "g500_msbfs_device_levels": [vp, C.POINTER(vp), i64p],
        "g500_msbfs_get_te": [vp, i64p],
        "g500_msbfs_gather_levels": [vp, i64p, i64, C.POINTER(C.c_uint8)],
        "g500_msbfs_gather_levels_noalloc": [vp, i64p, i64, C.POINTER(C.c_uint8)],
        "g500_msbfs_autotune": [vp, i64p, C.c_int, C.c_int, C.POINTER(MsbfsTuning),
                                C.POINTER(C.c_double)],
        "g500_msbfs_get_stats": [vp, C.POINTER(MsbfsStats)],
        "g500_msbfs_destroy": [vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.g500_last_error.argtypes = []
    L.g500_last_error.restype = C.c_char_p
    _lib = L
    return L


def device_memory(device=0):
    """(free, total) device bytes."""
    f, t = C.c_int64(), C.c_int64()
    _check(load().g500_device_memory(device, C.byref(f), C.byref(t)))
    return f.value, t.value


MSBFS_NO_LEVEL_ARRAYS = 1


def msbfs_footprint(graph, max_roots, level_arrays=True):
    b = C.c_int64()
    flags = 0 if level_arrays else MSBFS_NO_LEVEL_ARRAYS
    _check(load().g500_msbfs_footprint(graph.h, int(max_roots), flags, C.byref(b)))
    return b.value


def make_random_numbers(n, seed1, seed2, position):
    out = np.empty(n, np.float64)
    _check(load().g500_make_random_numbers(n, seed1, seed2, position, _p(out, C.c_double)))
    return out


def _check(rc):
    if rc != 0:
        raise G500Error(rc, load().g500_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Graph:
    """A device-resident Graph500 graph (edge list + CSR)."""

    def __init__(self, handle):
        self.h = handle
        self._info = None
        # engines built on this graph hold pointers into it: free() destroys
        # the live ones first (the library refuses to free a graph under them)
        self._engines = weakref.WeakSet()

    @classmethod
    def generate(cls, scale, edgefactor=16, seed1=2, seed2=3, device=0):
        h = C.c_void_p()
        _check(load().g500_graph_generate(device, scale, edgefactor, seed1, seed2, C.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, v0, v1, device=0):
        v0 = np.ascontiguousarray(v0, np.int64)
        v1 = np.ascontiguousarray(v1, np.int64)
        h = C.c_void_p()
        _check(load().g500_graph_from_edges(device, _p(v0, C.c_int64), _p(v1, C.c_int64),
                                            len(v0), C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, capacity, device=0):
        """An empty graph filled by append_edges() before construct()."""
        h = C.c_void_p()
        _check(load().g500_graph_create(device, int(capacity), C.byref(h)))
        return cls(h)

    def append_edges(self, v0, v1):
        v0 = np.ascontiguousarray(v0, np.int64)
        v1 = np.ascontiguousarray(v1, np.int64)
        _check(load().g500_graph_append_edges(self.h, _p(v0, C.c_int64), _p(v1, C.c_int64),
                                              len(v0)))
        self._info = None
        return self

    def construct(self, corebfs=False, symmetric_input=False):
        flags = (CONSTRUCT_COREBFS if corebfs else 0) | (CONSTRUCT_SYMMETRIC_INPUT if symmetric_input else 0)
        _check(load().g500_graph_construct(self.h, flags))
        self._info = None
        return self

    @property
    def info(self):
        if self._info is None:
            inf = GraphInfo()
            _check(load().g500_graph_get_info(self.h, C.byref(inf)))
            self._info = inf
        return self._info

    def has_edge(self, v):
        out = C.c_int()
        _check(load().g500_graph_has_edge(self.h, v, C.byref(out)))
        return bool(out.value)

    def edges(self, start=0, count=None):
        if count is None:
            count = self.info.nedges_input - start
        v0 = np.empty(count, np.int64)
        v1 = np.empty(count, np.int64)
        _check(load().g500_graph_copy_edges(self.h, start, count, _p(v0, C.c_int64),
                                            _p(v1, C.c_int64)))
        return v0, v1

    def csr(self):
        inf = self.info
        row = np.empty(inf.n + 1, np.int64)
        col = np.empty(max(inf.m, 1), np.uint32)
        deg = np.empty(inf.n, np.int64)
        n2o = np.empty(inf.n, np.uint32)
        o2n = np.empty(inf.n, np.uint32)
        _check(load().g500_graph_copy_csr(self.h, _p(row, C.c_int64), _p(col, C.c_uint32),
                                          _p(deg, C.c_int64), _p(n2o, C.c_uint32),
                                          _p(o2n, C.c_uint32)))
        return dict(row_off=row, col=col[:inf.m], deg_new=deg, new2orig=n2o, orig2new=o2n)

    def forest(self):
        n = self.info.n
        tp = np.empty(n, np.int64)
        hp = np.empty(n, np.uint8)
        _check(load().g500_graph_copy_forest(self.h, _p(tp, C.c_int64), _p(hp, C.c_uint8)))
        return tp, hp

    def refresh(self):
        self._info = None
        return self

    def find_roots(self, nroots, r1=0, r2=0):
        out = np.empty(max(nroots, 1), np.int64)
        found = C.c_int()
        _check(load().g500_find_roots(self.h, nroots, r1, r2, _p(out, C.c_int64),
                                      C.byref(found)))
        return out[:found.value].copy()

    def free(self):
        if self.h:
            for e in list(self._engines):
                e.destroy()
            _check(load().g500_graph_free(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Bfs:
    """BFS engine on a constructed graph (BfsBase's prepare/run/get_pred/end)."""

    def __init__(self, graph):
        self.graph = graph
        h = C.c_void_p()
        _check(load().g500_bfs_create(graph.h, C.byref(h)))
        self.h = h
        graph._engines.add(self)

    def run(self, root, edgefactor=16, alpha=10.0, beta=14.0):
        at = np.full(AT_NUM, -1, np.int64)
        _check(load().g500_bfs_run(self.h, int(root), edgefactor, alpha, beta,
                                   _p(at, C.c_int64)))
        return at

    def pred(self):
        out = np.empty(self.graph.info.n, np.int64)
        _check(load().g500_bfs_get_pred(self.h, _p(out, C.c_int64)))
        return out

    def levels(self):
        out = np.empty(self.graph.info.n, np.uint16)
        _check(load().g500_bfs_get_levels(self.h, _p(out, C.c_uint16)))
        return out

    def validate(self, root, host_pred=None):
        te = C.c_int64()
        errs = np.zeros(VERR_NUM, np.int64)
        ok = C.c_int()
        hp = None
        if host_pred is not None:
            host_pred = np.ascontiguousarray(host_pred, np.int64)
            hp = _p(host_pred, C.c_int64)
        _check(load().g500_bfs_validate(self.h, int(root), hp, C.byref(te),
                                        _p(errs, C.c_int64), C.byref(ok)))
        return bool(ok.value), te.value, dict(zip(VERR_NAMES, errs.tolist()))

    def set_exec_alpha(self, exec_alpha):
        _check(load().g500_bfs_set_exec_alpha(self.h, float(exec_alpha)))

    def set_timing(self, on=True):
        _check(load().g500_bfs_set_timing(self.h, int(on)))

    def set_counting(self, on=True):
        """Counting runs report the bytes their kernels need (stats())."""
        _check(load().g500_bfs_set_counting(self.h, int(on)))

    def stats(self, max_levels=256):
        st = RunStats()
        tr = (LevelTrace * max_levels)()
        _check(load().g500_bfs_get_stats(self.h, C.byref(st), tr, max_levels))
        levels = [dict(level=tr[i].level, top_down=tr[i].top_down, nq_size=tr[i].nq_size,
                       nq_edges=tr[i].nq_edges, gpu_ms=tr[i].gpu_ms,
                       exec_bottom_up=tr[i].exec_bottom_up)
                  for i in range(min(st.nlevels, max_levels))]
        return st, levels

    def destroy(self):
        if self.h:
            _check(load().g500_bfs_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class MultiBfs:
    """Batched bit-parallel BFS: up to 256 roots per launch, uint8 levels."""

    def __init__(self, graph, max_roots=MSBFS_MAX_ROOTS, level_arrays=True):
        self.graph = graph
        h = C.c_void_p()
        flags = 0 if level_arrays else MSBFS_NO_LEVEL_ARRAYS
        if flags:
            _check(load().g500_msbfs_create_ex(graph.h, int(max_roots), flags, C.byref(h)))
        else:
            _check(load().g500_msbfs_create(graph.h, int(max_roots), C.byref(h)))
        self.h = h
        self.nroots = 0
        graph._engines.add(self)

    def tuning(self):
        t = MsbfsTuning()
        _check(load().g500_msbfs_get_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in MsbfsTuning._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = MsbfsTuning(**cur)
        _check(load().g500_msbfs_set_tuning(self.h, C.byref(t)))

    def launch(self, roots):
        self._roots = np.ascontiguousarray(roots, np.int64)
        _check(load().g500_msbfs_launch(self.h, _p(self._roots, C.c_int64), len(self._roots)))
        self.nroots = len(self._roots)

    def done(self):
        d = C.c_int()
        _check(load().g500_msbfs_query(self.h, C.byref(d)))
        return bool(d.value)

    def wait(self):
        _check(load().g500_msbfs_wait(self.h))

    def run(self, roots):
        r = np.ascontiguousarray(roots, np.int64)
        _check(load().g500_msbfs_run(self.h, _p(r, C.c_int64), len(r)))
        self.nroots = len(r)

    def levels(self, index):
        out = np.empty(self.graph.info.n, np.uint8)
        _check(load().g500_msbfs_get_levels(self.h, int(index), _p(out, C.c_uint8)))
        return out

    def device_levels(self):
        p, stride = C.c_void_p(), C.c_int64()
        _check(load().g500_msbfs_device_levels(self.h, C.byref(p), C.byref(stride)))
        return p.value, stride.value

    def gather(self, root_index, vertices, noalloc=False):
        """Levels of `vertices` (original IDs) for batch roots `root_index`.
        noalloc=True stages them in the library's fixed device array
        (g500_msbfs_gather_levels_noalloc, at most NOALLOC_MAX_PAIRS)."""
        ri = np.broadcast_to(np.asarray(root_index, np.int64), np.shape(vertices))
        pairs = np.ascontiguousarray(np.stack([ri, np.asarray(vertices, np.int64)], axis=-1)
                                     .reshape(-1, 2))
        out = np.empty(len(pairs), np.uint8)
        fn = (load().g500_msbfs_gather_levels_noalloc if noalloc
              else load().g500_msbfs_gather_levels)
        if len(pairs) or noalloc:
            _check(fn(self.h, _p(pairs, C.c_int64), len(pairs), _p(out, C.c_uint8)))
        return out

    def autotune(self, roots, rounds=1):
        r = np.ascontiguousarray(roots, np.int64)
        t, ms = MsbfsTuning(), C.c_double()
        _check(load().g500_msbfs_autotune(self.h, _p(r, C.c_int64), len(r), int(rounds),
                                          C.byref(t), C.byref(ms)))
        self.nroots = len(r)
        return {f: getattr(t, f) for f, _ in MsbfsTuning._fields_}, ms.value

    def te(self):
        out = np.empty(max(self.nroots, 1), np.int64)
        _check(load().g500_msbfs_get_te(self.h, _p(out, C.c_int64)))
        return out[:self.nroots]

    def stats(self):
        st = MsbfsStats()
        _check(load().g500_msbfs_get_stats(self.h, C.byref(st)))
        return dict(nroots=st.nroots, nlevels=st.nlevels, gpu_ms=st.gpu_ms,
                    discovered=st.discovered,
                    lane_discoveries=list(st.lane_discoveries[:min(st.nlevels, 64)]))

    def destroy(self):
        if self.h:
            _check(load().g500_msbfs_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
