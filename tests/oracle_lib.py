"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only -- never by the product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_MRG_SO = os.path.join(ROOT, "oracle", "_ref", "libref_mrg.so")

ERR_NAMES = ["range", "root", "self_parent", "depth_map", "edge_tree",
             "edge_gap", "pred_depth", "no_parent_edge"]


class MrgState(C.Structure):
    _fields_ = [("z1", C.c_uint32), ("z2", C.c_uint32), ("z3", C.c_uint32),
                ("z4", C.c_uint32), ("z5", C.c_uint32)]


class RefMrgState(C.Structure):
    # generator/splittable_mrg.h: uint_fast32_t is 64-bit on x86-64 glibc
    _fields_ = [("z1", C.c_uint64), ("z2", C.c_uint64), ("z3", C.c_uint64),
                ("z4", C.c_uint64), ("z5", C.c_uint64)]


class GraphInfo(C.Structure):
    _fields_ = [("log_n", C.c_int), ("n", C.c_int64), ("n_c", C.c_int64),
                ("m", C.c_int64), ("m_sym", C.c_int64),
                ("max_used_vertex", C.c_int64)]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        i64p, u32p, i32p = C.POINTER(C.c_int64), C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
        L.orc_make_mrg_seed.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(MrgState)]
        L.orc_mrg_skip.argtypes = [C.POINTER(MrgState), C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_mrg_get_uint.argtypes = [C.POINTER(MrgState)]
        L.orc_mrg_get_uint.restype = C.c_uint32
        L.orc_mrg_get_double.argtypes = [C.POINTER(MrgState)]
        L.orc_mrg_get_double.restype = C.c_double
        L.orc_make_random_numbers.argtypes = [C.c_int64, C.c_uint64, C.c_uint64, C.c_int64,
                                              C.POINTER(C.c_double)]
        L.orc_scramble_values.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
        L.orc_scramble.argtypes = [C.c_int64, C.c_int, C.c_uint64, C.c_uint64]
        L.orc_scramble.restype = C.c_int64
        L.orc_generate_edges.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64,
                                         i64p, i64p]
        L.orc_graph_build.argtypes = [i64p, i64p, C.c_int64]
        L.orc_graph_build.restype = C.c_void_p
        L.orc_graph_wrap.argtypes = [C.c_int, C.c_int64, C.c_int64, i64p, u32p, i64p, u32p, u32p]
        L.orc_graph_wrap.restype = C.c_void_p
        L.orc_graph_free.argtypes = [C.c_void_p]
        L.orc_graph_get_info.argtypes = [C.c_void_p, C.POINTER(GraphInfo)]
        for nm, rt in [("orc_graph_row_off", i64p), ("orc_graph_col", u32p),
                       ("orc_graph_deg_new", i64p), ("orc_graph_new2orig", u32p),
                       ("orc_graph_orig2new", u32p)]:
            getattr(L, nm).argtypes = [C.c_void_p]
            getattr(L, nm).restype = rt
        L.orc_has_edge.argtypes = [C.c_void_p, C.c_int64]
        L.orc_bfs.argtypes = [C.c_void_p, C.c_int64, i32p, i64p]
        L.orc_do_bfs.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_double, i64p,
                                 i64p, C.POINTER(C.c_int8), C.c_int]
        L.orc_do_bfs_omp.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_double,
                                     i64p, i64p, C.POINTER(C.c_int8), C.c_int, C.c_int]
        L.orc_do_schedule.argtypes = [C.c_int64, C.c_int64, C.c_int, C.c_double, C.c_double,
                                      C.c_int, i64p, i64p, C.c_int, i64p,
                                      C.POINTER(C.c_int8), C.c_int]
        L.orc_validate.argtypes = [i64p, i64p, C.c_int64, C.c_int64, C.c_int64, C.c_int,
                                   C.c_int64, i64p, i64p, i64p]
        u8p = C.POINTER(C.c_uint8)
        L.orc_corebfs_prune.argtypes = [C.c_void_p, i64p, u8p, u8p, i64p]
        L.orc_corebfs_reset_unreachable.argtypes = [C.c_void_p, i64p, u8p, u8p, i64p]
        L.orc_corebfs_reset_unreachable.restype = C.c_int64
        L.orc_corebfs_run.argtypes = [C.c_void_p, i64p, u8p, u8p, C.c_int64, i64p]
        L.orc_find_roots.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, i64p]
        L.orc_harmonic_mean_teps.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        L.orc_harmonic_mean_teps.restype = C.c_double
        _lib = L
    return _lib


def ref_mrg():
    """The reference's own splittable_mrg.c, compiled by oracle/Makefile."""
    if not os.path.exists(REF_MRG_SO):
        return None
    L = C.CDLL(REF_MRG_SO)
    L.mrg_seed.argtypes = [C.POINTER(RefMrgState), C.POINTER(C.c_uint64)]
    L.mrg_skip.argtypes = [C.POINTER(RefMrgState), C.c_uint64, C.c_uint64, C.c_uint64]
    L.mrg_get_uint_orig.argtypes = [C.POINTER(RefMrgState)]
    L.mrg_get_uint_orig.restype = C.c_uint64
    L.mrg_get_double_orig.argtypes = [C.POINTER(RefMrgState)]
    L.mrg_get_double_orig.restype = C.c_double
    return L


# ---------------------------------------------------------------- helpers
def make_seed(s1, s2):
    st = MrgState()
    lib().orc_make_mrg_seed(s1, s2, C.byref(st))
    return st


def generate_edges(scale, s1=2, s2=3, start=0, end=None, ef=16):
    if end is None:
        end = ef << scale
    v0 = np.empty(end - start, np.int64)
    v1 = np.empty(end - start, np.int64)
    lib().orc_generate_edges(scale, s1, s2, start, end, _p(v0, C.c_int64), _p(v1, C.c_int64))
    return v0, v1


def make_random_numbers(n, s1, s2, pos):
    out = np.empty(n, np.float64)
    lib().orc_make_random_numbers(n, s1, s2, pos, _p(out, C.c_double))
    return out


class Graph:
    """Oracle CSR (owned by the C library)."""

    def __init__(self, v0, v1):
        v0 = np.ascontiguousarray(v0, np.int64)
        v1 = np.ascontiguousarray(v1, np.int64)
        self.v0, self.v1 = v0, v1
        self.h = lib().orc_graph_build(_p(v0, C.c_int64), _p(v1, C.c_int64), len(v0))
        if not self.h:
            raise MemoryError("orc_graph_build failed")
        self.info = GraphInfo()
        lib().orc_graph_get_info(self.h, C.byref(self.info))

    @classmethod
    def wrap(cls, log_n, n_c, m, csr):
        """The oracle over CSR arrays built elsewhere (row_off, col, deg_new,
        new2orig, orig2new as int64/uint32 numpy arrays, kept alive here)."""
        self = cls.__new__(cls)
        self.v0 = self.v1 = None
        self._keep = {k: np.ascontiguousarray(csr[k], np.int64 if k in ("row_off", "deg_new")
                                               else np.uint32)
                      for k in ("row_off", "col", "deg_new", "new2orig", "orig2new")}
        k = self._keep
        self.h = lib().orc_graph_wrap(log_n, n_c, m, _p(k["row_off"], C.c_int64),
                                      _p(k["col"], C.c_uint32), _p(k["deg_new"], C.c_int64),
                                      _p(k["new2orig"], C.c_uint32), _p(k["orig2new"], C.c_uint32))
        if not self.h:
            raise MemoryError("orc_graph_wrap failed")
        self.info = GraphInfo()
        lib().orc_graph_get_info(self.h, C.byref(self.info))
        return self

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_graph_free(self.h)
            self.h = None

    @property
    def n(self):
        return self.info.n

    def arrays(self):
        L, n, m = lib(), self.info.n, self.info.m
        row = np.ctypeslib.as_array(L.orc_graph_row_off(self.h), (n + 1,)).copy()
        col = np.ctypeslib.as_array(L.orc_graph_col(self.h), (max(m, 1),))[:m].copy()
        deg = np.ctypeslib.as_array(L.orc_graph_deg_new(self.h), (n,)).copy()
        n2o = np.ctypeslib.as_array(L.orc_graph_new2orig(self.h), (n,)).copy()
        o2n = np.ctypeslib.as_array(L.orc_graph_orig2new(self.h), (n,)).copy()
        return dict(row_off=row, col=col, deg_new=deg, new2orig=n2o, orig2new=o2n)

    def has_edge(self, v):
        return bool(lib().orc_has_edge(self.h, v))

    def bfs(self, root, want_pred=False):
        lev = np.empty(self.info.n, np.int32)
        pred = np.empty(self.info.n, np.int64) if want_pred else None
        lib().orc_bfs(self.h, root, _p(lev, C.c_int32),
                      _p(pred, C.c_int64) if want_pred else None)
        return (lev, pred) if want_pred else lev

    def do_bfs(self, root, ef=16, alpha=10.0, beta=14.0):
        pred = np.empty(self.info.n, np.int64)
        at = np.empty(9, np.int64)
        dirs = np.zeros(256, np.int8)
        k = lib().orc_do_bfs(self.h, root, ef, alpha, beta, _p(pred, C.c_int64),
                             _p(at, C.c_int64), _p(dirs, C.c_int8), 256)
        return pred, at, dirs[:min(k, 256)].copy()

    def do_bfs_omp(self, root, nthreads, ef=16, alpha=10.0, beta=14.0):
        pred = np.empty(self.info.n, np.int64)
        at = np.empty(9, np.int64)
        dirs = np.zeros(256, np.int8)
        k = lib().orc_do_bfs_omp(self.h, root, ef, alpha, beta, _p(pred, C.c_int64),
                                 _p(at, C.c_int64), _p(dirs, C.c_int8), 256, nthreads)
        return pred, at, dirs[:min(k, 256)].copy()

    def find_roots(self, nroots, r1=0, r2=0):
        out = np.empty(nroots, np.int64)
        k = lib().orc_find_roots(self.h, nroots, r1, r2, _p(out, C.c_int64))
        return out[:k].copy()

    def corebfs(self):
        """Prune + reset_unreachable; returns a dict of arrays and stats."""
        n = self.info.n
        tp = np.empty(n, np.int64)
        hp = np.empty(n, np.uint8)
        ic = np.empty(n, np.uint8)
        st = np.empty(4, np.int64)
        lib().orc_corebfs_prune(self.h, _p(tp, C.c_int64), _p(hp, C.c_uint8), _p(ic, C.c_uint8),
                                _p(st, C.c_int64))
        hd = np.zeros(1, np.int64)
        removed = lib().orc_corebfs_reset_unreachable(self.h, _p(tp, C.c_int64), _p(hp, C.c_uint8),
                                                      _p(ic, C.c_uint8), _p(hd, C.c_int64))
        return dict(tree_parent=tp, has_parent=hp, in_core=ic, core_vertices=int(st[0]),
                    core_edges=int(st[1]), tree_vertices=int(st[2]), iterations=int(st[3]),
                    removed=int(removed), hd_root=int(hd[0]))

    def corebfs_run(self, cb, root):
        pred = np.empty(self.info.n, np.int64)
        lv = lib().orc_corebfs_run(self.h, _p(cb["tree_parent"], C.c_int64),
                                   _p(cb["has_parent"], C.c_uint8), _p(cb["in_core"], C.c_uint8),
                                   int(root), _p(pred, C.c_int64))
        return pred, lv

    def validate(self, root, pred, corebfs=False, nglobal=None):
        pred = np.ascontiguousarray(pred, np.int64).copy()
        if nglobal is None:
            nglobal = self.info.max_used_vertex + 1
        return validate(self.v0, self.v1, nglobal, self.info.n, root, pred, corebfs)


def validate(v0, v1, nglobal, nlocal, root, pred, corebfs=False):
    """Returns (ok, te, errs dict); pred is modified in place for corebfs."""
    te = np.zeros(1, np.int64)
    errs = np.zeros(8, np.int64)
    ok = lib().orc_validate(_p(v0, C.c_int64), _p(v1, C.c_int64), len(v0), nglobal, nlocal,
                            int(corebfs), root, _p(pred, C.c_int64), _p(te, C.c_int64),
                            _p(errs, C.c_int64))
    return bool(ok), int(te[0]), dict(zip(ERR_NAMES, errs.tolist()))


def do_schedule(m, n_c, nq_size, nq_edges, ef=16, alpha=10.0, beta=14.0, root_level=0):
    nq_size = np.ascontiguousarray(nq_size, np.int64)
    nq_edges = np.ascontiguousarray(nq_edges, np.int64)
    at = np.empty(9, np.int64)
    dirs = np.zeros(256, np.int8)
    k = lib().orc_do_schedule(m, n_c, ef, alpha, beta, root_level, _p(nq_size, C.c_int64),
                              _p(nq_edges, C.c_int64), len(nq_size), _p(at, C.c_int64),
                              _p(dirs, C.c_int8), 256)
    return at, dirs[:min(k, 256)].copy()
