"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement (a wrong
MRG constant or skip, an off-by-one in scramble, a wrong quadrant split, a
dropped dedupe, a wrong direction-switch comparison, a missing validator
check) fails at least one of them:
  * MRG            -> the reference's own splittable_mrg.c (oracle/_ref)
  * generator/CSR  -> pf_nedge (mpi/utils.hpp:3688-3733) + logged edge counts
  * BFS levels     -> scipy.sparse.csgraph (textbook library BFS)
  * validator      -> planted mistakes, one per check, plus valid trees
  * roots          -> formula re-evaluated with the reference's doubles
  * DO schedule    -> hand-traced sequences and the logged SCALE-20 trace
"""
import ctypes as C
import os
import random

import numpy as np
import pytest

import oracle_lib as O


# ------------------------------------------------------------------ MRG
def _ref_or_skip():
    R = O.ref_mrg()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return R


def test_mrg_skip_matches_reference_build():
    R = _ref_or_skip()
    rnd = random.Random(7)
    for trial in range(300):
        s1, s2 = rnd.randrange(1 << 64), rnd.randrange(1 << 64)
        hi = rnd.choice([0, 2, 50, rnd.randrange(1 << 64)])
        mid = rnd.choice([0, 7, rnd.randrange(1 << 64)])
        lo = rnd.choice([0, 1, rnd.randrange(1 << 64)])
        st = O.make_seed(s1, s2)
        rs = O.RefMrgState(st.z1, st.z2, st.z3, st.z4, st.z5)
        O.lib().orc_mrg_skip(C.byref(st), hi, mid, lo)
        R.mrg_skip(C.byref(rs), hi, mid, lo)
        assert (st.z1, st.z2, st.z3, st.z4, st.z5) == (rs.z1, rs.z2, rs.z3, rs.z4, rs.z5)
        for _ in range(4):
            assert O.lib().orc_mrg_get_uint(C.byref(st)) == R.mrg_get_uint_orig(C.byref(rs))
            assert O.lib().orc_mrg_get_double(C.byref(st)) == R.mrg_get_double_orig(C.byref(rs))


def test_mrg_skip_is_additive_and_one_step():
    L = O.lib()
    a = O.make_seed(2, 3)
    b = O.make_seed(2, 3)
    L.orc_mrg_skip(C.byref(a), 0, 5, 12345)
    L.orc_mrg_skip(C.byref(a), 1, 3, 999)
    L.orc_mrg_skip(C.byref(b), 1, 8, 12345 + 999)
    assert (a.z1, a.z2, a.z3, a.z4, a.z5) == (b.z1, b.z2, b.z3, b.z4, b.z5)
    c = O.make_seed(2, 3)
    d = O.make_seed(2, 3)
    L.orc_mrg_skip(C.byref(c), 0, 0, 1)
    x = L.orc_mrg_get_uint(C.byref(d))
    assert c.z1 == x and (c.z2, c.z3, c.z4, c.z5) == (d.z2, d.z3, d.z4, d.z5)
    # the recurrence itself: z1' = 107374182*z1 + 104480*z5 mod 2^31-1
    e = O.make_seed(11, 13)
    z1, z5 = e.z1, e.z5
    assert L.orc_mrg_get_uint(C.byref(e)) == (107374182 * z1 + 104480 * z5) % 0x7FFFFFFF


def test_mrg_seed_layout():
    # generator/graph_generator.hpp:492-498
    st = O.make_seed(2, 3)
    assert (st.z1, st.z2, st.z3, st.z4, st.z5) == (3, 1, 4, 1, 1)
    s1, s2 = (0xF << 60) | (5 << 30) | 7, (0xA << 60) | (9 << 30) | 11
    st = O.make_seed(s1, s2)
    assert (st.z1, st.z2, st.z3, st.z4, st.z5) == (8, 6, 12, 10, (0xA << 4) + 0xF + 1)


# ------------------------------------------------------------ generator
def test_scramble_is_a_bijection():
    sv = (C.c_uint64 * 2)()
    O.lib().orc_scramble_values(2, 3, sv)
    for scale in (1, 5, 10, 13):
        out = {O.lib().orc_scramble(v, scale, sv[0], sv[1]) for v in range(1 << scale)}
        assert out == set(range(1 << scale))


def test_edges_in_range_and_deterministic():
    v0, v1 = O.generate_edges(10)
    assert len(v0) == 16 << 10
    assert v0.min() >= 0 and v1.min() >= 0 and max(v0.max(), v1.max()) < 1 << 10
    a0, a1 = O.generate_edges(10, start=1000, end=3000)
    assert np.array_equal(a0, v0[1000:3000]) and np.array_equal(a1, v1[1000:3000])


def test_quadrant_frequencies():
    # R-MAT A/B/C/D = .57/.19/.19/.05 gives a heavily skewed degree
    # distribution (the quadrants themselves are hidden by scramble).
    v0, v1 = O.generate_edges(12)
    deg = np.bincount(np.concatenate([v0[v0 != v1], v1[v0 != v1]]), minlength=1 << 12)
    assert deg.max() > 20 * deg.mean()  # power-law-like skew


@pytest.mark.parametrize("scale", [8, 10, 12, 14, 16])
def test_pf_nedge_small_scales(scale, golden):
    v0, v1 = O.generate_edges(scale)
    g = O.Graph(v0, v1)
    roots = g.find_roots(4)
    for r in roots:
        lev, pred = g.bfs(int(r), want_pred=True)
        ok, te, errs = g.validate(int(r), pred)
        assert ok, errs
        assert te == golden["pf_nedge"][scale]


def test_reference_logged_counts_scale16(golden):
    v0, v1 = O.generate_edges(16)
    g = O.Graph(v0, v1)
    assert g.info.log_n == 16
    assert g.info.m_sym == golden["edge_counts"]["16"]["sym"]
    assert g.info.m == golden["edge_counts"]["16"]["dedup"]


def test_reference_logged_counts_scale20(golden):
    v0, v1 = O.generate_edges(20)
    g = O.Graph(v0, v1)
    ref = golden["edge_counts"]["20"]
    assert (g.info.m_sym, g.info.m, g.info.n_c) == (ref["sym"], ref["dedup"], ref["n_c"])
    r = int(g.find_roots(1)[0])
    pred, at, dirs = g.do_bfs(r)
    ok, te, _ = g.validate(r, pred)
    assert ok and te == golden["pf_nedge"][20]
    trace = ["TD" if d else "BU" for d in dirs]
    assert trace == golden["scale20_direction_trace_example"]


# ------------------------------------------------------------------ CSR
def test_csr_invariants():
    v0, v1 = O.generate_edges(12)
    g = O.Graph(v0, v1)
    a = g.arrays()
    n, nc, m = g.info.n, g.info.n_c, g.info.m
    row, col = a["row_off"], a["col"]
    assert row[0] == 0 and row[-1] == m and np.all(np.diff(row) >= 0)
    assert np.all(np.diff(row)[nc:] == 0) and np.all(np.diff(row)[:nc] > 0)
    # rows strictly increasing (sorted + unique), no loops, symmetric
    src = np.repeat(np.arange(n), np.diff(row))
    assert np.all(src != col)
    key = src.astype(np.int64) * n + col
    assert np.all(np.diff(key) > 0)
    rkey = np.sort(col.astype(np.int64) * n + src)
    assert np.array_equal(rkey, key)
    # relabel: degree descending, ties by original ID
    deg = a["deg_new"]
    n2o = a["new2orig"].astype(np.int64)
    assert np.all(np.diff(deg) <= 0)
    ties = (np.diff(deg) == 0)
    assert np.all(np.diff(n2o)[ties] > 0)
    assert np.array_equal(a["orig2new"][n2o], np.arange(n))
    # degree counts duplicates, the CSR does not
    nl = v0 != v1
    d_orig = np.bincount(np.concatenate([v0[nl], v1[nl]]), minlength=n)
    assert np.array_equal(deg, d_orig[n2o])
    uniq = np.unique(np.concatenate([v0[nl] * n + v1[nl], v1[nl] * n + v0[nl]]))
    assert len(uniq) == m


# ------------------------------------------------------------ BFS levels
@pytest.mark.parametrize("scale", [10, 13])
def test_levels_match_scipy(scale):
    sp = pytest.importorskip("scipy.sparse")
    csg = pytest.importorskip("scipy.sparse.csgraph")
    v0, v1 = O.generate_edges(scale)
    g = O.Graph(v0, v1)
    n = g.info.n
    A = sp.coo_matrix((np.ones(len(v0)), (v0, v1)), shape=(n, n)).tocsr()
    A = ((A + A.T) > 0).astype(np.int8)
    for r in g.find_roots(5):
        lev = g.bfs(int(r))
        order, _ = csg.breadth_first_order(A, int(r), directed=False)
        dist = csg.shortest_path(A, unweighted=True, directed=False, indices=int(r))
        exp = np.where(np.isinf(dist), -1, dist).astype(np.int64)
        assert np.array_equal(lev, exp)
        assert set(order.tolist()) == set(np.nonzero(exp >= 0)[0].tolist())


def test_do_bfs_levels_equal_textbook():
    v0, v1 = O.generate_edges(14)
    g = O.Graph(v0, v1)
    for r in g.find_roots(6):
        lev = g.bfs(int(r))
        pred, at, dirs = g.do_bfs(int(r))
        d = (pred.astype(np.uint64) >> np.uint64(48)).astype(np.int64)
        d[d == 0xFFFF] = -1
        assert np.array_equal(d, lev)
        ok, te, errs = g.validate(int(r), pred)
        assert ok, errs


def test_do_bfs_omp_matches_serial():
    """The OpenMP restatement (the CPU baseline) keeps the serial one's
    schedule, at_data and depths; its parents pass the validator.  Alpha and
    beta extremes force pure top-down and never-return runs too."""
    v0, v1 = O.generate_edges(15)
    g = O.Graph(v0, v1)
    for alpha, beta in ((10.0, 14.0), (1e9, 14.0), (10.0, 1e-6)):
        for r in g.find_roots(4):
            r = int(r)
            pred_s, at_s, dirs_s = g.do_bfs(r, alpha=alpha, beta=beta)
            for nt in (1, 4):
                pred, at, dirs = g.do_bfs_omp(r, nt, alpha=alpha, beta=beta)
                assert at.tolist() == at_s.tolist() and dirs.tolist() == dirs_s.tolist()
                depth = lambda p: np.where(p < 0, -1, (p.astype(np.uint64) >> np.uint64(48)).astype(np.int64))  # noqa: E731
                assert np.array_equal(depth(pred), depth(pred_s))
                ok, te, errs = g.validate(r, pred)
                assert ok, errs


# ------------------------------------------------------- DO schedule
def test_do_schedule_hand_traced():
    # m = 1000 edges, alpha 10 -> switch to BU when a growing TD level has
    # nq_edges > 100; n_c = 448, beta 14, ef 16 -> back to TD when a
    # shrinking BU level has nq_size < 448/(14*16*2) = 1.
    nq = [0, 5, 50, 200, 100, 0]
    ne = [0, 40, 150, 999, 999, 0]
    at, dirs = O.do_schedule(1000, 448, nq, ne)
    assert dirs.tolist() == [1, 1, 0, 0, 0]
    assert at[2] == 2 and at[0] == 150 and at[1] == 1000  # T2B at level 2
    assert at[6] == 40  # last non-switching growing TD level
    assert at[8] == 5 and at[5] == -1
    # with beta large enough the shrinking BU level switches back
    at, dirs = O.do_schedule(1000, 448 * 200, nq, ne)
    assert dirs.tolist() == [1, 1, 0, 0, 1]
    assert at[5] == 4 and at[3] == 100 and at[4] == 448 * 200
    # a growing BU level never switches; a shrinking TD level never switches
    at, dirs = O.do_schedule(1000, 10 ** 9, [0, 5, 3, 200, 0], [0, 40, 30, 999, 0])
    assert dirs.tolist() == [1, 1, 1, 0]


# ------------------------------------------------------------ validator
def _tree(scale=12):
    v0, v1 = O.generate_edges(scale)
    g = O.Graph(v0, v1)
    r = int(g.find_roots(1)[0])
    lev, pred = g.bfs(r, want_pred=True)
    return g, r, lev, pred


def _parent(p):
    return int((np.int64(p) << np.int64(16)) >> np.int64(16))


def _entry(parent, depth):
    u = (int(depth) << 48) | (int(parent) & 0xFFFFFFFFFFFF)
    return np.int64(u - (1 << 64) if u >= 1 << 63 else u)


def test_validator_accepts_valid_tree():
    g, r, lev, pred = _tree()
    ok, te, errs = g.validate(r, pred)
    assert ok and sum(errs.values()) == 0
    assert te == 65535  # pf_nedge[12]


def _nbrs(g):
    v0, v1 = g.v0, g.v1
    nb = {}
    for a, b in zip(v0.tolist(), v1.tolist()):
        if a != b:
            nb.setdefault(a, set()).add(b)
            nb.setdefault(b, set()).add(a)
    return nb


def test_validator_rejects_planted_mistakes():
    g, r, lev, pred = _tree()
    nb = _nbrs(g)
    reached = np.nonzero(lev > 0)[0]
    parents = {_parent(p) for p in pred[lev > 0]}
    leaves = [int(v) for v in reached if int(v) not in parents]

    def check(p, expect, corebfs=False):
        ok, te, errs = g.validate(r, p, corebfs=corebfs)
        assert not ok
        assert errs[expect] > 0, (expect, errs)

    v = leaves[0]
    p = pred.copy(); p[v] = _entry(g.info.max_used_vertex + 7, lev[v]); check(p, "range")
    p = pred.copy(); p[r] = _entry(v, 0); check(p, "root")
    p = pred.copy(); p[v] = _entry(v, lev[v]); check(p, "self_parent")
    p = pred.copy(); p[v] = _entry(_parent(pred[v]), lev[v] + 1); check(p, "depth_map")
    unreached = np.nonzero(lev < 0)[0]
    if len(unreached):
        u = int(unreached[0])
        p = pred.copy(); p[u] = _entry(r, 0xFFFF); check(p, "depth_map")
    p = pred.copy(); p[v] = -1; check(p, "edge_tree")
    # depth gap: leaf re-parented to a same-depth neighbour
    done_gap = False
    for v2 in leaves:
        same = [w for w in nb.get(v2, ()) if lev[w] == lev[v2]]
        if same:
            p = pred.copy(); p[v2] = _entry(same[0], lev[v2] + 1); check(p, "edge_gap")
            done_gap = True
            break
    assert done_gap
    # parent that is not a neighbour but sits at the right depth
    d = lev[v]
    far = [int(w) for w in np.nonzero(lev == d - 1)[0] if int(w) not in nb.get(v, ())]
    p = pred.copy(); p[v] = _entry(far[0], d); check(p, "no_parent_edge")
    # CoreBFS mode keeps explicit depths: a wrong one trips check_pred_depth
    p = pred.copy(); p[v] = _entry(_parent(pred[v]), d + 1); check(p, "pred_depth", corebfs=True)


def test_validator_corebfs_rebuilds_depths():
    g, r, lev, pred = _tree()
    p = pred.copy()
    mask = lev > 1
    p[mask] = (p[mask].astype(np.uint64) | np.uint64(0xFFFF << 48)).view(np.int64)
    ok, te, errs = O.validate(g.v0, g.v1, g.info.max_used_vertex + 1, g.info.n, r, p, True)
    assert ok, errs
    assert np.array_equal(p >> 48, pred >> 48)  # depths written back in place


# ------------------------------------------------------------------ roots
def test_find_roots_formula():
    R = _ref_or_skip()
    v0, v1 = O.generate_edges(12)
    g = O.Graph(v0, v1)
    N = 1 << g.info.log_n

    def ref_double(s1, s2, pos):
        st = O.RefMrgState()
        seed = O.make_seed(s1, s2)
        st.z1, st.z2, st.z3, st.z4, st.z5 = seed.z1, seed.z2, seed.z3, seed.z4, seed.z5
        R.mrg_skip(C.byref(st), 2, 0, 2 * pos)
        return R.mrg_get_double_orig(C.byref(st)), R.mrg_get_double_orig(C.byref(st))

    for r1 in (0, 3):
        roots, counter = [], 0
        while len(roots) < 16:
            d0, d1 = ref_double(2 + r1, 3, counter)
            root = int((d0 + d1) * N) % N
            counter += 2
            if root in roots:
                continue
            if g.has_edge(root):
                roots.append(root)
        assert g.find_roots(16, r1=r1).tolist() == roots


def test_harmonic_mean():
    t = np.array([1.0, 2.0, 4.0])
    e = np.array([10.0, 10.0, 10.0])
    hm = O.lib().orc_harmonic_mean_teps(t.ctypes.data_as(C.POINTER(C.c_double)),
                                        e.ctypes.data_as(C.POINTER(C.c_double)), 3)
    assert abs(hm - 3.0 / (0.1 + 0.2 + 0.4)) < 1e-12


# ---------------------------------------------------------------- CoreBFS
def test_corebfs_matches_reference_log_scale20(golden):
    v0, v1 = O.generate_edges(20)
    g = O.Graph(v0, v1)
    cb = g.corebfs()
    ref = golden["corebfs_scale20"]
    assert cb["core_vertices"] == ref["core_vertices"]
    assert cb["core_edges"] == ref["core_edges"]
    assert cb["removed"] == ref["removed_parents"]
    assert cb["hd_root"] == ref["highest_degree_vertex"]


def _core_has_edge(g, cb, v):
    return bool(cb["in_core"][v]) or bool(cb["has_parent"][v])


@pytest.mark.parametrize("scale", [12, 16])
def test_corebfs_levels_equal_plain_bfs(scale, golden):
    v0, v1 = O.generate_edges(scale)
    g = O.Graph(v0, v1)
    cb = g.corebfs()
    roots = [r for r in g.find_roots(8)]
    for r in roots:
        pred, lv = g.corebfs_run(cb, int(r))
        # tree vertices carry depth 0xFFFF until the validator rebuilds them
        ok, te, errs = O.validate(g.v0, g.v1, g.info.max_used_vertex + 1, g.info.n, int(r),
                                  pred, True)
        assert ok, errs
        assert te == golden["pf_nedge"][scale]
        d = (pred.astype(np.uint64) >> np.uint64(48)).astype(np.int64)
        d[d == 0xFFFF] = -1
        assert np.array_equal(d, g.bfs(int(r)))


def test_corebfs_small_forest_by_hand():
    # cycle 0-1-2-0 with a path 2-3-4 and a leaf 1-5; separate edge 7-8;
    # separate star 10-{11,12,13}
    e = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (1, 5), (7, 8), (10, 11), (10, 12), (10, 13)]
    v0 = np.array([a for a, b in e], np.int64)
    v1 = np.array([b for a, b in e], np.int64)
    g = O.Graph(v0, v1)
    cb = g.corebfs()
    tp, hp, ic = cb["tree_parent"], cb["has_parent"], cb["in_core"]
    assert ic[[0, 1, 2]].all() and not ic[[3, 4, 5, 7, 8, 10, 11]].any()
    assert tp[4] == 3 and tp[3] == 2 and tp[5] == 1          # forest toward the core
    assert hp[7] and hp[8] and tp[7] == -1 and tp[8] == -1   # two-vertex component reset
    assert hp[11] and tp[11] == -1 and not hp[10]            # star: leaves reset, centre none
    assert cb["removed"] == 5 and cb["iterations"] == 3
    pred, lv = g.corebfs_run(cb, 4)                          # root deep in a tree
    assert lv == 2                                           # 4 -> 3 -> 2 (core)
    assert (pred[4] & 0xFFFFFFFFFFFF) == 4 and (pred[4] >> 48) == 0
    assert (pred[3] & 0xFFFFFFFFFFFF) == 4 and (pred[3] >> 48) == 1
    assert (pred[2] & 0xFFFFFFFFFFFF) == 3 and (pred[2] >> 48) == 2
    assert (pred[5] >> 48) & 0xFFFF == 0xFFFF and (pred[5] & 0xFFFFFFFFFFFF) == 1
    ok, te, errs = O.validate(v0, v1, 14, g.info.n, 4, pred, True)
    assert ok, errs
    assert (pred[5] >> 48) == 4 and (pred[0] >> 48) == 3     # rebuilt depths


# ------------------------------------- the reference driver's own -A log
def _ref_traces():
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "ref_autotune_traces.json")) as f:
        return json.load(f)["cases"]


def _roots_by_reference_rule(g, nroots, scale):
    """find_roots, redrawn with the next seed while any root reaches <= 3
    levels (mpi/main.cc:334-359)."""
    r1 = 0
    while True:
        roots = g.find_roots(nroots, r1=r1)
        if scale <= 6 or all(g.do_bfs(int(r))[1][8] > 3 for r in roots):
            return roots
        r1 += 1


@pytest.mark.parametrize("scale", [16, 18])
def test_schedule_matches_reference_driver_log(scale, golden):
    """Per root index, the auto_tuning_data the reference's own driver printed
    with -A (ref_runnable, built from mpi/main.cc): the T2B fields at alpha 10
    and the B2T fields at the alpha it then settled on.  This pins the root
    list (drawing, -ffast-math formula and redraw rule) and the direction
    rule for every root, not only the first."""
    case = next(c for c in _ref_traces() if c["scale"] == scale and not c["corebfs"])
    v0, v1 = O.generate_edges(scale)
    g = O.Graph(v0, v1)
    roots = _roots_by_reference_rule(g, case["nroots"], scale)
    assert len(roots) == case["nroots"]
    tuned = float(case["tuned_alpha"])
    for (i, t2b, nqe, m, pre), (j, b2t, nqs, nc, ef, pres) in zip(case["alpha_block"],
                                                                 case["beta_block"]):
        assert i == j
        _, at, _ = g.do_bfs(int(roots[i]), ef=16, alpha=10.0, beta=14.0)
        assert [at[2], at[0], at[1], at[6]] == [t2b, nqe, m, pre], (scale, i)
        _, at, _ = g.do_bfs(int(roots[i]), ef=16, alpha=tuned, beta=14.0)
        assert [at[5], at[3], at[4], at[7]] == [b2t, nqs, nc, pres], (scale, i)
    assert case["traversed_edges"] == [golden["pf_nedge"][scale]] * case["nroots"]
