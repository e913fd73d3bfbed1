"""GPU parity: the HIP path (through the C ABI) against the oracle.

Integer work, so every comparison is bit-exact: generated edges, the CSR,
per-root BFS levels, roots, the direction schedule and the validator's
verdict.  Parents may legally differ from the oracle's, so parent trees are
checked with both validators instead.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import g500bfs
    g500bfs.load()
    return g500bfs


def _oracle_levels(og, root):
    lev = og.bfs(int(root))
    return np.where(lev < 0, 0xFFFF, lev).astype(np.uint16)


@pytest.mark.parametrize("scale", [1, 6, 12, 16])
def test_generator_bit_exact(G, scale):
    print(f"\n[gen] scale={scale}", flush=True)
    g = G.Graph.generate(scale)
    d0, d1 = g.edges()
    o0, o1 = O.generate_edges(scale)
    assert np.array_equal(d0, o0) and np.array_equal(d1, o1)


@pytest.mark.parametrize("scale", [6, 12, 16])
def test_csr_bit_exact(G, scale):
    print(f"\n[csr] scale={scale}", flush=True)
    g = G.Graph.generate(scale).construct()
    o0, o1 = O.generate_edges(scale)
    og = O.Graph(o0, o1)
    inf = g.info
    assert (inf.log_n, inf.n, inf.n_c, inf.m, inf.m_sym, inf.max_used_vertex) == (
        og.info.log_n, og.info.n, og.info.n_c, og.info.m, og.info.m_sym,
        og.info.max_used_vertex)
    a, b = g.csr(), og.arrays()
    for k in ("row_off", "col", "deg_new", "new2orig", "orig2new"):
        assert np.array_equal(a[k].astype(np.int64), b[k].astype(np.int64)), k


def _check_roots(G, g, og, roots, golden_te=None, alpha=10.0, beta=14.0):
    bfs = G.Bfs(g)
    for r in roots:
        r = int(r)
        at = bfs.run(r, alpha=alpha, beta=beta)
        lev = bfs.levels()
        assert np.array_equal(lev, _oracle_levels(og, r)), f"levels differ for root {r}"
        ok, te, errs = bfs.validate(r)
        assert ok, (r, errs)
        pred = bfs.pred()
        ok2, te2, errs2 = og.validate(r, pred)
        assert ok2, (r, errs2)
        assert te == te2
        if golden_te is not None:
            assert te == golden_te
        # same direction schedule and auto-tuning slots as the reference loop
        _, at_o, dirs_o = og.do_bfs(r, alpha=alpha, beta=beta)
        st, trace = bfs.stats()
        assert [t["top_down"] for t in trace] == dirs_o.tolist()
        assert at.tolist() == at_o.tolist()
    return bfs


def test_bfs_scale16_reference_roots(G, golden):
    print("\n[bfs] scale=16 roots=16", flush=True)
    g = G.Graph.generate(16).construct()
    o0, o1 = O.generate_edges(16)
    og = O.Graph(o0, o1)
    roots = g.find_roots(16)
    assert roots.tolist() == og.find_roots(16).tolist()
    _check_roots(G, g, og, roots, golden_te=golden["pf_nedge"][16])


def test_bfs_scale20_te(G, golden):
    print("\n[bfs] scale=20 roots=4", flush=True)
    g = G.Graph.generate(20).construct()
    inf = g.info
    ref = golden["edge_counts"]["20"]
    assert (inf.m_sym, inf.m, inf.n_c) == (ref["sym"], ref["dedup"], ref["n_c"])
    o0, o1 = O.generate_edges(20)
    og = O.Graph(o0, o1)
    roots = g.find_roots(4)
    _check_roots(G, g, og, roots, golden_te=golden["pf_nedge"][20])


def test_bfs_scale24_drop_in(G, golden):
    """The SCALE 24 drop-in case: the reference's logged edge counts, 64 roots
    by the reference's rule, every parent tree valid with the reference's
    traversed-edge count, and for four roots levels bit-exact against the
    oracle's textbook BFS over the same CSR (the CSR build itself is compared
    with the oracle's at SCALE 16)."""
    print("\n[bfs] scale=24 roots=64", flush=True)
    g = G.Graph.generate(24).construct()
    inf = g.info
    ref = golden["edge_counts"]["24"]
    assert (inf.m_sym, inf.m) == (ref["sym"], ref["dedup"])
    roots = g.find_roots(64)
    assert len(roots) == 64
    bfs = G.Bfs(g)
    for r in roots.tolist():
        bfs.run(r)
        ok, te, errs = bfs.validate(r)
        assert ok and te == golden["pf_nedge"][24], (r, errs, te)
    og = O.Graph.wrap(inf.log_n, inf.n_c, inf.m, g.csr())
    for r in roots[:4].tolist():
        bfs.run(r)
        assert np.array_equal(bfs.levels(), _oracle_levels(og, r)), f"levels differ for root {r}"


@pytest.mark.parametrize("alpha,beta", [(1e9, 14.0), (1e-3, 14.0), (10.0, 1e-6), (0.5, 1e6)])
def test_direction_extremes(G, alpha, beta):
    """Pure top-down, immediate bottom-up, never-return and eager-return."""
    print(f"\n[dir] alpha={alpha} beta={beta}", flush=True)
    g = G.Graph.generate(14).construct()
    o0, o1 = O.generate_edges(14)
    og = O.Graph(o0, o1)
    _check_roots(G, g, og, g.find_roots(3), alpha=alpha, beta=beta)


def test_top_down_frontier_sizes(G):
    """Pure top-down at SCALE 18: frontiers of at most 512 vertices (staged
    by the expansion itself) and larger ones (block prefix sums, a last-block
    scan of up to 32 and of more than 32 block totals, then tile owners) all
    occur."""
    print("\n[dir] scale=18 pure top-down", flush=True)
    g = G.Graph.generate(18).construct()
    o0, o1 = O.generate_edges(18)
    og = O.Graph(o0, o1)
    bfs = _check_roots(G, g, og, g.find_roots(2), alpha=1e9, beta=14.0)
    sizes = [t["nq_size"] for t in bfs.stats()[1]]
    assert max(sizes) > 65536 and any(512 < s <= 65536 for s in sizes)


def _small_cases():
    rng = np.random.default_rng(5)
    path = np.arange(9)
    star_c = np.zeros(30, np.int64)
    two = (np.array([0, 1, 2, 10, 11]), np.array([1, 2, 0, 11, 12]))
    dup = (np.array([3, 3, 3, 4, 4, 5]), np.array([4, 4, 4, 3, 5, 5]))
    rnd = rng.integers(0, 300, size=(2, 900))
    return {
        "path": (path[:-1], path[1:]),
        "star": (star_c, np.arange(1, 31)),
        "two_components": two,
        "dups_and_loops": dup,
        "random_sparse": (rnd[0], rnd[1]),
        "single_edge": (np.array([7]), np.array([2])),
    }


@pytest.mark.parametrize("name", list(_small_cases()))
def test_small_graphs(G, name):
    print(f"\n[small] {name}", flush=True)
    v0, v1 = _small_cases()[name]
    v0 = np.asarray(v0, np.int64)
    v1 = np.asarray(v1, np.int64)
    g = G.Graph.from_edges(v0, v1).construct()
    og = O.Graph(v0, v1)
    assert (g.info.n, g.info.n_c, g.info.m) == (og.info.n, og.info.n_c, og.info.m)
    a, b = g.csr(), og.arrays()
    for k in ("row_off", "col", "new2orig"):
        assert np.array_equal(a[k].astype(np.int64), b[k].astype(np.int64)), k
    roots = [v for v in range(og.info.n) if og.has_edge(v)][:5]
    _check_roots(G, g, og, roots)


def test_degenerate_inputs(G):
    print("\n[degenerate]", flush=True)
    # empty edge list and an all-self-loop list build an edgeless graph
    for v0, v1 in ((np.zeros(0, np.int64), np.zeros(0, np.int64)),
                   (np.array([4, 4, 9]), np.array([4, 4, 9]))):
        g = G.Graph.from_edges(v0, v1).construct()
        assert g.info.n_c == 0 and g.info.m == 0
        assert g.find_roots(4).tolist() == O.Graph(v0, v1).find_roots(4).tolist()
        bfs = G.Bfs(g)
        bfs.run(0)
        assert bfs.pred()[0] == 0  # root is its own parent, nothing else
    with pytest.raises(G.G500Error):
        G.Graph.from_edges(np.array([-1]), np.array([2]))
    g = G.Graph.generate(8).construct()
    with pytest.raises(G.G500Error):
        G.Bfs(g).run(1 << 20)


def test_validator_agrees_on_planted_mistakes(G):
    print("\n[validator] planted", flush=True)
    g = G.Graph.generate(12).construct()
    o0, o1 = O.generate_edges(12)
    og = O.Graph(o0, o1)
    r = int(g.find_roots(1)[0])
    bfs = G.Bfs(g)
    bfs.run(r)
    pred = bfs.pred()
    lev = og.bfs(r)
    reached = np.nonzero(lev > 0)[0]
    parents = {int((p << 16) >> 16) for p in pred[lev > 0]}
    v = int([x for x in reached if int(x) not in parents][0])

    def ent(parent, depth):
        u = (int(depth) << 48) | (int(parent) & 0xFFFFFFFFFFFF)
        return np.int64(u - (1 << 64) if u >= 1 << 63 else u)

    plants = []
    p = pred.copy(); p[v] = ent(og.info.max_used_vertex + 3, lev[v]); plants.append(p)
    p = pred.copy(); p[r] = ent(v, 0); plants.append(p)
    p = pred.copy(); p[v] = ent(v, lev[v]); plants.append(p)
    p = pred.copy(); p[v] = ent((int(pred[v]) << 16) >> 16, lev[v] + 1); plants.append(p)
    p = pred.copy(); p[v] = -1; plants.append(p)
    far = [int(w) for w in np.nonzero(lev == lev[v] - 1)[0]]
    nbrs = set(o1[o0 == v].tolist()) | set(o0[o1 == v].tolist())
    far = [w for w in far if w not in nbrs]
    p = pred.copy(); p[v] = ent(far[0], lev[v]); plants.append(p)
    for p in plants:
        ok_d, te_d, e_d = bfs.validate(r, host_pred=p)
        ok_o, te_o, e_o = og.validate(r, p)
        assert not ok_d and not ok_o
        first = lambda e: next(k for k in G.VERR_NAMES if e[k])  # noqa: E731
        assert first(e_d) == first(e_o), (e_d, e_o)
    ok, te, _ = bfs.validate(r, host_pred=pred)
    assert ok and te == 65535
    # a run after uploads of other parent arrays writes every entry afresh
    bfs.validate(r, host_pred=np.full_like(pred, 5))
    r2 = int(g.find_roots(2)[1])
    bfs.run(r2)
    ok, te, errs = bfs.validate(r2)
    assert ok and te == 65535, errs
    assert og.validate(r2, bfs.pred())[0]


def test_runs_alternating_between_components(G):
    """One engine runs roots of a small component and of the giant one in
    turn.  The giant-component runs switch back to top-down, so their pred
    gather starts early and covers only the vertices inside its snapshot;
    the fixup must then reset the entries the small-component run wrote
    (their vertices are outside the snapshot and unreached), and the next
    small-component run must clear what the giant run wrote."""
    print("\n[components] scale=16", flush=True)
    g = G.Graph.generate(16).construct()
    o0, o1 = O.generate_edges(16)
    og = O.Graph(o0, o1)
    big = int(g.find_roots(1)[0])
    lev_big = og.bfs(big)
    small = [v for v in range(og.info.n) if og.has_edge(v) and lev_big[v] < 0]
    assert small, "the graph has vertices with edges outside the giant component"
    bfs = G.Bfs(g)
    for r in (small[0], big, small[-1], big, small[0]):
        bfs.run(r)
        assert np.array_equal(bfs.levels(), _oracle_levels(og, r)), r
        pred = bfs.pred()
        ok, _, errs = og.validate(r, pred)
        assert ok, (r, errs)
        if r == big:
            # a top-down level after a bottom-up one: the early gather ran
            dirs = [t["top_down"] for t in bfs.stats()[1]]
            assert 0 in dirs and 1 in dirs[dirs.index(0):], dirs
            assert all(pred[v] == -1 for v in small), r


def test_roots_without_edges(G):
    """Roots that have no edges reach only themselves; the next run, from
    any root, leaves no trace of them in the parent array."""
    print("\n[isolated roots]", flush=True)
    v0, v1 = _small_cases()["two_components"]
    v0 = np.asarray(v0, np.int64)
    v1 = np.asarray(v1, np.int64)
    g = G.Graph.from_edges(v0, v1).construct()
    og = O.Graph(v0, v1)
    bfs = G.Bfs(g)
    for r in (5, 7, 7, 0, 6, 11, 2):
        bfs.run(r)
        assert np.array_equal(bfs.levels(), _oracle_levels(og, r)), r
        pred = bfs.pred()
        ok, _, errs = og.validate(r, pred)
        assert ok, (r, errs)
        assert bfs.validate(r)[0], r
        lone = [v for v in range(og.info.n) if not og.has_edge(v)]
        assert all(pred[v] == (r if v == r else -1) for v in lone), (r, pred.tolist())


@pytest.mark.parametrize("exec_alpha", [0.0, 1e9, None])
def test_execution_mode_keeps_reference_schedule(G, exec_alpha, golden):
    """Top-down levels run as top-down (0), as bottom-up sweeps (1e9) or as
    the default decides (over m/25 edges: bottom-up): the levels, the
    direction schedule and auto_tuning_data stay identical."""
    print(f"\n[exec] alpha_exec={exec_alpha}", flush=True)
    g = G.Graph.generate(16).construct()
    o0, o1 = O.generate_edges(16)
    og = O.Graph(o0, o1)
    bfs = G.Bfs(g)
    if exec_alpha is not None:
        bfs.set_exec_alpha(exec_alpha)
    for r in g.find_roots(6):
        r = int(r)
        at = bfs.run(r)
        assert np.array_equal(bfs.levels(), _oracle_levels(og, r))
        ok, te, errs = bfs.validate(r)
        assert ok and te == golden["pf_nedge"][16], errs
        _, at_o, dirs_o = og.do_bfs(r)
        st, trace = bfs.stats()
        assert [t["top_down"] for t in trace] == dirs_o.tolist()
        assert at.tolist() == at_o.tolist()
        if exec_alpha == 0.0:
            assert all(t["exec_bottom_up"] == (0 if t["top_down"] else 1) for t in trace)
        elif exec_alpha is None:
            m = g.info.m
            for prev, t in zip(trace, trace[1:]):
                if t["top_down"]:
                    heavy = prev["top_down"] and prev["nq_edges"] > m // 25
                    assert t["exec_bottom_up"] == int(heavy), (r, t)
        else:
            assert all(t["exec_bottom_up"] == 1 for t in trace[1:])


@pytest.mark.parametrize("depth,guess,corebfs", [("0", "", False), ("1", "wrong", False),
                                                 ("2", "wrong", False), ("3", "", False),
                                                 ("1", "wrong", True)])
def test_step_guesses_never_change_results(G, depth, guess, corebfs, golden, monkeypatch):
    """The host queues guessed steps (and, on guesses, the pred gather behind
    a conversion and the fixup behind the last step) before the step ahead
    has reported; every kernel checks the loop state.  With no guesses, with
    deeper speculation, and with every guess turned into a wrong one of
    another kind (G500_GUESS=wrong: a bottom-up level guessed to be followed
    by the conversion, a top-down level by the end of the run, ...), the
    levels, the direction schedule, auto_tuning_data and the validated tree
    stay those of the oracle; wrong guesses are counted.  With CoreBFS (the
    gather that also writes the forest's entries) the levels after the depth
    rebuild and the validated tree are checked."""
    print(f"\n[guess] depth={depth} guess={guess!r} corebfs={corebfs}", flush=True)
    monkeypatch.setenv("G500_SPECULATE", depth)
    monkeypatch.setenv("G500_GUESS", guess)
    g = G.Graph.generate(16).construct(corebfs=corebfs)
    o0, o1 = O.generate_edges(16)
    og = O.Graph(o0, o1)
    bfs = G.Bfs(g)  # reads both variables
    wrong = 0
    for r in g.find_roots(8):
        r = int(r)
        at = bfs.run(r)
        ok, te, errs = bfs.validate(r)
        assert ok and te == golden["pf_nedge"][16], (r, errs)
        st, trace = bfs.stats()
        wrong += st.mispredicted
        if corebfs:
            lev = bfs.levels().astype(np.int64)
            lev[lev == 0xFFFF] = -1
            assert np.array_equal(lev, og.bfs(r)), r
            continue
        assert np.array_equal(bfs.levels(), _oracle_levels(og, r)), r
        _, at_o, dirs_o = og.do_bfs(r)
        assert [t["top_down"] for t in trace] == dirs_o.tolist(), r
        assert at.tolist() == at_o.tolist(), r
    if depth == "0":
        assert wrong == 0
    if guess == "wrong":
        assert wrong >= 8  # at least one per run
    bfs.destroy()
    g.free()


def _ref_case(scale, corebfs):
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "ref_autotune_traces.json")) as f:
        return next(c for c in json.load(f)["cases"]
                    if c["scale"] == scale and c["corebfs"] == corebfs)


@pytest.mark.parametrize("scale,corebfs", [(16, False), (18, False), (20, False),
                                           (16, True), (18, True)])
def test_auto_tuning_data_matches_reference_driver_log(G, scale, corebfs, golden):
    """Every root's auto_tuning_data against what the reference's own CPU
    driver printed with -A (mpi/main.cc:60-78, built from its sources): roots
    drawn by the reference rule, the T2B fields at alpha 10 and the B2T fields
    at the alpha the reference's tuner settled on; with CoreBFS the core BFS's
    fields.  Each run's tree validates with the reference's TE."""
    print(f"\n[reference -A log] scale={scale} corebfs={corebfs}", flush=True)
    case = _ref_case(scale, corebfs)
    g = G.Graph.generate(scale).construct(corebfs=corebfs)
    bfs = G.Bfs(g)
    r1 = 0
    while True:
        roots = g.find_roots(case["nroots"], r1=r1)
        if all(bfs.run(int(r))[8] > 3 for r in roots):
            break
        r1 += 1
    tuned = float(case["tuned_alpha"])
    for (i, t2b, nqe, m, pre), (j, b2t, nqs, nc, ef, pres) in zip(case["alpha_block"],
                                                                 case["beta_block"]):
        r = int(roots[i])
        at = bfs.run(r, alpha=10.0, beta=14.0)
        assert [at[2], at[0], at[1], at[6]] == [t2b, nqe, m, pre], (i, r)
        at = bfs.run(r, alpha=tuned, beta=14.0)
        assert [at[5], at[3], at[4], at[7]] == [b2t, nqs, nc, pres], (i, r)
        ok, te, errs = bfs.validate(r)
        assert ok and te == case["traversed_edges"][i] == golden["pf_nedge"][scale], (i, errs)
    bfs.destroy()
    g.free()
