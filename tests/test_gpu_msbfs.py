"""GPU parity of the batched bit-parallel BFS (g500_msbfs_*).

Each root's uint8 level array must equal, bit for bit, the levels of an
independent BFS from that root: the oracle's textbook BFS on small graphs,
the plain GPU engine on larger ones.  Traversed-edge counts must equal the
validator's TE for the same root.  Covered: 1..256 roots (ragged lane words,
duplicate roots, roots without edges), every degree-bin layout, graph replay
against eager launches, every direction mode (top-down, streaming pull,
hub pull + push, per-vertex pull) forced on every level, the split level (light
roots push while the others pull) at every push threshold, the asynchronous
launch/query/wait form, BFS depths beyond the captured level count, and the
uint8 depth limit.
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


def _u8(lev):
    return np.where(lev < 0, 0xFF, lev).astype(np.uint8)


def _oracle_te(og, root):
    lev, pred = og.bfs(int(root), want_pred=True)
    ok, te, _ = og.validate(int(root), pred)
    assert ok
    return te


@pytest.fixture(scope="module")
def s12(G):
    o0, o1 = O.generate_edges(12)
    og = O.Graph(o0, o1)
    g = G.Graph.generate(12).construct()
    return g, og


def _mixed_roots(g, og, count, seed):
    """Random roots over all vertices: isolated ones, duplicates, hubs."""
    rng = np.random.default_rng(seed)
    n = g.info.n
    roots = rng.integers(0, n, size=count)
    if count >= 4:
        roots[1] = roots[0]                       # duplicate
        iso = [v for v in range(n) if not og.has_edge(v)]
        if iso:
            roots[2] = iso[seed % len(iso)]       # no edges
        roots[3] = int(g.csr()["new2orig"][0])    # highest-degree vertex
    return roots


@pytest.mark.parametrize("count", [1, 5, 63, 64, 65, 130, 200, 256])
def test_levels_equal_oracle_scale12(G, s12, count):
    print(f"\n[msbfs] scale=12 roots={count}", flush=True)
    g, og = s12
    roots = _mixed_roots(g, og, count, seed=count)
    mb = G.MultiBfs(g, max_roots=256)
    mb.run(roots)
    te = mb.te()
    for i, r in enumerate(roots):
        exp = _u8(og.bfs(int(r)))
        got = mb.levels(i)
        assert np.array_equal(got, exp), f"root index {i} ({r})"
    for i in range(min(count, 6)):
        assert te[i] == _oracle_te(og, roots[i]), f"TE of root index {i}"
    st = mb.stats()
    assert st["nroots"] == count
    assert st["discovered"] == sum(int((mb.levels(i) != 0xFF).sum()) for i in range(count)) - \
        sum(1 for r in roots if not og.has_edge(int(r)))
    mb.destroy()


def test_scale16_256_roots_equal_plain_engine(G, golden):
    print("\n[msbfs] scale=16 roots=256 vs plain engine", flush=True)
    g = G.Graph.generate(16).construct()
    roots = g.find_roots(256)
    assert len(roots) == 256
    mb = G.MultiBfs(g)
    mb.run(roots)
    te = mb.te()
    bfs = G.Bfs(g)
    for i, r in enumerate(roots):
        bfs.run(int(r))
        lev16 = bfs.levels()
        exp = np.where(lev16 == 0xFFFF, 0xFF, lev16).astype(np.uint8)
        assert np.array_equal(mb.levels(i), exp), f"root index {i}"
        if i < 8:
            ok, te_plain, _ = bfs.validate(int(r))
            assert ok and te[i] == te_plain
    giant = golden["pf_nedge"][16]
    assert (te == giant).sum() >= 250  # roots in the giant component


@pytest.mark.parametrize("dense_algo", [1, 2])
def test_batch_with_roots_in_tiny_components(G, dense_algo):
    """Roots on isolated edges finish after one level while the others sweep
    the giant component: their lanes leave the frontier, and the bottom-up
    kernels count absent lanes as covered.  Every root's levels must still
    equal the oracle's, in the pull, streaming and hybrid dense forms."""
    print(f"\n[msbfs] scale=16 tiny-component roots dense_algo={dense_algo}", flush=True)
    g = G.Graph.generate(16).construct()
    o0, o1 = O.generate_edges(16)
    og = O.Graph(o0, o1)
    a = og.arrays()
    n_c = og.info.n_c
    deg = (a["row_off"][1:n_c + 1] - a["row_off"][:n_c]).astype(np.int64)
    lone = [int(a["new2orig"][v]) for v in range(n_c)
            if deg[v] == 1 and deg[int(a["col"][a["row_off"][v]])] == 1][:3]
    assert len(lone) == 3, "SCALE 16 has six isolated edges"
    roots = np.array(list(g.find_roots(125)) + lone, np.int64)
    mb = G.MultiBfs(g)
    mb.set_tuning(dense_algo=dense_algo, stream_frac=0.2 if dense_algo == 1 else 0.0)
    mb.run(roots)
    for i, r in enumerate(roots.tolist()):
        ref = og.bfs(r)
        assert np.array_equal(mb.levels(i), np.where(ref < 0, 0xFF, ref).astype(np.uint8)), (i, r)
    mb.destroy()


@pytest.mark.parametrize("tuning", [
    dict(heavy_degree=1, medium_degree=1, medium_group=8),          # all one wave each
    dict(heavy_degree=1 << 40, medium_degree=1 << 40, medium_group=8),  # all one lane each
    dict(heavy_degree=64, medium_degree=2, medium_group=2),
    dict(heavy_degree=512, medium_degree=16, medium_group=4),
    dict(heavy_degree=256, medium_degree=32, medium_group=16),
    dict(heavy_degree=128, medium_degree=8, medium_group=32, blocks_per_cu=1),
    dict(use_graph=0),
    dict(td_alpha=1e-9),                          # every level top-down
    dict(td_alpha=1e12, stream_frac=0.0),         # every level per-vertex pull
    dict(td_alpha=1e12, stream_frac=1.0, dense_algo=1),   # every level streaming pull
    dict(td_alpha=1e12, stream_frac=1.0, dense_algo=1, heavy_degree=1, medium_degree=1),
    dict(td_alpha=1e12, stream_frac=1.0, dense_algo=2, hot_bytes=4096),  # hub pull + push
    dict(td_alpha=1e12, stream_frac=1.0, dense_algo=2, hot_bytes=0),     # push only
    dict(td_alpha=1e12, stream_frac=1.0, dense_algo=2, hot_bytes=1 << 40),  # pull only
    dict(td_alpha=1e12, stream_frac=1.0, dense_algo=2, hot_bytes=512, heavy_degree=8,
         medium_degree=4, heavy_chunk=256, medium_group=4),
    dict(td_alpha=4.0, stream_frac=0.3, dense_algo=2, hot_bytes=2048),
    dict(heavy_degree=16, medium_degree=4, heavy_chunk=256),
    dict(heavy_degree=16, medium_degree=4, heavy_chunk=65536, td_alpha=1.0),
    dict(split_bp=200),                           # light roots push at the first pull level
    dict(split_bp=10000),                         # every root pushes there
    dict(split_bp=1),
    dict(split_bp=500, heavy_degree=16, medium_degree=4, heavy_chunk=256, medium_group=4),
    dict(split_bp=10000, td_alpha=100.0, heavy_degree=1, medium_degree=1),
])
def test_bins_and_modes_do_not_change_levels(G, s12, tuning):
    print(f"\n[msbfs] tuning {tuning}", flush=True)
    g, og = s12
    roots = _mixed_roots(g, og, 100, seed=7)
    ref = G.MultiBfs(g, max_roots=128)
    ref.set_tuning(split_bp=0)  # every bottom-up level a plain per-vertex pull
    ref.run(roots)
    mb = G.MultiBfs(g, max_roots=128)
    mb.set_tuning(**tuning)
    for _ in range(2):  # a replayed graph must give the same answer again
        mb.run(roots)
        for i in range(len(roots)):
            assert np.array_equal(mb.levels(i), ref.levels(i)), (tuning, i)
        assert np.array_equal(mb.te(), ref.te())


def test_async_launch_and_smaller_second_batch(G, s12):
    print("\n[msbfs] launch/query/wait, 200 then 3 roots", flush=True)
    g, og = s12
    mb = G.MultiBfs(g)
    roots = _mixed_roots(g, og, 200, seed=3)
    mb.launch(roots)
    while not mb.done():
        pass
    mb.wait()
    for i in (0, 63, 64, 199):
        assert np.array_equal(mb.levels(i), _u8(og.bfs(int(roots[i]))))
    small = roots[[5, 6, 7]]
    mb.run(small)
    assert mb.nroots == 3
    for i in range(3):
        assert np.array_equal(mb.levels(i), _u8(og.bfs(int(small[i]))))
    with pytest.raises(G.G500Error):
        mb.levels(3)


def _path_graph(G, length):
    v = np.arange(length - 1, dtype=np.int64)
    g = G.Graph.from_edges(v, v + 1).construct()
    return g


@pytest.mark.parametrize("use_graph", [1, 0])
def test_depth_beyond_captured_levels(G, use_graph):
    """A 200-vertex path: depths up to 199 need more level buffers than the
    first capture holds; the batch finishes level by level on the host."""
    print(f"\n[msbfs] path 200 use_graph={use_graph}", flush=True)
    g = _path_graph(G, 200)
    mb = G.MultiBfs(g, max_roots=4)
    mb.set_tuning(use_graph=use_graph)
    roots = np.array([0, 199, 100, 37])
    for _ in range(2):
        mb.run(roots)
        n = g.info.n
        for i, r in enumerate(roots):
            exp = np.full(n, 0xFF, np.uint8)
            exp[:200] = np.abs(np.arange(200) - r)
            assert np.array_equal(mb.levels(i), exp), i
        assert mb.stats()["nlevels"] == 200
        assert list(mb.te()) == [199] * 4


def _host_levels(n, src, dst, root):
    adj = [[] for _ in range(n)]
    for a, b in zip(src.tolist(), dst.tolist()):
        adj[a].append(b)
        adj[b].append(a)
    lev = np.full(n, 0xFF, np.uint8)
    lev[root] = 0
    frontier = [root]
    d = 0
    while frontier:
        d += 1
        nxt = []
        for u in frontier:
            for w in adj[u]:
                if lev[w] == 0xFF:
                    lev[w] = d
                    nxt.append(w)
        frontier = nxt
    return lev


@pytest.mark.parametrize("plen", [15, 16])
@pytest.mark.parametrize("nroots", [3, 65])
def test_level_planes_at_their_depth_limit(G, plen, nroots):
    """The level arrays of one or two root words go through four bit planes
    per vertex while a batch has at most 15 levels (value 15 marks an
    unreached lane), else straight from the level frontiers: a path of 15
    or 16 vertices with 285 leaves on its second vertex (n = 512) gives 15
    or 16 levels from the path's ends."""
    print(f"\n[msbfs] planes path {plen} roots {nroots}", flush=True)
    src = np.concatenate([np.arange(plen - 1), np.ones(300 - plen, np.int64)])
    dst = np.concatenate([np.arange(1, plen), np.arange(plen, 300)])
    g = G.Graph.from_edges(src.astype(np.int64), dst.astype(np.int64)).construct()
    n = g.info.n
    assert n % 256 == 0
    mb = G.MultiBfs(g, max_roots=128)
    ends = [0, plen - 1, plen // 2, plen, 299]
    roots = np.array([ends[i % len(ends)] for i in range(nroots)], np.int64)
    mb.run(roots)
    for i in sorted({0, 1, 2, nroots - 1}):
        assert np.array_equal(mb.levels(i), _host_levels(n, src, dst, int(roots[i]))), i


def test_depth_limit_is_an_error(G):
    print("\n[msbfs] path 300: depth 299 > 254", flush=True)
    g = _path_graph(G, 300)
    mb = G.MultiBfs(g, max_roots=2)
    with pytest.raises(G.G500Error) as e:
        mb.run(np.array([0, 150]))
    assert "depth" in str(e.value)
    mb.run(np.array([150]))  # depth 150: fine on the same handle
    lev = mb.levels(0)
    assert lev[0] == 150 and lev[299] == 149 and lev[150] == 0


def test_bad_arguments(G, s12):
    g, _ = s12
    mb = G.MultiBfs(g, max_roots=8)
    with pytest.raises(G.G500Error):
        mb.run(np.arange(9))
    with pytest.raises(G.G500Error):
        mb.run(np.array([g.info.n]))
    with pytest.raises(G.G500Error):
        mb.set_tuning(medium_group=3)
    cg = G.Graph.generate(10).construct(corebfs=True)
    with pytest.raises(G.G500Error):
        G.MultiBfs(cg)


def test_two_vertex_graph(G):
    """n = 2 (not a multiple of 4): the one-vertex-per-lane output path."""
    print("\n[msbfs] two vertices", flush=True)
    g = G.Graph.from_edges(np.array([0]), np.array([1])).construct()
    assert g.info.n == 2
    mb = G.MultiBfs(g, max_roots=3)
    mb.run(np.array([0, 1, 0]))
    assert mb.levels(0).tolist() == [0, 1]
    assert mb.levels(1).tolist() == [1, 0]
    assert mb.levels(2).tolist() == [0, 1]
    assert list(mb.te()) == [1, 1, 1]


def test_gather_levels(G, s12):
    print("\n[msbfs] gather", flush=True)
    g, og = s12
    roots = _mixed_roots(g, og, 70, seed=11)
    mb = G.MultiBfs(g)
    mb.run(roots)
    rng = np.random.default_rng(5)
    ri = rng.integers(0, 70, 500)
    vs = rng.integers(-3, g.info.n + 3, 500)
    got = mb.gather(ri, vs)
    for i, r, v in zip(range(500), ri, vs):
        exp = 0xFF if v < 0 or v >= g.info.n else mb.levels(int(r))[v]
        assert got[i] == exp
    with pytest.raises(G.G500Error):
        mb.gather([70], [0])


def test_autotune_keeps_levels(G, s12):
    print("\n[msbfs] autotune", flush=True)
    g, og = s12
    roots = _mixed_roots(g, og, 64, seed=13)
    mb = G.MultiBfs(g, max_roots=64)
    best, ms = mb.autotune(roots, rounds=1)
    assert ms > 0 and best == mb.tuning()
    mb.run(roots)
    for i in (0, 2, 3, 40, 63):
        assert np.array_equal(mb.levels(i), _u8(og.bfs(int(roots[i]))))


def test_engine_without_level_arrays(G, s12):
    print("\n[msbfs] no level arrays", flush=True)
    g, og = s12
    roots = _mixed_roots(g, og, 90, seed=17)
    # the level arrays, and their bit planes (four per vertex, two lane words)
    assert G.msbfs_footprint(g, 90, level_arrays=False) + 90 * g.info.n + \
        max(g.info.n_c, 1) * 4 * 2 * 8 == G.msbfs_footprint(g, 90)
    mb = G.MultiBfs(g, max_roots=90, level_arrays=False)
    mb.run(roots)
    ref = G.MultiBfs(g, max_roots=90)
    ref.run(roots)
    assert np.array_equal(mb.te(), ref.te())
    rng = np.random.default_rng(2)
    ri = rng.integers(0, 90, 400)
    vs = rng.integers(0, g.info.n, 400)
    vs[:90] = roots  # every root looked up at itself (level 0, also without edges)
    ri[:90] = np.arange(90)
    assert np.array_equal(mb.gather(ri, vs), ref.gather(ri, vs))
    assert np.all(mb.gather(np.arange(90), roots) == 0)
    for i in (0, 2, 50):
        exp = _u8(og.bfs(int(roots[i])))
        assert np.array_equal(mb.gather(np.full(g.info.n, i), np.arange(g.info.n)), exp)
    with pytest.raises(G.G500Error):
        mb.levels(0)
