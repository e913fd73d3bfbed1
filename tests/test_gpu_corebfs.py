"""GPU parity of CoreBFS (-C): forest, core, per-root levels and pred."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import g500bfs
    g500bfs.load()
    return g500bfs


def test_corebfs_scale20_reference_log(G, golden):
    print("\n[corebfs] scale 20", flush=True)
    g = G.Graph.generate(20).construct(corebfs=True)
    ref = golden["corebfs_scale20"]
    assert g.info.core_vertices == g.info.n_c == ref["core_vertices"]
    assert g.info.m == ref["core_edges"]
    bfs = G.Bfs(g)  # prepare_bfs runs reset_unreachable
    inf = g.refresh().info
    assert inf.removed_parents == ref["removed_parents"]
    assert inf.hd_root == ref["highest_degree_vertex"]
    for r in g.find_roots(4):
        bfs.run(int(r))
        ok, te, errs = bfs.validate(int(r))
        assert ok, errs
        assert te == golden["pf_nedge"][20]


def _hand_edges():
    e = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (1, 5), (7, 8), (10, 11), (10, 12), (10, 13)]
    return (np.array([a for a, b in e], np.int64), np.array([b for a, b in e], np.int64))


@pytest.mark.parametrize("case", ["hand", "s12", "s16"])
def test_corebfs_forest_levels_and_pred(G, golden, case):
    print(f"\n[corebfs] {case}", flush=True)
    if case == "hand":
        v0, v1 = _hand_edges()
        g = G.Graph.from_edges(v0, v1).construct(corebfs=True)
    else:
        scale = int(case[1:])
        v0, v1 = O.generate_edges(scale)
        g = G.Graph.generate(scale).construct(corebfs=True)
    og = O.Graph(v0, v1)
    cb = og.corebfs()
    bfs = G.Bfs(g)
    inf = g.refresh().info
    assert (inf.core_vertices, inf.m) == (cb["core_vertices"], cb["core_edges"])
    assert inf.tree_vertices == cb["tree_vertices"] and inf.removed_parents == cb["removed"]
    assert inf.hd_root == cb["hd_root"]
    tp, hp = g.forest()
    assert np.array_equal(tp, cb["tree_parent"]) and np.array_equal(hp, cb["has_parent"])
    # has_edge = core vertex or forest vertex (mpi/bfs.hpp:167-170)
    for v in range(min(og.info.n, 4096)):
        assert g.has_edge(v) == bool(cb["in_core"][v] or cb["has_parent"][v])
    roots = g.find_roots(6) if case != "hand" else np.array([4, 0, 5, 3])
    tree = (cb["has_parent"] == 1) & (cb["in_core"] == 0)
    for r in roots:
        r = int(r)
        at = bfs.run(r)
        pred = bfs.pred()
        ref_pred, lv = og.corebfs_run(cb, r)
        # forest part and the walked path are fully determined
        assert np.array_equal(pred[tree], ref_pred[tree]), r
        assert bfs.stats()[0].root_level == lv
        assert at[8] >= lv + 1
        okd, ted, errs_d = bfs.validate(r)
        p2 = pred.copy()
        oko, teo, errs_o = O.validate(og.v0, og.v1, og.info.max_used_vertex + 1, og.info.n, r,
                                      p2, True)
        assert okd and oko, (errs_d, errs_o)
        assert ted == teo
        if case != "hand":
            assert ted == golden["pf_nedge"][int(case[1:])]
            # after the depth rebuild every level equals a plain BFS level
            lev = bfs.levels().astype(np.int64)
            lev[lev == 0xFFFF] = -1
            assert np.array_equal(lev, og.bfs(r))


def test_driver_corebfs(G, golden):
    import os
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(G.HERE), "graph500-bfs_amd", "bin", "g500_runnable")
    print("\n[driver] 16 -C", flush=True)
    p = subprocess.run([exe, "16", "-n", "8", "-C"], capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr
    body = p.stdout.split("============= Result ==============")[1]
    b = {k: float(v) for k, v in re.findall(r"^([a-z_A-Z]+):\s+(\S+)$", body, re.M)}
    assert b["min_nedge"] == b["max_nedge"] == golden["pf_nedge"][16]
