"""Chunked upload of an input edge list (g500_graph_create +
g500_graph_append_edges, the path BfsOnMI355X::construct takes for the
reference's EdgeListStorage chunks) against the one-call upload, and the
library's refusal to free a graph under a live engine."""
import numpy as np
import pytest

import g500bfs
import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_chunked_upload_builds_the_same_csr(golden):
    print("append: SCALE 13 in uneven chunks", flush=True)
    v0, v1 = O.generate_edges(13)
    whole = g500bfs.Graph.from_edges(v0, v1).construct()
    g = g500bfs.Graph.create(len(v0) + 100)  # spare capacity is fine
    cuts = [0, 1, 1000, 1000, 77777, len(v0)]
    for a, b in zip(cuts[:-1], cuts[1:]):
        g.append_edges(v0[a:b], v1[a:b])
    g.construct()
    a, b = whole.csr(), g.csr()
    for k in ("row_off", "col", "deg_new", "new2orig", "orig2new"):
        assert np.array_equal(a[k], b[k]), k
    assert g.info.nedges_input == len(v0)
    ea = g.edges()
    assert np.array_equal(ea[0], v0) and np.array_equal(ea[1], v1)
    bfs = g500bfs.Bfs(g)
    r = int(g.find_roots(1)[0])
    bfs.run(r)
    ok, te, _ = bfs.validate(r)
    assert ok and te == golden["pf_nedge"][13]
    bfs.destroy()
    g.free()
    whole.free()


def test_append_limits():
    g = g500bfs.Graph.create(10)
    g.append_edges([1, 2, 3], [4, 5, 6])
    with pytest.raises(g500bfs.G500Error, match="capacity"):
        g.append_edges(np.arange(8), np.arange(8))
    with pytest.raises(g500bfs.G500Error, match="outside"):
        g.append_edges([-1], [2])
    assert g.info.nedges_input == 3
    g.construct()
    with pytest.raises(g500bfs.G500Error, match="constructed"):
        g.append_edges([1], [2])
    g.free()


def test_graph_free_waits_for_its_engines():
    L = g500bfs.load()
    g = g500bfs.Graph.generate(10).construct()
    b = g500bfs.Bfs(g)
    mb = g500bfs.MultiBfs(g, max_roots=8)
    # the library refuses while engines hold the graph
    assert L.g500_graph_free(g.h) == -4
    assert b"engine" in L.g500_last_error()
    mb.destroy()
    assert L.g500_graph_free(g.h) == -4
    # the binding destroys the remaining engine first
    g.free()
    assert b.h is None and g.h is None
