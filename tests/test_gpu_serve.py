"""GPU test of the serving loop: every answer (TE and target levels) equals
the answer of the same query run alone, whatever batch it shared, and equals
the oracle's BFS.  Covers mixed graphs, many engines in flight, a budget that
allows one engine per graph, and a budget too small for any engine."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import g500bfs
    import serve
    g500bfs.load()
    graphs, oracles, pools = [], [], []
    for i, scale in enumerate((9, 10, 11, 12)):
        s1, s2 = 2 + i, 3 + i
        g = g500bfs.Graph.generate(scale, seed1=s1, seed2=s2).construct()
        v0, v1 = O.generate_edges(scale, s1, s2)
        graphs.append(g)
        oracles.append(O.Graph(v0, v1))
        pools.append(g.find_roots(48))
    return g500bfs, serve, graphs, oracles, pools


def _oracle_answer(og, q):
    lev, pred = og.bfs(q.root, want_pred=True)
    ok, te, _ = og.validate(q.root, pred)
    assert ok
    exp = np.where(lev < 0, 0xFF, lev)
    return te, exp[q.targets].astype(np.uint8)


@pytest.mark.parametrize("max_batch,engines", [(64, 4), (256, 2), (8, 3)])
def test_mixed_stream_answers_equal_alone(setup, max_batch, engines):
    print(f"\n[serve] max_batch={max_batch} engines={engines}", flush=True)
    G, serve, graphs, oracles, pools = setup
    qs = serve.poisson_workload(graphs, pools, 500, rate_qps=2e5, seed=max_batch)
    srv = serve.Server(graphs, max_batch=max_batch, max_engines_per_graph=engines)
    done = serve.run_open_loop(srv, qs)
    assert len(srv.done) == 500 and srv.batches < 500  # batched, not one by one
    srv.close()
    alone = [serve.Query(q.qid, q.gid, q.root, q.targets) for q in done]
    serve.one_per_launch(graphs, alone)
    for q, a in zip(done, alone):
        assert q.te == a.te and np.array_equal(q.levels, a.levels), q.qid
    for q in done[:60]:
        te, lev = _oracle_answer(oracles[q.gid], q)
        assert q.te == te and np.array_equal(q.levels, lev), q.qid
    s = serve.latency_summary(done)
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_budget_for_one_engine_per_graph(setup):
    print("\n[serve] budget = one engine per graph", flush=True)
    G, serve, graphs, oracles, pools = setup
    need = sum(G.msbfs_footprint(g, 64, level_arrays=False) for g in graphs)
    srv = serve.Server(graphs, max_batch=64, hbm_budget=need)
    qs = serve.poisson_workload(graphs, pools, 300, rate_qps=1e5, seed=9)
    done = serve.run_open_loop(srv, qs)
    assert len(done) == 300
    assert all(len(es) == 1 for es in srv.engines)
    assert srv.used <= need
    for q in done[:40]:
        te, lev = _oracle_answer(oracles[q.gid], q)
        assert q.te == te and np.array_equal(q.levels, lev)
    srv.close()


def test_budget_too_small(setup):
    G, serve, graphs, oracles, pools = setup
    srv = serve.Server(graphs, max_batch=64, hbm_budget=1024)
    srv.submit(serve.Query(0, 0, pools[0][0], [0], 0.0))
    with pytest.raises(MemoryError):
        srv.poll()
