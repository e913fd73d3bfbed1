"""Host logic of the serving loop (no GPU): continuous batching, batch size
limit, free-memory admission with a reservation for graphs that have no
engine yet, and latency bookkeeping, driven through a stand-in engine whose
batches finish after a set number of polls."""
import numpy as np
import pytest

import serve


class FakeGraph:
    class info:
        n = 1 << 10
        m = 1000


class FakeEngine:
    launches = []

    def __init__(self, graph, max_roots, polls=2):
        self.graph, self.max_roots, self.polls = graph, max_roots, polls
        self.left = 0
        self.nroots = 0

    def launch(self, roots):
        assert 1 <= len(roots) <= self.max_roots
        FakeEngine.launches.append((id(self), len(roots)))
        self.roots = np.asarray(roots)
        self.nroots = len(roots)
        self.left = self.polls

    def done(self):
        self.left -= 1
        return self.left <= 0

    def wait(self):
        pass

    def te(self):
        return self.roots * 10  # a function of the root alone

    def gather(self, idx, vertices):
        idx = np.asarray(idx, np.int64)
        return ((self.roots[idx] + np.asarray(vertices)) % 251).astype(np.uint8)

    def destroy(self):
        pass


def make_server(ngraphs=2, max_batch=4, budget=100, footprint=10, engines=3):
    FakeEngine.launches = []
    graphs = [FakeGraph() for _ in range(ngraphs)]
    return serve.Server(graphs, max_batch=max_batch, hbm_budget=budget,
                        max_engines_per_graph=engines,
                        engine_factory=lambda g, r: FakeEngine(g, r),
                        footprint=[footprint] * ngraphs)


def test_batches_respect_size_and_answers_follow_the_query():
    srv = make_server(ngraphs=1, max_batch=4)
    qs = [serve.Query(i, 0, 100 + i, [i, i + 1], 0.0) for i in range(10)]
    for q in qs:
        srv.submit(q)
    srv.drain(clock=lambda: 1.0)
    assert len(srv.done) == 10
    assert all(n <= 4 for _, n in FakeEngine.launches)
    assert sum(n for _, n in FakeEngine.launches) == 10
    for q in qs:
        assert q.te == q.root * 10
        assert q.levels.tolist() == [(q.root + t) % 251 for t in q.targets]


def test_continuous_batching_uses_more_engines_while_others_run():
    srv = make_server(ngraphs=1, max_batch=2, budget=30, footprint=10, engines=3)
    for i in range(6):
        srv.submit(serve.Query(i, 0, i, [0], 0.0))
    srv.poll(clock=lambda: 0.0)   # three engines start, each with two queries
    assert len(srv.engines[0]) == 3
    assert sum(e.batch is not None for e in srv.engines[0]) == 3
    srv.drain(clock=lambda: 0.0)
    assert len(srv.done) == 6


def test_partial_batch_waits_while_its_graph_runs():
    """While a batch of a graph is in flight, a second batch of that graph
    starts only when a full one is waiting; a partial queue waits for the
    running batch and then goes out at once."""
    srv = make_server(ngraphs=1, max_batch=4, budget=30, footprint=10, engines=3)
    for i in range(6):
        srv.submit(serve.Query(i, 0, i, [0], 0.0))
    srv.poll(clock=lambda: 0.0)   # one batch of 4; the 2 left wait
    assert [n for _, n in FakeEngine.launches] == [4]
    for i in range(6, 9):        # 5 queued: a full batch overlaps
        srv.submit(serve.Query(i, 0, i, [0], 0.0))
    srv.poll(clock=lambda: 0.0)
    assert [n for _, n in FakeEngine.launches] == [4, 4]
    assert sum(e.batch is not None for e in srv.engines[0]) == 2
    srv.drain(clock=lambda: 0.0)
    assert [n for _, n in FakeEngine.launches] == [4, 4, 1]
    assert len(srv.done) == 9


def test_extra_engines_leave_room_for_idle_graphs():
    srv = make_server(ngraphs=3, max_batch=1, budget=30, footprint=10, engines=4)
    for i in range(5):
        srv.submit(serve.Query(i, 0, i, [0], 0.0))
    srv.poll(clock=lambda: 0.0)
    # graph 0 may not take the memory graphs 1 and 2 need for a first engine
    assert len(srv.engines[0]) == 1 and srv.used == 10
    srv.submit(serve.Query(9, 2, 9, [0], 0.0))
    srv.poll(clock=lambda: 0.0)
    assert len(srv.engines[2]) == 1
    srv.drain(clock=lambda: 0.0)
    assert len(srv.done) == 6


def test_budget_below_one_engine_is_refused():
    srv = make_server(ngraphs=1, budget=5, footprint=10)
    srv.submit(serve.Query(0, 0, 1, [0], 0.0))
    with pytest.raises(MemoryError):
        srv.poll(clock=lambda: 0.0)


def test_poisson_workload_and_latency_summary():
    graphs = [FakeGraph(), FakeGraph()]
    graphs[1].info = type("I", (), {"n": 64, "m": 3000})
    pools = [np.arange(5), np.arange(7)]
    qs = serve.poisson_workload(graphs, pools, 4000, rate_qps=1000.0, seed=3)
    t = np.array([q.t_arrive for q in qs])
    assert np.all(np.diff(t) >= 0)
    assert abs(len(qs) / t[-1] - 1000.0) < 100.0
    share = np.mean([q.gid == 1 for q in qs])
    assert abs(share - 0.75) < 0.05  # proportional to edge counts 1000 : 3000
    assert all(q.root in pools[q.gid] for q in qs)
    assert all(np.all(q.targets < graphs[q.gid].info.n) for q in qs)
    for i, q in enumerate(qs[:100]):
        q.t_done = q.t_arrive + (i + 1) * 1e-3
    s = serve.latency_summary(qs[:100])
    assert s["p50_ms"] == pytest.approx(50.5, rel=0.02) and s["p99_ms"] > s["p50_ms"]


def test_latency_summary_steady_state():
    """Queries every 1 ms for 4 s with a constant 30 ms latency: the steady
    window (middle half of the arrivals) completes 1000 q/s at 30 ms, the
    first and last quarters have the same mean; a latency that grows with
    arrival time shows as a rise from the first quarter to the last."""
    qs = [serve.Query(i, 0, 0, [], t_arrive=i * 1e-3) for i in range(4000)]
    for q in qs:
        q.t_done = q.t_arrive + 0.030
    s = serve.latency_summary(qs)
    assert s["steady_window_s"] == pytest.approx(2.0, rel=0.01)
    assert s["steady_arrival_qps"] == pytest.approx(1000.0, rel=0.01)
    assert s["steady_completion_qps"] == pytest.approx(1000.0, rel=0.02)
    assert s["steady_p50_ms"] == pytest.approx(30.0, rel=1e-6)
    assert s["first_quarter_mean_ms"] == pytest.approx(s["last_quarter_mean_ms"], rel=1e-6)
    assert s["qps"] < 1000.0  # the whole-run figure pays the last latency
    for i, q in enumerate(qs):
        q.t_done = q.t_arrive + 0.030 + i * 1e-5  # queue growing by 10 us per query
    g = serve.latency_summary(qs)
    assert g["last_quarter_mean_ms"] > g["first_quarter_mean_ms"] + 25.0
    assert g["steady_completion_qps"] < g["steady_arrival_qps"]
