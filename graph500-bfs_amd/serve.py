"""BFS query serving on resident graphs with continuous batching.

A query names a resident graph, a root and a few target vertices; its answer
is the root's traversed-edge count (the validator's TE) and the BFS levels
of the targets.  Each graph has a pool of batched-BFS engines
(g500_msbfs_*), each with its own HIP stream and at most one batch in
flight.  The loop

  * admits arriving queries into a per-graph queue,
  * launches a batch on an idle engine as soon as its graph has queued
    queries and none of its batches is in flight (up to max_batch of them,
    never waiting for a batch to fill); while one is in flight, another
    batch of the same graph starts only when a full one (max_batch queries)
    is waiting -- overlapping batches of one graph share the device's
    bandwidth, so small ones would only shrink every batch,
  * creates another engine for a graph whose engines are all busy while the
    free-device-memory budget covers one more engine's footprint and still
    leaves room for a first engine of every graph that has none,
  * polls the engines without blocking and completes the queries of every
    finished batch (TE and target levels fetched from the device).

Batches of different engines overlap on the device.  Every result is a
function of (graph, root, targets) alone: the batched engine gives each root
the levels of an independent BFS, whatever else is in its batch.
"""
import collections
import time

import numpy as np

import g500bfs as G


class Query:
    __slots__ = ("qid", "gid", "root", "targets", "t_arrive", "t_launch", "t_done",
                 "te", "levels")

    def __init__(self, qid, gid, root, targets, t_arrive=0.0):
        self.qid, self.gid, self.root = qid, gid, int(root)
        self.targets = np.asarray(targets, np.int64)
        self.t_arrive = t_arrive
        self.t_launch = self.t_done = None
        self.te = None
        self.levels = None


class _Engine:
    def __init__(self, mb):
        self.mb = mb        # a g500bfs.MultiBfs (or a stand-in with its methods)
        self.batch = None   # queries of the batch in flight


class Server:
    """Continuous-batching BFS server over resident graphs.

    graphs: list of constructed g500bfs.Graph (no CoreBFS).
    hbm_budget: device bytes the engines may take (default: free memory
    minus 8 GiB when the server starts).
    """

    def __init__(self, graphs, max_batch=256, hbm_budget=None, max_engines_per_graph=4,
                 device=0, engine_factory=None, footprint=None):
        self.graphs = list(graphs)
        self.max_batch = int(max_batch)
        if hbm_budget is None:
            free, _ = G.device_memory(device)
            hbm_budget = max(0, free - (8 << 30))
        self.budget = int(hbm_budget)
        self.used = 0
        self.max_engines = int(max_engines_per_graph)
        # answers need point lookups only, so engines keep no level arrays
        self.new_engine = engine_factory or \
            (lambda g, r: G.MultiBfs(g, max_roots=r, level_arrays=False))
        self.footprint = list(footprint) if footprint is not None else \
            [G.msbfs_footprint(g, self.max_batch, level_arrays=False) for g in self.graphs]
        self.engines = [[] for _ in self.graphs]
        self.queues = [collections.deque() for _ in self.graphs]
        self.done = []
        self.batches = 0
        self.refused_engines = 0

    # -- admission ---------------------------------------------------------
    def submit(self, q, now=None):
        if q.gid < 0 or q.gid >= len(self.graphs):
            raise ValueError(f"query {q.qid}: no graph {q.gid}")
        if now is not None and q.t_arrive is None:
            q.t_arrive = now
        self.queues[q.gid].append(q)

    def _idle_engine(self, gid):
        for e in self.engines[gid]:
            if e.batch is None:
                return e
        # a graph's extra engine must leave room for the first engine of
        # every graph that has none yet
        reserve = 0
        if self.engines[gid]:
            reserve = sum(f for g, f in enumerate(self.footprint)
                          if g != gid and not self.engines[g])
        if len(self.engines[gid]) < self.max_engines and \
                self.used + self.footprint[gid] + reserve <= self.budget:
            e = _Engine(self.new_engine(self.graphs[gid], self.max_batch))
            self.engines[gid].append(e)
            self.used += self.footprint[gid]
            return e
        if not self.engines[gid]:
            raise MemoryError(f"graph {gid}: one engine needs {self.footprint[gid]} bytes, "
                              f"budget left {self.budget - self.used}")
        self.refused_engines += 1
        return None

    def _launch(self, now):
        for gid, qu in enumerate(self.queues):
            while qu:
                busy = any(e.batch is not None for e in self.engines[gid])
                if busy and len(qu) < self.max_batch:
                    break  # the running batch finishes first; a fuller one follows
                e = self._idle_engine(gid)
                if e is None:
                    break
                n = min(len(qu), self.max_batch)
                batch = [qu.popleft() for _ in range(n)]
                e.mb.launch(np.array([q.root for q in batch], np.int64))
                for q in batch:
                    q.t_launch = now
                e.batch = batch
                self.batches += 1

    def _complete(self, e, clock):
        e.mb.wait()
        batch, e.batch = e.batch, None
        te = e.mb.te()
        idx = np.concatenate([np.full(len(q.targets), i, np.int64) for i, q in enumerate(batch)]) \
            if batch else np.zeros(0, np.int64)
        tg = np.concatenate([q.targets for q in batch]) if batch else np.zeros(0, np.int64)
        lev = e.mb.gather(idx, tg)
        t = clock()
        pos = 0
        for i, q in enumerate(batch):
            q.te = int(te[i])
            q.levels = lev[pos:pos + len(q.targets)].copy()
            pos += len(q.targets)
            q.t_done = t
            self.done.append(q)

    def poll(self, clock=time.perf_counter):
        """Completes finished batches and launches queued work; returns the
        number of batches in flight afterwards."""
        for engines in self.engines:
            for e in engines:
                if e.batch is not None and e.mb.done():
                    self._complete(e, clock)
        self._launch(clock())
        return sum(1 for es in self.engines for e in es if e.batch is not None)

    def prewarm(self, engines_per_graph=1, root_pools=None):
        """Creates engines up front (within the budget) and runs one batch of
        each lane-word count on them, so that no allocation or graph capture
        lands inside a measured run."""
        for gid, g in enumerate(self.graphs):
            for _ in range(engines_per_graph):
                if len(self.engines[gid]) >= self.max_engines or \
                        self.used + self.footprint[gid] > self.budget:
                    break
                e = _Engine(self.new_engine(g, self.max_batch))
                self.engines[gid].append(e)
                self.used += self.footprint[gid]
            pool = root_pools[gid] if root_pools is not None else g.find_roots(1)
            for e in self.engines[gid]:
                for n in range(64, self.max_batch + 64, 64):
                    n = min(n, self.max_batch)
                    e.mb.run(np.resize(np.asarray(pool, np.int64), n))
                e.mb.run(np.resize(np.asarray(pool, np.int64), 1))

    def drain(self, clock=time.perf_counter):
        while any(self.queues) or any(e.batch is not None for es in self.engines for e in es):
            self.poll(clock)

    def close(self):
        for es in self.engines:
            for e in es:
                if e.batch is not None:
                    e.mb.wait()
                e.mb.destroy()
        self.engines = [[] for _ in self.graphs]
        self.used = 0


def poisson_workload(graphs, root_pools, n_queries, rate_qps, targets_per_query=4, seed=1):
    """Open arrival stream: exponential gaps at rate_qps, graph chosen with
    probability proportional to its edge count, root uniform from the graph's
    pool of non-isolated vertices, targets uniform over its vertex IDs."""
    rng = np.random.default_rng(seed)
    m = np.array([g.info.m for g in graphs], float)
    p = m / m.sum()
    gaps = rng.exponential(1.0 / rate_qps, n_queries)
    t = np.cumsum(gaps)
    gids = rng.choice(len(graphs), size=n_queries, p=p)
    out = []
    for i in range(n_queries):
        g = int(gids[i])
        pool = root_pools[g]
        root = int(pool[rng.integers(0, len(pool))])
        tg = rng.integers(0, graphs[g].info.n, targets_per_query)
        out.append(Query(i, g, root, tg, float(t[i])))
    return out


def run_open_loop(server, queries, clock=time.perf_counter):
    """Replays queries at their arrival times (relative seconds) and returns
    the completed queries with t_arrive / t_done on the same clock."""
    qs = sorted(queries, key=lambda q: q.t_arrive)
    t0 = clock()
    for q in qs:
        q.t_arrive += t0
    i = 0
    while i < len(qs):
        now = clock()
        while i < len(qs) and qs[i].t_arrive <= now:
            server.submit(qs[i])
            i += 1
        server.poll(clock)
    server.drain(clock)
    return qs


def latency_summary(queries):
    """Whole-run figures (qps = queries over the span from the first arrival
    to the last completion, so the last batches' latency counts against it)
    and steady-state ones: the queries arriving in the middle half of the
    arrival period, their latency percentiles, the completion rate over that
    window, and the mean latency of the first and the last quarter of
    arrivals (a queue that grows shows as a rise between the two)."""
    arr = np.array([q.t_arrive for q in queries])
    done = np.array([q.t_done for q in queries])
    lat = (done - arr) * 1e3
    span = done.max() - arr.min()
    out = {"queries": len(queries), "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
           "qps": len(queries) / span}
    a0, a1 = np.percentile(arr, 25), np.percentile(arr, 75)
    mid = (arr >= a0) & (arr < a1)
    if a1 > a0 and mid.any():
        order = np.argsort(arr)
        q4 = max(1, len(order) // 4)
        out.update({
            "steady_window_s": float(a1 - a0),
            "steady_arrival_qps": float(mid.sum() / (a1 - a0)),
            "steady_completion_qps": float(((done >= a0) & (done < a1)).sum() / (a1 - a0)),
            "steady_p50_ms": float(np.percentile(lat[mid], 50)),
            "steady_p99_ms": float(np.percentile(lat[mid], 99)),
            "first_quarter_mean_ms": float(lat[order[:q4]].mean()),
            "last_quarter_mean_ms": float(lat[order[-q4:]].mean()),
        })
    return out


def one_per_launch(graphs, queries, clock=time.perf_counter):
    """Baseline: every query alone, one launch each, back to back (plain
    single-root engine for the BFS, batch-of-one engine for the answer)."""
    bfs = [G.Bfs(g) for g in graphs]
    one = [G.MultiBfs(g, max_roots=1, level_arrays=False) for g in graphs]
    service = []
    plain = []
    for q in queries:
        t0 = clock()
        bfs[q.gid].run(q.root)
        t1 = clock()
        one[q.gid].run(np.array([q.root]))
        te = int(one[q.gid].te()[0])
        lev = one[q.gid].gather(np.zeros(len(q.targets), np.int64), q.targets)
        t2 = clock()
        plain.append(t1 - t0)
        service.append(t2 - t1)
        q.te, q.levels = te, lev
    for b in bfs:
        b.destroy()
    for o in one:
        o.destroy()
    plain = np.array(plain) * 1e3
    service = np.array(service) * 1e3
    return {"plain_p50_ms": float(np.percentile(plain, 50)),
            "plain_p99_ms": float(np.percentile(plain, 99)),
            "plain_qps": float(1e3 / plain.mean()),
            "single_p50_ms": float(np.percentile(service, 50)),
            "single_p99_ms": float(np.percentile(service, 99)),
            "single_qps": float(1e3 / service.mean())}
