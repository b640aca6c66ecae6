"""Host logic of the serving loop (compresstemporalgraph_amd/serving.py) on
the CPU: request generation, the memory budget, latency statistics, and the
admission / batching / result bookkeeping of serve_batched and serve_plain
against a stand-in graph whose queries finish after a known number of steps.
The GPU parity of the served results is in test_gpu_parity.py."""
import numpy as np

from compresstemporalgraph_amd import serving


def test_poisson_requests_shape():
    req = serving.poisson_requests(rate=1000, count=400, n=5000, t_lo=100, t_hi=1 << 20, seed=9)
    assert len(req) == 400
    arr = np.array([r.arrival for r in req])
    assert np.all(np.diff(arr) > 0)
    # mean inter-arrival close to 1/rate
    assert abs(arr[-1] / len(req) - 1e-3) < 3e-4
    lens = np.array([r.window[1] - r.window[0] for r in req])
    assert lens.min() >= 2 ** 10 - 1 and lens.max() <= 2 ** 24
    assert all(100 <= r.window[0] < (1 << 20) for r in req)
    assert all(0 <= r.src < 5000 and r.kind == "BFS" for r in req)
    # seeded
    again = serving.poisson_requests(rate=1000, count=400, n=5000, t_lo=100, t_hi=1 << 20, seed=9)
    assert [(r.arrival, r.src, r.window) for r in again] == [(r.arrival, r.src, r.window) for r in req]


def test_mixed_requests_share():
    base = serving.poisson_requests(2000, 3000, 100, 0, 1 << 24, seed=4)
    mix = serving.mixed_requests(2000, 3000, 100, 0, 1 << 24, dc_share=0.25, seed=4)
    assert [(r.arrival, r.src, r.window) for r in mix] == [(r.arrival, r.src, r.window) for r in base]
    kinds = [r.kind for r in mix]
    assert set(kinds) == {"BFS", "DC"}
    assert abs(kinds.count("DC") / len(kinds) - 0.25) < 0.03


def test_stats_percentiles_and_kinds():
    req = [serving.Request(arrival=float(i), src=0, window=(0, 1), kind="DC" if i % 4 == 0 else "BFS")
           for i in range(100)]
    lat = np.linspace(0.001, 0.100, 100)
    done = [r.arrival + l for r, l in zip(req, lat)]
    st = serving._stats(req, done)
    assert st["queries"] == 100
    assert abs(st["p50_ms"] - np.percentile(lat, 50) * 1e3) < 1e-6
    assert abs(st["p99_ms"] - np.percentile(lat, 99) * 1e3) < 1e-6
    assert abs(st["qps"] - round(100 / (done[-1] - 0.0), 1)) < 1e-9
    assert st["DC"]["queries"] == 25 and st["BFS"]["queries"] == 75
    assert abs(st["DC"]["p50_ms"] - np.percentile(lat[::4], 50) * 1e3) < 1e-6


class _FakeServer:
    """Finishes the query admitted to a lane after `key % 3 + 1` steps (key:
    the BFS source, else the window start)."""

    LANES = 64

    def __init__(self, kind, log, graph, cost):
        self.kind, self.log, self.graph, self.cost = kind, log, graph, cost
        self.left, self.query = {}, {}
        graph.bytes += cost
        self.closed = False

    def _take(self, key, query):
        lane = min(set(range(self.LANES)) - set(self.left) - set(self.query))
        self.left[lane], self.query[lane] = key % 3 + 1, query
        self.log.append(("admit", self.kind, query))
        return lane

    def admit(self, *args, **kw):
        if self.kind == "BFS":
            src, window = args
            return self._take(src, ("bfs", src, window))
        window = args[0]
        if self.kind == "PR":
            assert kw["iters"] == serving.PR_ITERS
            return self._take(window[0], ("pr", window))
        return self._take(window[0], ("cc", window))

    def step(self):
        assert self.left, "step with no running query"
        self.log.append(("step", self.kind))
        fin = []
        for lane in list(self.left):
            self.left[lane] -= 1
            if self.left[lane] == 0:
                fin.append(lane)
                del self.left[lane]
        return fin

    def result(self, lane, out=None):
        assert lane not in self.left
        return self.query.pop(lane)

    def close(self):
        assert not self.closed
        self.closed = True
        self.graph.bytes -= self.cost


class _FakeGraph:
    n = 10
    device = 0
    COST = {"BFS": 300, "CC": 500, "PR": 900}

    def __init__(self):
        self.log, self.dc_calls, self.dc_plain, self.plain = [], [], 0, []
        self.bytes = 1000  # the handle's own state
        self.servers = []

    def device_bytes(self):
        return self.bytes

    def _server(self, kind):
        s = _FakeServer(kind, self.log, self, self.COST[kind])
        self.servers.append(s)
        return s

    def bfs_server(self):
        return self._server("BFS")

    def cc_server(self):
        return self._server("CC")

    def pagerank_server(self, damping=0.85):
        assert damping == 0.85
        return self._server("PR")

    def degree_centrality_batch(self, wins, out=None, out_centrality=None):
        assert 1 <= len(wins) <= 64
        self.dc_calls.append(list(wins))
        return [("cnt", w) for w in wins], [("cen", w) for w in wins]

    def bfs(self, src, window=None, out=None):
        self.plain.append("BFS")
        return ("bfs", src, window)

    def degree_centrality(self, window=None, centrality=False, out=None, out_centrality=None):
        self.dc_plain += 1
        return ("cnt", window), ("cen", window)

    def pagerank(self, window=None, iters=20, out=None):
        self.plain.append("PR")
        return ("pr", window)

    def connected_components(self, window=None, out=None):
        self.plain.append("CC")
        return ("cc", window)


WANT = {"BFS": lambda r: ("bfs", r.src, r.window), "CC": lambda r: ("cc", r.window),
        "PR": lambda r: ("pr", r.window), "DC": lambda r: (("cnt", r.window), ("cen", r.window))}


def test_serve_loop_bookkeeping():
    """Every request is answered exactly once with its own result, in both
    loops; DC requests go through batched calls of at most 64; BFS requests
    are admitted in arrival order; the server is closed at the end."""
    req = serving.mixed_requests(rate=1e6, count=300, n=1000, t_lo=0, t_hi=1 << 20, dc_share=0.3, seed=2)
    g = _FakeGraph()
    kb, kp = {}, {}
    sb = serving.serve_batched(g, req, keep=kb)
    sp = serving.serve_plain(g, req, keep=kp)
    assert sorted(kb) == list(range(len(req))) == sorted(kp)
    for i, r in enumerate(req):
        assert kb[i] == WANT[r.kind](r) == kp[i]
    # every DC request went through a batched call (of 2..64) or, alone, a
    # plain one (serve_plain's calls are on the same stand-in: subtract them)
    ndc = sum(r.kind == "DC" for r in req)
    assert sum(len(c) for c in g.dc_calls) + g.dc_plain - ndc == ndc
    assert all(serving.MIN_BATCH["DC"] <= len(c) <= 64 for c in g.dc_calls)
    order = sorted((r.arrival, r.src) for r in req if r.kind == "BFS")
    assert [q[1] for k, kind, q in (e for e in g.log if e[0] == "admit")] == [s for _, s in order]
    assert sb["queries"] == sp["queries"] == len(req) and sb["lanes"] == {"BFS": 64}
    assert [s.kind for s in g.servers] == ["BFS"] and all(s.closed for s in g.servers)
    assert sb["state_bytes"]["BFS"] == 300 and g.bytes == 1000


def test_mixed_requests_four_kinds():
    mix = serving.mixed_requests(1e4, 4000, 100, 0, 1 << 24, seed=8,
                                 shares={"DC": 0.1, "PR": 0.3, "CC": 0.2})
    kinds = [r.kind for r in mix]
    for k, share in (("BFS", 0.4), ("CC", 0.2), ("PR", 0.3), ("DC", 0.1)):
        assert abs(kinds.count(k) / len(kinds) - share) < 0.03, k


def test_serve_loop_all_kinds(monkeypatch):
    """BFS, CC, PR and DC requests: each answered once with its own result;
    with every server started by a single query, BFS, CC and PR all go
    through their servers (admitted in arrival order per kind), DC in
    batched calls of at most 64."""
    monkeypatch.setattr(serving, "SERVE_FROM", {"BFS": 1, "CC": 1, "PR": 1})
    req = serving.mixed_requests(rate=1e5, count=400, n=1000, t_lo=0, t_hi=1 << 20, seed=6,
                                 shares={"DC": 0.1, "PR": 0.3, "CC": 0.2})
    g = _FakeGraph()
    kb, kp = {}, {}
    st = serving.serve_batched(g, req, keep=kb)
    serving.serve_plain(g, req, keep=kp)
    for i, r in enumerate(req):
        assert kb[i] == WANT[r.kind](r) == kp[i], (i, r.kind)
    assert sorted(s.kind for s in g.servers) == ["BFS", "CC", "PR"]
    assert st["state_bytes"] == {"BFS": 300, "CC": 500, "PR": 900, "DC": 64 * 8 * g.n}
    for kind in ("BFS", "CC", "PR"):
        admitted = [e[2] for e in g.log if e[0] == "admit" and e[1] == kind]
        assert admitted == [WANT[kind](r) for r in sorted(req, key=lambda r: r.arrival) if r.kind == kind]
    # the batched loop made no plain call of a served kind
    assert len(g.plain) == sum(r.kind != "DC" for r in req)


def test_serve_loop_plain_below_threshold():
    """An idle CC or PR server starts only once SERVE_FROM queries of its kind
    wait: a burst (all arrived at once) goes through the servers, sparse
    arrivals (each answered before the next comes) through plain calls."""
    burst = serving.mixed_requests(rate=1e9, count=200, n=1000, t_lo=0, t_hi=1 << 20, seed=4,
                                   shares={"PR": 0.5, "CC": 0.5})
    g = _FakeGraph()
    kb = {}
    serving.serve_batched(g, burst, keep=kb)
    assert all(kb[i] == WANT[r.kind](r) for i, r in enumerate(burst))
    assert not g.plain and sorted(s.kind for s in g.servers) == ["CC", "PR"]
    sparse = serving.mixed_requests(rate=500, count=12, n=1000, t_lo=0, t_hi=1 << 20, seed=4,
                                    shares={"PR": 0.5, "CC": 0.5})
    g = _FakeGraph()
    kb = {}
    serving.serve_batched(g, sparse, keep=kb)
    assert all(kb[i] == WANT[r.kind](r) for i, r in enumerate(sparse))
    assert sorted(g.plain) == sorted(r.kind for r in sparse)
    assert not [e for e in g.log if e[0] == "admit"]


def test_serve_loop_fair_share(monkeypatch):
    """Weighted fair queuing: while PR steps cost 3 ms each and BFS steps
    almost nothing, the loop keeps stepping BFS between PR steps instead of
    alternating the two."""
    import time as _t
    monkeypatch.setattr(serving, "SERVE_FROM", {"BFS": 1, "CC": 1, "PR": 1})
    req = [serving.Request(arrival=0.0, src=3 * k + 2, window=(3 * k + 2, 9), kind="BFS" if k % 2 else "PR")
           for k in range(40)]
    g = _FakeGraph()
    orig = _FakeServer.step

    def step(self):
        if self.kind == "PR":
            _t.sleep(0.003)
        return orig(self)

    monkeypatch.setattr(_FakeServer, "step", step)
    kb = {}
    serving.serve_batched(g, req, keep=kb)
    assert all(kb[i] == WANT[r.kind](r) for i, r in enumerate(req))
    steps = [e[1] for e in g.log if e[0] == "step"]
    # every query takes 3 steps: the 20 BFS queries finish within the first
    # PR step or two, not interleaved one-for-one with PR's three
    last_bfs = max(i for i, k in enumerate(steps) if k == "BFS")
    assert steps[:last_bfs + 1].count("PR") <= 2, steps


def test_serve_loop_budget_fallback():
    """A budget that holds the BFS server but not the CC or PR state: those
    kinds are answered by plain calls (their servers are opened, measured and
    closed again), and DC batches shrink to what is left."""
    req = serving.mixed_requests(rate=1e5, count=200, n=1000, t_lo=0, t_hi=1 << 20, seed=3,
                                 shares={"DC": 0.2, "PR": 0.2, "CC": 0.2})
    g = _FakeGraph()
    kb = {}
    budget = 300 + 5 * 8 * g.n + 50  # BFS, then 450 left: neither CC (500) nor PR (900)
    st = serving.serve_batched(g, req, keep=kb, budget_bytes=budget)
    for i, r in enumerate(req):
        assert kb[i] == WANT[r.kind](r), (i, r.kind)
    assert st["lanes"] == {"BFS": 64}
    assert st["state_bytes"]["DC"] == 5 * 8 * g.n
    assert all(len(c) <= 5 for c in g.dc_calls)
    assert sorted(g.plain) == sorted(r.kind for r in req if r.kind in ("CC", "PR"))
    assert all(s.closed for s in g.servers) and g.bytes == 1000
