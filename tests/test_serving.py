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


def test_lanes_for_budget():
    n = 1 << 20
    shared = 3 * 8 * n
    assert serving.lanes_for_budget(n, 8 << 30) == 64
    assert serving.lanes_for_budget(n, shared + 10 * 4 * n + 5) == 10
    assert serving.lanes_for_budget(n, 0) == 1


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
    """Finishes the query admitted to a lane after `src % 3 + 1` steps."""

    LANES = 64

    def __init__(self, log):
        self.left, self.src, self.log = {}, {}, log

    def admit(self, src, window):
        lane = min(set(range(self.LANES)) - set(self.left))
        self.left[lane], self.src[lane] = src % 3 + 1, (src, window)
        self.log.append(("admit", src))
        return lane

    def step(self):
        assert self.left, "step with no running query"
        fin = []
        for lane in list(self.left):
            self.left[lane] -= 1
            if self.left[lane] == 0:
                fin.append(lane)
                del self.left[lane]
        return fin

    def result(self, lane, out=None):
        return ("bfs",) + self.src.pop(lane)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        pass


class _FakeGraph:
    n = 10
    device = 0

    def __init__(self):
        self.log, self.dc_calls, self.dc_plain = [], [], 0

    def bfs_server(self):
        return _FakeServer(self.log)

    def degree_centrality_batch(self, wins, out=None, out_centrality=None):
        assert 1 <= len(wins) <= 64
        self.dc_calls.append(list(wins))
        return [("cnt", w) for w in wins], [("cen", w) for w in wins]

    def bfs(self, src, window=None, out=None):
        return ("bfs", src, window)

    def degree_centrality(self, window=None, centrality=False, out=None, out_centrality=None):
        self.dc_plain += 1
        return ("cnt", window), ("cen", window)

    def pagerank_batch(self, wins, iters=20, out=None):
        assert 1 <= len(wins) <= 64 and iters == serving.PR_ITERS
        self.pr_calls = getattr(self, "pr_calls", []) + [list(wins)]
        return [("pr", w) for w in wins]

    def pagerank(self, window=None, iters=20, out=None):
        return ("pr", window)

    def connected_components(self, window=None, out=None):
        return ("cc", window)

    def connected_components_batch(self, wins, out=None):
        assert serving.MIN_BATCH["CC"] <= len(wins) <= 64
        return [("cc", w) for w in wins]


def test_serve_loop_bookkeeping():
    """Every request is answered exactly once with its own result, in both
    loops; DC requests go through batched calls of at most 64; BFS requests
    are admitted in arrival order."""
    req = serving.mixed_requests(rate=1e6, count=300, n=1000, t_lo=0, t_hi=1 << 20, dc_share=0.3, seed=2)
    g = _FakeGraph()
    kb, kp = {}, {}
    sb = serving.serve_batched(g, req, keep=kb)
    sp = serving.serve_plain(g, req, keep=kp)
    assert sorted(kb) == list(range(len(req))) == sorted(kp)
    for i, r in enumerate(req):
        if r.kind == "DC":
            assert kb[i] == (("cnt", r.window), ("cen", r.window)) == kp[i]
        else:
            assert kb[i] == ("bfs", r.src, r.window) == kp[i]
    # every DC request went through a batched call (of 2..64) or, alone, a
    # plain one (serve_plain's calls are on the same stand-in: subtract them)
    ndc = sum(r.kind == "DC" for r in req)
    assert sum(len(c) for c in g.dc_calls) + g.dc_plain - ndc == ndc
    assert all(serving.MIN_BATCH["DC"] <= len(c) <= 64 for c in g.dc_calls)
    order = sorted((r.arrival, r.src) for r in req if r.kind == "BFS")
    assert [s for k, s in g.log if k == "admit"] == [s for _, s in order]
    assert sb["queries"] == sp["queries"] == len(req) and sb["lanes"] == 64


def test_serve_loop_dc_worker():
    """The same stream with DC batches on a second handle in a worker thread:
    every request answered once with its own result, DC calls only on the
    second handle, at most 64 windows each."""
    req = serving.mixed_requests(rate=1e5, count=300, n=1000, t_lo=0, t_hi=1 << 20, dc_share=0.3, seed=5)
    g, g2 = _FakeGraph(), _FakeGraph()
    kb = {}
    st = serving.serve_batched(g, req, keep=kb, workers={"DC": g2})
    assert sorted(kb) == list(range(len(req))) and st["queries"] == len(req)
    for i, r in enumerate(req):
        want = (("cnt", r.window), ("cen", r.window)) if r.kind == "DC" else ("bfs", r.src, r.window)
        assert kb[i] == want
    assert not g.dc_calls and not g.dc_plain
    assert sum(len(c) for c in g2.dc_calls) + g2.dc_plain == sum(r.kind == "DC" for r in req)
    assert all(serving.MIN_BATCH["DC"] <= len(c) <= 64 for c in g2.dc_calls)


def test_mixed_requests_four_kinds():
    mix = serving.mixed_requests(1e4, 4000, 100, 0, 1 << 24, seed=8,
                                 shares={"DC": 0.1, "PR": 0.3, "CC": 0.2})
    kinds = [r.kind for r in mix]
    for k, share in (("BFS", 0.4), ("CC", 0.2), ("PR", 0.3), ("DC", 0.1)):
        assert abs(kinds.count(k) / len(kinds) - share) < 0.03, k


def test_serve_loop_all_kinds():
    """BFS, CC, PR and DC requests, inline and with the worker: each answered
    once with its own result; PR and DC in batched calls of at most 64."""
    req = serving.mixed_requests(rate=1e5, count=400, n=1000, t_lo=0, t_hi=1 << 20, seed=6,
                                 shares={"DC": 0.1, "PR": 0.3, "CC": 0.2})
    want = {"BFS": lambda r: ("bfs", r.src, r.window), "CC": lambda r: ("cc", r.window),
            "PR": lambda r: ("pr", r.window), "DC": lambda r: (("cnt", r.window), ("cen", r.window))}
    for workers in (0, 1, 3):
        g, g2, g3, g4 = _FakeGraph(), _FakeGraph(), _FakeGraph(), _FakeGraph()
        kb, kp = {}, {}
        wmap = {0: None, 1: {"DC": g2, "PR": g2, "CC": g2}, 3: {"DC": g2, "PR": g3, "CC": g4}}[workers]
        serving.serve_batched(g, req, keep=kb, workers=wmap)
        serving.serve_plain(g, req, keep=kp)
        for i, r in enumerate(req):
            assert kb[i] == want[r.kind](r) == kp[i], (workers, i, r.kind)
        gb = {0: g, 1: g2, 3: g3}[workers]
        calls = getattr(gb, "pr_calls", [])
        assert all(serving.MIN_BATCH["PR"] <= len(c) <= 64 for c in calls)
