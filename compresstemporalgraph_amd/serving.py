"""Serving form: BFS, CC, PageRank (PR) and degree-centrality (DC) window
queries arriving over time, served by continuous batching on one GPU.

Requests are (arrival time, kind, source, window).  The loop admits every BFS
request that has arrived into a free lane of the BFS server (up to 64 at
once, fewer when the per-query state would exceed the memory budget),
advances all running queries by one BFS level per decode pass, and completes
a query as soon as its frontier empties -- so a query that arrives while
others are on the GPU joins them at the next level instead of waiting for the
batch to drain.  DC and PR requests that have arrived are answered together
by one batched call per kind (up to 64 windows; PR 20 iterations), and so
are CC requests.  Given more handles of the graph (one per kind, or shared),
these run in worker threads beside the BFS loop, else between its steps.  The plain baseline serves the
same arrivals one query at a time (TemporalGraph.bfs, connected_components,
pagerank, degree_centrality), in arrival order.

Latency is completion time minus arrival time; throughput is completed
queries over the span from the first arrival to the last completion.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .engine import BfsServer, TemporalGraph


@dataclass
class Request:
    arrival: float  # seconds after the start of the run
    src: int
    window: tuple
    kind: str = "BFS"  # "BFS", "CC", "PR" or "DC"


def poisson_requests(rate: float, count: int, n: int, t_lo: int, t_hi: int, seed: int = 0) -> list:
    """`count` requests at Poisson rate `rate` (per second), random sources and
    windows with log-uniform lengths in [2^10, 2^24] and uniform starts."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), count)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), count)
    srcs = rng.integers(0, n, count)
    return [Request(float(a), int(s), (int(b), int(b + l))) for a, s, b, l in zip(arrivals, srcs, starts, lens)]


def mixed_requests(rate: float, count: int, n: int, t_lo: int, t_hi: int, dc_share: float = 0.2,
                   seed: int = 0, shares: dict = None) -> list:
    """poisson_requests with shares of other kinds: `shares` maps "DC", "PR"
    and "CC" to fractions (the rest stays BFS); without it, dc_share of DC.
    src is used by BFS only."""
    req = poisson_requests(rate, count, n, t_lo, t_hi, seed)
    shares = shares if shares is not None else {"DC": dc_share}
    rng = np.random.default_rng(seed + 1)
    for r, u in zip(req, rng.random(count)):
        acc = 0.0
        for kind in ("DC", "PR", "CC"):
            acc += shares.get(kind, 0.0)
            if u < acc:
                r.kind = kind
                break
    return req


def lanes_for_budget(n: int, budget_bytes: int) -> int:
    """Queries that fit the memory budget: each holds an int32 level row of n
    entries (the three shared u64 masks are charged once)."""
    shared = 3 * 8 * n
    return int(max(1, min(BfsServer.LANES, (budget_bytes - shared) // (4 * n))))


def _stats(req, done_at) -> dict:
    lat = np.array([done_at[i] - r.arrival for i, r in enumerate(req)])
    span = max(done_at) - min(r.arrival for r in req)
    out = {"queries": len(req), "qps": round(len(req) / span, 1),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}
    kinds = sorted({r.kind for r in req})
    if len(kinds) > 1:
        for k in kinds:
            lk = np.array([lat[i] for i, r in enumerate(req) if r.kind == k])
            out[k] = {"queries": int(len(lk)), "p50_ms": round(float(np.percentile(lk, 50)) * 1e3, 3),
                      "p99_ms": round(float(np.percentile(lk, 99)) * 1e3, 3)}
    return out


def _sink(g: TemporalGraph, device_out: bool, rows: int = 0):
    """Output buffer for one result: a host array, or a device row (results
    then stay in HBM, as a server handing them to GPU consumers would); with
    rows, a (rows, n) block."""
    if not device_out:
        return lambda: None
    import torch
    shape = (rows, g.n) if rows else (g.n,)
    return lambda: torch.empty(shape, dtype=torch.int32, device=f"cuda:{g.device}")


PR_ITERS = 20


def _dc_batch(g: TemporalGraph, req: list, idx: list, device_out: bool):
    """One batched DC call for requests idx; returns [(count, centrality)] rows."""
    wins = [req[i].window for i in idx]
    if device_out:
        blk = _sink(g, True, len(idx))
        cnt, cen = g.degree_centrality_batch(wins, out=blk(), out_centrality=blk())
    else:
        cnt, cen = g.degree_centrality_batch(wins)
    return [(cnt[j], cen[j]) for j in range(len(idx))]


def _pr_batch(g: TemporalGraph, req: list, idx: list, device_out: bool):
    """One batched PageRank call for requests idx; returns its rows."""
    wins = [req[i].window for i in idx]
    if device_out:
        import torch
        out = torch.empty((len(idx), g.n), dtype=torch.float32, device=f"cuda:{g.device}")
        rank = g.pagerank_batch(wins, iters=PR_ITERS, out=out)
    else:
        rank = g.pagerank_batch(wins, iters=PR_ITERS)
    return [rank[j] for j in range(len(idx))]


def _cc_batch(g: TemporalGraph, req: list, idx: list, device_out: bool):
    """One batched CC call for requests idx; returns its rows."""
    wins = [req[i].window for i in idx]
    lab = g.connected_components_batch(wins, out=_sink(g, True, len(idx))() if device_out else None)
    return [lab[j] for j in range(len(idx))]


def _plain_one(g: TemporalGraph, r, device_out: bool):
    """One request answered by the plain entry point of its kind."""
    sink = _sink(g, device_out)
    if r.kind == "DC":
        return g.degree_centrality(r.window, centrality=True, out=sink(), out_centrality=sink())
    if r.kind == "PR":
        out = None
        if device_out:
            import torch
            out = torch.empty(g.n, dtype=torch.float32, device=f"cuda:{g.device}")
        return g.pagerank(r.window, iters=PR_ITERS, out=out)
    if r.kind == "CC":
        return g.connected_components(r.window, out=sink())
    return g.bfs(r.src, r.window, out=sink())


# Fewest pending queries of a kind worth one batched call: a batched call
# costs about the same whatever its width (passes over the graph for 64
# columns), so below these counts the plain calls are cheaper (on a 2^25
# contact graph a windowed batch costs about 24 plain PR, 9 plain CC or 2
# plain DC queries).  CC waits for 16: a batch made while the other kinds
# run takes several times its time alone and holds back every queue, so CC
# batches only under a real backlog.
MIN_BATCH = {"DC": 2, "PR": 16, "CC": 16}


def _answer(g: TemporalGraph, req: list, idx: list, device_out: bool) -> list:
    """Results for the non-BFS requests idx, batched per kind (up to 64
    each; plain calls below MIN_BATCH); returns [(i, result)]."""
    out = []
    for kind, fn in (("DC", _dc_batch), ("PR", _pr_batch), ("CC", _cc_batch)):
        sel = [i for i in idx if req[i].kind == kind]
        for b in range(0, len(sel), 64):
            part = sel[b:b + 64]
            if len(part) >= MIN_BATCH[kind]:
                out += list(zip(part, fn(g, req, part, device_out)))
            else:
                out += [(i, _plain_one(g, req[i], device_out)) for i in part]
    return out


class _BatchWorker:
    """Answers the non-BFS requests on a second handle of the same graph, in
    a thread of its own: each wake-up takes every pending request (DC and PR
    batched per kind, CC one at a time).  The handle has its own stream, so
    this work overlaps the BFS levels of the main loop on the device (the
    engine's calls release the interpreter lock while they wait for the
    GPU)."""

    def __init__(self, g: TemporalGraph, req: list, device_out: bool, t0: float, done_at: list, keep):
        import threading
        self.g, self.req, self.device_out, self.t0 = g, req, device_out, t0
        self.done_at, self.keep = done_at, keep
        self.cv = threading.Condition()
        self.pending, self.open, self.outstanding, self.error = [], True, 0, None
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def put(self, i: int) -> None:
        with self.cv:
            self.pending.append(i)
            self.outstanding += 1
            self.cv.notify()

    def busy(self) -> bool:
        with self.cv:
            if self.error is not None:
                raise self.error
            return self.outstanding > 0

    def close(self) -> None:
        with self.cv:
            self.open = False
            self.cv.notify()
        self.thread.join()

    def _run(self) -> None:
        while True:
            with self.cv:
                while self.open and not self.pending:
                    self.cv.wait()
                if not self.pending:
                    return
                idx, self.pending = self.pending, []
            try:
                res = _answer(self.g, self.req, idx, self.device_out)
            except Exception as e:  # surfaced by busy()
                with self.cv:
                    self.error, self.outstanding = e, 0
                return
            t = time.perf_counter() - self.t0
            with self.cv:
                for i, row in res:
                    self.done_at[i] = t
                    if self.keep is not None:
                        self.keep[i] = row
                self.outstanding -= len(idx)


def serve_batched(g: TemporalGraph, req: list, budget_bytes: int = 8 << 30, keep=None,
                  device_out: bool = False, workers: dict = None) -> dict:
    """Continuous batching.  `keep` (a dict) receives the results per request
    index when given (for identity checks).  `workers` maps the kinds "DC",
    "PR" and "CC" to other handles of the same graph: each distinct handle
    gets a worker thread that answers its kinds beside the BFS loop (the
    handles' streams overlap on the device); kinds without one are answered
    between the BFS steps."""
    lanes = lanes_for_budget(g.n, budget_bytes)
    sink = _sink(g, device_out)
    done_at = [0.0] * len(req)
    order = sorted(range(len(req)), key=lambda i: req[i].arrival)
    with g.bfs_server() as srv:
        lane_of, dc_wait = {}, []
        nxt = 0
        t0 = time.perf_counter()
        by_handle = {}
        for kind, h in (workers or {}).items():
            if id(h) not in by_handle:
                by_handle[id(h)] = _BatchWorker(h, req, device_out, t0, done_at, keep)
        worker_of = {kind: by_handle[id(h)] for kind, h in (workers or {}).items()}
        pool = list(by_handle.values())
        try:
            while len(done_at) and (nxt < len(order) or lane_of or dc_wait or any(w.busy() for w in pool)):
                now = time.perf_counter() - t0
                while nxt < len(order) and req[order[nxt]].arrival <= now:
                    i = order[nxt]
                    if req[i].kind != "BFS":
                        if req[i].kind in worker_of:
                            worker_of[req[i].kind].put(i)
                        else:
                            dc_wait.append(i)
                    elif len(lane_of) < lanes:
                        lane_of[srv.admit(req[i].src, req[i].window)] = i
                    else:
                        break  # BFS lanes full: this request waits (arrival order kept)
                    nxt += 1
                if not lane_of and not dc_wait:  # idle until the next arrival (or the workers)
                    wake = req[order[nxt]].arrival - now if nxt < len(order) else 1e-4
                    time.sleep(max(0.0, min(wake, 1e-3) if pool else wake))
                    continue
                if dc_wait:
                    idx, dc_wait = dc_wait, []
                    res = _answer(g, req, idx, device_out)
                    t = time.perf_counter() - t0
                    for i, row in res:
                        done_at[i] = t
                        if keep is not None:
                            keep[i] = row
                if lane_of:
                    for lane in srv.step():
                        i = lane_of.pop(lane)
                        levels = srv.result(lane, out=sink())
                        done_at[i] = time.perf_counter() - t0
                        if keep is not None:
                            keep[i] = levels
        finally:
            for w in pool:
                w.close()
    out = _stats(req, done_at)
    out["lanes"] = lanes
    return out


def serve_plain(g: TemporalGraph, req: list, keep=None, device_out: bool = False) -> dict:
    """One query at a time, in arrival order."""
    done_at = [0.0] * len(req)
    order = sorted(range(len(req)), key=lambda i: req[i].arrival)
    t0 = time.perf_counter()
    for i in order:
        now = time.perf_counter() - t0
        if req[i].arrival > now:
            time.sleep(req[i].arrival - now)
        res = _plain_one(g, req[i], device_out)
        done_at[i] = time.perf_counter() - t0
        if keep is not None:
            keep[i] = res
    return _stats(req, done_at)
