"""Serving form: BFS window queries arriving over time, served by continuous
batching on one GPU.

Requests are (arrival time, source, window).  The loop admits every request
that has arrived into a free lane of the BFS server (up to 64 at once, fewer
when the per-query state would exceed the memory budget), advances all
running queries by one BFS level per decode pass, and completes a query as
soon as its frontier empties -- so a query that arrives while others are on
the GPU joins them at the next level instead of waiting for the batch to
drain.  The plain baseline serves the same arrivals one query at a time with
TemporalGraph.bfs, in arrival order.

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


def poisson_requests(rate: float, count: int, n: int, t_lo: int, t_hi: int, seed: int = 0) -> list:
    """`count` requests at Poisson rate `rate` (per second), random sources and
    windows with log-uniform lengths in [2^10, 2^24] and uniform starts."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), count)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), count)
    srcs = rng.integers(0, n, count)
    return [Request(float(a), int(s), (int(b), int(b + l))) for a, s, b, l in zip(arrivals, srcs, starts, lens)]


def lanes_for_budget(n: int, budget_bytes: int) -> int:
    """Queries that fit the memory budget: each holds an int32 level row of n
    entries (the three shared u64 masks are charged once)."""
    shared = 3 * 8 * n
    return int(max(1, min(BfsServer.LANES, (budget_bytes - shared) // (4 * n))))


def _stats(req, done_at) -> dict:
    lat = np.array([done_at[i] - r.arrival for i, r in enumerate(req)])
    span = max(done_at) - min(r.arrival for r in req)
    return {"queries": len(req), "qps": round(len(req) / span, 1),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}


def _sink(g: TemporalGraph, device_out: bool):
    """Output buffer for one result: a host array, or a device row (results
    then stay in HBM, as a server handing them to GPU consumers would)."""
    if not device_out:
        return lambda: None
    import torch
    return lambda: torch.empty(g.n, dtype=torch.int32, device=f"cuda:{g.device}")


def serve_batched(g: TemporalGraph, req: list, budget_bytes: int = 8 << 30, keep=None,
                  device_out: bool = False) -> dict:
    """Continuous batching.  `keep` (a dict) receives the results per request
    index when given (for identity checks)."""
    lanes = lanes_for_budget(g.n, budget_bytes)
    sink = _sink(g, device_out)
    done_at = [0.0] * len(req)
    order = sorted(range(len(req)), key=lambda i: req[i].arrival)
    with g.bfs_server() as srv:
        lane_of = {}
        nxt = 0
        t0 = time.perf_counter()
        while len(done_at) and (nxt < len(order) or lane_of):
            now = time.perf_counter() - t0
            while nxt < len(order) and req[order[nxt]].arrival <= now and len(lane_of) < lanes:
                i = order[nxt]
                lane_of[srv.admit(req[i].src, req[i].window)] = i
                nxt += 1
            if not lane_of:  # idle until the next arrival
                time.sleep(max(0.0, req[order[nxt]].arrival - now))
                continue
            for lane in srv.step():
                i = lane_of.pop(lane)
                levels = srv.result(lane, out=sink())
                done_at[i] = time.perf_counter() - t0
                if keep is not None:
                    keep[i] = levels
    out = _stats(req, done_at)
    out["lanes"] = lanes
    return out


def serve_plain(g: TemporalGraph, req: list, keep=None, device_out: bool = False) -> dict:
    """One query at a time, in arrival order."""
    sink = _sink(g, device_out)
    done_at = [0.0] * len(req)
    order = sorted(range(len(req)), key=lambda i: req[i].arrival)
    t0 = time.perf_counter()
    for i in order:
        now = time.perf_counter() - t0
        if req[i].arrival > now:
            time.sleep(req[i].arrival - now)
        levels = g.bfs(req[i].src, req[i].window, out=sink())
        done_at[i] = time.perf_counter() - t0
        if keep is not None:
            keep[i] = levels
    return _stats(req, done_at)
