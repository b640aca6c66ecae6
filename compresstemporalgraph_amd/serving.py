"""Serving form: BFS, CC, PageRank (PR) and degree-centrality (DC) window
queries arriving over time, served by continuous batching on one handle.

Requests are (arrival time, kind, source, window).  BFS, CC and PR each have a
query server of 64 lanes: a request that has arrived is admitted into a free
lane of its kind's server and joins the queries already running there at the
server's next step -- one BFS level, one group of CC label rounds, or one PR
iteration (20 per query) -- so it never waits for a batch to drain.  DC (a
single pass per query) is answered by batched calls of the DC requests that
have arrived (up to 64).  The kinds share the handle's one stream by weighted
fair queuing (serve_batched), and a query completes as soon as its server
reports it finished.  Each kind's device state is charged against
the memory budget when its server opens (the handle's device bytes before
and after); a kind whose state does not fit is answered by plain calls, one
per turn of the loop.  The plain baseline serves the same arrivals one query
at a time (TemporalGraph.bfs, connected_components, pagerank,
degree_centrality), in arrival order.

Latency is completion time minus arrival time; throughput is completed
queries over the span from the first arrival to the last completion.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from .engine import TemporalGraph


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
# plain DC queries).  The serving loop batches DC by this rule; BFS, CC and
# PR go through their servers.
MIN_BATCH = {"DC": 2, "PR": 16, "CC": 16}

SERVED = ("BFS", "CC", "PR")  # kinds with a query server


def _open_server(g: TemporalGraph, kind: str):
    if kind == "BFS":
        return g.bfs_server()
    if kind == "CC":
        return g.cc_server()
    return g.pagerank_server(damping=0.85)


def _admit(srv, kind: str, r) -> int:
    if kind == "BFS":
        return srv.admit(r.src, r.window)
    if kind == "CC":
        return srv.admit(r.window)
    return srv.admit(r.window, iters=PR_ITERS)


def open_servers(g: TemporalGraph, kinds, budget_bytes: int):
    """Opens the servers of `kinds` (in SERVED order) while their device state
    fits the budget; returns ({kind: server}, {kind: bytes charged}).  A
    server's charge is the handle's device bytes after its opening minus
    before (its lanes, plus the batched run structure CC and PR share when the
    handle has not built it yet, charged to whichever opens first; it stays
    with the handle); one that does not fit is closed again and its kind is
    left to plain calls.  DC is charged its widest batch: count and
    centrality rows of 64 queries (fewer when the rest of the budget is
    smaller)."""
    servers, charged = {}, {}
    left = budget_bytes
    for kind in SERVED:
        if kind not in kinds:
            continue
        before = g.device_bytes()
        srv = _open_server(g, kind)
        cost = g.device_bytes() - before
        if cost > left:
            srv.close()
            continue
        servers[kind], charged[kind] = srv, cost
        left -= cost
    if "DC" in kinds:
        rows = int(max(1, min(64, left // (8 * g.n))))
        charged["DC"] = rows * 8 * g.n
    return servers, charged


# An idle CC or PR server starts once this many queries of its kind wait:
# its step costs about the same at any lane count, so fewer are cheaper as
# plain calls.  A running server takes every arrival of its kind.  The BFS
# server's step costs what its frontiers hold, so it always serves.
SERVE_FROM = {"BFS": 1, "CC": 8, "PR": 16}


def serve_batched(g: TemporalGraph, req: list, budget_bytes: int = 8 << 30, keep=None,
                  device_out: bool = False) -> dict:
    """Continuous batching on one handle.  `keep` (a dict) receives the
    results per request index when given (for identity checks).

    The kinds with work share the GPU by weighted fair queuing: each turn
    runs one action (a server step, a DC batch, or one plain call) of the
    kind whose virtual time is least, and charges it the action's measured
    time divided by the number of queries the action advanced -- so every
    query in the system gets an equal share of device time, and a kind of
    cheap steps (BFS levels) is not held to the pace of costly ones (PR
    iterations, CC round groups).  A kind that becomes busy again starts at
    the virtual time of the last action, with no credit for its idle time."""
    sink = _sink(g, device_out)
    pr_sink = sink
    if device_out:
        import torch
        pr_sink = lambda: torch.empty(g.n, dtype=torch.float32, device=f"cuda:{g.device}")
    done_at = [0.0] * len(req)
    order = sorted(range(len(req)), key=lambda i: req[i].arrival)
    kinds = {r.kind for r in req}
    servers, charged = open_servers(g, kinds, budget_bytes)
    dc_rows = charged.get("DC", 8 * g.n * 64) // (8 * g.n)
    queue = {k: deque() for k in ("BFS", "CC", "PR", "DC")}
    lane_of = {k: {} for k in servers}
    vt = {k: 0.0 for k in queue}
    vnow = 0.0
    busy = set()
    left = len(req)
    nxt = 0
    t0 = time.perf_counter()

    def finish(i, res):
        nonlocal left
        done_at[i] = time.perf_counter() - t0
        left -= 1
        if keep is not None:
            keep[i] = res

    def act(kind) -> int:
        """One action of `kind`; returns the number of queries it advanced."""
        if kind in servers and lane_of[kind]:
            srv, lanes = servers[kind], lane_of[kind]
            n = len(lanes)
            for lane in srv.step():
                i = lanes.pop(lane)
                finish(i, srv.result(lane, out=(pr_sink if kind == "PR" else sink)()))
            return n
        q = queue[kind]
        if kind == "DC" and len(q) >= MIN_BATCH["DC"]:
            idx = [q.popleft() for _ in range(min(dc_rows, len(q)))]
            for i, row in zip(idx, _dc_batch(g, req, idx, device_out)):
                finish(i, row)
            return len(idx)
        i = q.popleft()
        finish(i, _plain_one(g, req[i], device_out))
        return 1

    try:
        while left:
            now = time.perf_counter() - t0
            while nxt < len(order) and req[order[nxt]].arrival <= now:
                queue[req[order[nxt]].kind].append(order[nxt])
                nxt += 1
            for kind, srv in servers.items():  # admission in arrival order per kind
                q, lanes = queue[kind], lane_of[kind]
                if not lanes and len(q) < SERVE_FROM[kind]:
                    continue  # idle server, short queue: plain calls
                while q and len(lanes) < srv.LANES:
                    i = q.popleft()
                    lanes[_admit(srv, kind, req[i])] = i
            ready = [k for k in queue if queue[k] or (k in lane_of and lane_of[k])]
            if not ready:  # idle until the next arrival
                time.sleep(max(0.0, req[order[nxt]].arrival - (time.perf_counter() - t0)))
                continue
            for k in ready:
                if k not in busy:
                    vt[k] = max(vt[k], vnow)
            busy = set(ready)
            kind = min(ready, key=lambda k: vt[k])
            t = time.perf_counter()
            n = act(kind)
            vt[kind] += (time.perf_counter() - t) / max(1, n)
            vnow = vt[kind]
    finally:
        for srv in servers.values():
            srv.close()
    out = _stats(req, done_at)
    out["lanes"] = {k: srv.LANES for k, srv in servers.items()}
    out["state_bytes"] = charged
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
