"""Raw-contact brute force — TEST INFRASTRUCTURE ONLY.

Every function here works on the raw contact list ``(src, dst, t)`` sorted by
(src, dst, t), never on the compressed streams, so it checks the C oracle's
decoder and algorithm code independently.  The rules are the algorithm
contracts stated in oracle/oracle.c (BaselineRunner.java:273-608 and
gpu_cuda/shared per-contact blocks, made deterministic).
"""
from __future__ import annotations

import numpy as np

INF32 = 0x7FFFFFFF
FULL_WINDOW = (-(1 << 31), (1 << 31) - 1)
LD_UNREACHED = -(((1 << 31) - 1) // 2)  # -1073741823


def contact_hashes(src, dst, t) -> np.ndarray:
    """Per-contact checksum terms: h = (t ^ r*0x9E3779B1 ^ x*0x85EBCA77) *
    0x7FEB352D (uint32 arithmetic)."""
    x = np.asarray(src).astype(np.uint32)
    r = np.asarray(dst).astype(np.uint32)
    tt = np.asarray(t).astype(np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        h = tt ^ (r * np.uint32(0x9E3779B1)) ^ (x * np.uint32(0x85EBCA77))
        h *= np.uint32(0x7FEB352D)
    return h.astype(np.uint64)


def checksum(src, dst, t) -> int:
    h = contact_hashes(src, dst, t)
    return int(np.sum(h, dtype=np.uint64))


class Raw:
    def __init__(self, src, dst, t, n, window=None, active=None):
        self.n = n
        src = np.asarray(src, np.int64)
        dst = np.asarray(dst, np.int64)
        t = np.asarray(t, np.int64)
        ws, we = window if window is not None else FULL_WINDOW
        ab, ae = active if active is not None else (0, n)
        self.ab, self.ae = ab, ae
        self.active = np.zeros(n, bool)
        self.active[ab:min(ae, n)] = True
        self.src, self.dst, self.t = src, dst, t
        self.window = (ws, we)
        self.inw = (t >= ws) & (t <= we)
        # window edges between active vertices
        keep = self.inw & self.active[src] & self.active[dst]
        self.esrc, self.edst, self.et = src[keep], dst[keep], t[keep]

    # -- DC
    def degree(self):
        keep = self.inw & self.active[self.dst] & self.active[self.src]
        return np.bincount(self.src[keep], minlength=self.n).astype(np.int32)

    # -- BFS levels
    def bfs(self, s):
        level = np.full(self.n, INF32, np.int64)
        level[s] = 0
        if not self.active[s]:
            return level.astype(np.int32)
        frontier = np.array([s])
        lvl = 0
        order = np.argsort(self.esrc, kind="stable")
        es, ed = self.esrc[order], self.edst[order]
        starts = np.searchsorted(es, np.arange(self.n + 1))
        while len(frontier):
            nxt = np.concatenate([ed[starts[x]:starts[x + 1]] for x in frontier]) if len(frontier) else np.array([])
            nxt = np.unique(nxt.astype(np.int64))
            nxt = nxt[level[nxt] == INF32]
            lvl += 1
            level[nxt] = lvl
            frontier = nxt
        return level.astype(np.int32)

    # -- latest departure to target z: the latest t1 of a path x -t1-> x1 -t2->
    #    ... -> z over window contacts between active vertices with
    #    t1 <= t2 <= ...; ld[z] = the window end (clamped to int32, also for an
    #    inactive z, whose in-contacts are dropped); LD_UNREACHED where z is
    #    unreachable (-1073741823, temporal_ligra LatestDeparture-Compute.C:116-119).
    #    Fixpoint: ld[x] = max over contacts (x, r, t) with r reached and
    #    t <= ld[r] of t.
    def latest_departure(self, z):
        neg = LD_UNREACHED
        ld = np.full(self.n, neg, np.int64)
        ld[z] = min(self.window[1], (1 << 31) - 1)
        while True:
            lr = ld[self.edst]
            ok = (lr != neg) & (self.et <= lr)
            cand = np.full(self.n, neg, np.int64)
            np.maximum.at(cand, self.esrc[ok], self.et[ok])
            new = np.maximum(ld, cand)
            if np.array_equal(new, ld):
                return ld.astype(np.int32)
            ld = new

    # -- earliest arrival from s (the GPU BFS rule of bfs.cu:1601-1619 as a
    #    deterministic fixpoint): arr[s] = max(window start, 0); each distinct
    #    window pair (x, r) between active vertices is used with its FIRST
    #    window contact t only (the reference skips repeated neighbors,
    #    bfs.cu:1452), and gives r the arrival t when arr[x] <= t < arr[r].
    def earliest_arrival(self, s):
        arr = np.full(self.n, INF32, np.int64)
        arr[s] = max(self.window[0], 0)
        if not self.active[s]:
            return arr.astype(np.int32)
        es, ed, et = self.esrc, self.edst, self.et  # sorted by (src, dst, t)
        first = np.ones(len(es), bool)
        first[1:] = (es[1:] != es[:-1]) | (ed[1:] != ed[:-1])
        ps, pd, pt = es[first], ed[first], et[first]
        while True:
            ok = (arr[ps] <= pt) & (pt < arr[pd])
            cand = np.full(self.n, INF32, np.int64)
            np.minimum.at(cand, pd[ok], pt[ok])
            new = np.minimum(arr, cand)
            if np.array_equal(new, arr):
                return arr.astype(np.int32)
            arr = new

    # -- CC min-label by fixpoint propagation
    def cc(self):
        label = np.where(self.active, np.arange(self.n), -1).astype(np.int64)
        while True:
            cand = np.full(self.n, np.iinfo(np.int64).max)
            np.minimum.at(cand, self.edst, label[self.esrc])
            new = np.where(self.active, np.minimum(label, cand), -1)
            if np.array_equal(new, label):
                return label.astype(np.int32)
            label = new

    # -- GC rounds
    def gc(self, prio, rounds=512):
        prio = np.asarray(prio, np.int64)
        color = np.full(self.n, INF32, np.int64)
        u, v = self.edst, self.esrc  # v's out-neighbor u
        higher = (prio[u] > prio[v]) | ((prio[u] == prio[v]) & (u < v))
        for c in range(rounds):
            unc = self.active & (color == INF32)
            if not unc.any():
                break
            blocked = np.zeros(self.n, bool)
            sel = higher & unc[u] & unc[v]
            blocked[v[sel]] = True
            take = unc & ~blocked
            color[take] = c
        return color.astype(np.int32)

    # -- distinct window pairs with their first window contact
    def _pairs(self):
        keep = self.inw & self.active[self.src]
        s, d, t = self.src[keep], self.dst[keep], self.t[keep]
        first = np.ones(len(s), bool)
        first[1:] = (s[1:] != s[:-1]) | (d[1:] != d[:-1])
        return s[first], d[first], t[first], np.bincount(s, minlength=self.n)

    def pagerank(self, iters=20, damping=0.85, tol=0.0, dtype=np.float64):
        """Storage and per-contact arithmetic in ``dtype``; sums in float64."""
        ps, pd, pt, du = self._pairs()
        keep = self.active[pd]
        ps, pd, pt = ps[keep], pd[keep], pt[keep]
        N = int(self.active.sum())
        one = dtype(1)
        w = (one - np.exp(dtype(-0.1) * pt.astype(dtype))).astype(dtype)
        rank = np.where(self.active, one / dtype(N), dtype(0)).astype(dtype)
        base = (one - dtype(damping)) / dtype(N)
        it = 0
        for it in range(1, iters + 1):
            acc = np.where(self.active, float(base), 0.0)
            coef = np.zeros(self.n, dtype)
            nz = du > 0
            coef[nz] = (dtype(damping) * rank[nz] / du[nz].astype(dtype)).astype(dtype)
            np.add.at(acc, pd, (coef[ps] * w).astype(dtype).astype(np.float64))
            s = acc.sum()
            new = (acc / s if s > 0 else acc).astype(dtype)
            diff = np.abs(new.astype(np.float64) - rank.astype(np.float64)).sum()
            rank = new
            if tol > 0 and diff < tol:
                break
        return rank, it

    def hits(self, iters=20, dtype=np.float64):
        """Storage in ``dtype``; sums and norms in float64."""
        ps, pd, _, _ = self._pairs()
        keep = self.active[pd]
        ps, pd = ps[keep], pd[keep]
        hub = np.where(self.active, 1, 0).astype(dtype)
        auth = hub.copy()
        for _ in range(iters):
            acc = np.zeros(self.n)
            np.add.at(acc, ps, auth[pd].astype(np.float64))
            nrm = np.sqrt((acc * acc).sum())
            hub = (acc / nrm if nrm != 0 else acc).astype(dtype)
            acc = np.zeros(self.n)
            np.add.at(acc, pd, hub[ps].astype(np.float64))
            nrm = np.sqrt((acc * acc).sum())
            auth = (acc / nrm if nrm != 0 else acc).astype(dtype)
        return hub, auth
