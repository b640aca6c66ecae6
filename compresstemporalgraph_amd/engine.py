"""Host-side mirror of the reference's workload surface over ``libtga.so``.

The reference runs one workload per executable (gpu_cuda/<Dataset>/<algo>.cu)
and its CPU driver names them ``RA BFS CC PR HITS DC GC``
(cpu_java/.../BaselineRunner.java:50-88).  ``TemporalGraph`` opens one dataset
on one GPU and exposes the seven entry points of ``include/tga.h`` with the
same names; ``run(workload)`` accepts that CPU driver's workload spellings.

Outputs are numpy arrays by default.  Passing ``out=`` a CUDA tensor (or any
object with ``data_ptr()`` on the graph's device) writes the result into
device memory instead (``TGA_OUT_DEVICE``), which keeps timed loops free of
host copies.

There is no CPU fallback: if ``libtga.so`` is missing or no GPU is present,
opening a graph raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

TGA_OUT_DEVICE = 1
FULL_WINDOW = (-(1 << 31), (1 << 31) - 1)
INF32 = 0x7FFFFFFF
TGA_FROM_FILE = -(1 << 63)

_LIB = None


class EngineError(RuntimeError):
    pass


class Window(C.Structure):
    _fields_ = [("start", C.c_int64), ("end", C.c_int64)]


class DatasetDesc(C.Structure):
    _fields_ = [("max_label", C.c_int64), ("active_begin", C.c_int64), ("active_end", C.c_int64),
                ("min_timestamp", C.c_int64)]


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("contacts", C.c_int64), ("min_timestamp", C.c_int64),
                ("graph_bytes", C.c_uint64), ("timestamps_bytes", C.c_uint64), ("index_bytes", C.c_uint64),
                ("l_graph", C.c_int32), ("l_timestamps", C.c_int32), ("active_begin", C.c_int64),
                ("active_end", C.c_int64), ("device_bytes", C.c_int64), ("max_degree", C.c_int64),
                ("chunks", C.c_int64)]


class TuneReport(C.Structure):
    _fields_ = [("candidates", C.c_int32), ("burst", C.c_int32), ("flat_stride", C.c_int32),
                ("setting", C.c_int32 * 8), ("pass_ms", C.c_double * 8)]


def lib():
    """Loads libtga.so (built in-tree by ``build.build_engine``); raises if absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(_build.ENGINE_LIB):
            raise EngineError(f"{_build.ENGINE_LIB} is not built; run compresstemporalgraph_amd.build")
        L = C.CDLL(_build.ENGINE_LIB)
        vp, p32, p64, pu64 = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)
        L.tga_last_error.restype = C.c_char_p
        L.tga_open.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(DatasetDesc), C.c_int, C.POINTER(vp)]
        L.tga_open_dir.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(DatasetDesc), C.c_int, C.POINTER(vp)]
        L.tga_close.argtypes = [vp]
        L.tga_close.restype = None
        L.tga_info.argtypes = [vp, C.POINTER(GraphInfo)]
        L.tga_neighbor_access.argtypes = [vp, p32, C.c_int64, pu64, p64]
        L.tga_degree_centrality.argtypes = [vp, Window, vp, vp, C.c_uint32]
        L.tga_degree_centrality_batch.argtypes = [vp, C.POINTER(Window), C.c_int32, vp, vp, C.c_uint32]
        L.tga_bfs.argtypes = [vp, Window, C.c_int32, vp, C.c_uint32]
        L.tga_bfs_batch.argtypes = [vp, C.POINTER(Window), p32, C.c_int32, vp, C.c_uint32]
        L.tga_bfs_server_open.argtypes = [vp, C.POINTER(vp)]
        L.tga_bfs_server_admit.argtypes = [vp, C.c_int32, Window, p32]
        L.tga_bfs_server_step.argtypes = [vp, pu64]
        L.tga_bfs_server_result.argtypes = [vp, C.c_int32, vp, C.c_uint32]
        L.tga_bfs_server_close.argtypes = [vp]
        L.tga_bfs_server_close.restype = None
        L.tga_cc_server_open.argtypes = [vp, C.POINTER(vp)]
        L.tga_cc_server_admit.argtypes = [vp, Window, p32]
        L.tga_cc_server_step.argtypes = [vp, pu64]
        L.tga_cc_server_result.argtypes = [vp, C.c_int32, vp, C.c_uint32]
        L.tga_cc_server_close.argtypes = [vp]
        L.tga_cc_server_close.restype = None
        L.tga_pr_server_open.argtypes = [vp, C.c_float, C.POINTER(vp)]
        L.tga_pr_server_admit.argtypes = [vp, Window, C.c_int32, p32]
        L.tga_pr_server_step.argtypes = [vp, pu64]
        L.tga_pr_server_result.argtypes = [vp, C.c_int32, vp, C.c_uint32]
        L.tga_pr_server_close.argtypes = [vp]
        L.tga_pr_server_close.restype = None
        L.tga_earliest_arrival.argtypes = [vp, Window, C.c_int32, vp, C.c_uint32]
        L.tga_latest_departure.argtypes = [vp, Window, C.c_int32, vp, vp, C.c_uint32]
        L.tga_connected_components.argtypes = [vp, Window, vp, C.c_uint32]
        L.tga_connected_components_batch.argtypes = [vp, C.POINTER(Window), C.c_int32, vp, C.c_uint32]
        L.tga_graph_coloring.argtypes = [vp, Window, C.c_int32, vp, vp, C.c_uint32]
        L.tga_pagerank.argtypes = [vp, Window, C.c_int32, C.c_float, C.c_float, vp, p32, C.c_uint32]
        L.tga_pagerank_batch.argtypes = [vp, C.POINTER(Window), C.c_int32, C.c_int32, C.c_float, vp, C.c_uint32]
        L.tga_hits.argtypes = [vp, Window, C.c_int32, vp, vp, C.c_uint32]
        L.tga_hits_batch.argtypes = [vp, C.POINTER(Window), C.c_int32, C.c_int32, vp, vp, C.c_uint32]
        L.tga_profile.argtypes = [vp, C.c_int]
        L.tga_profile_read.argtypes = [vp, C.POINTER(C.c_double), p64, C.c_int]
        L.tga_encode_timestamps.argtypes = [p32, p64, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.POINTER(vp), pu64]
        L.tga_encode_gamma_gaps.argtypes = [pu64, C.c_int64, C.c_int, C.POINTER(vp)]
        L.tga_encode_ef_index.argtypes = [pu64, C.c_int64, C.c_int32, C.c_int, C.POINTER(vp), C.POINTER(vp)]
        L.tga_stream_bytes.argtypes = [vp]
        L.tga_stream_bytes.restype = C.c_uint64
        L.tga_stream_read.argtypes = [vp, C.POINTER(C.c_uint8), C.c_uint64]
        L.tga_stream_free.argtypes = [vp]
        L.tga_stream_free.restype = None
        L.tga_debug_fail_alloc.argtypes = [vp, C.c_int64]
        L.tga_autotune.argtypes = [vp, C.c_int32, C.POINTER(TuneReport)]
        L.tga_ts_decode_bench.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        L.tga_neighbor_access_noalloc.argtypes = [vp, p32, C.c_int64, C.POINTER(C.c_uint64), p64]
        L.tga_graph_coloring_batch.argtypes = [vp, C.POINTER(Window), C.c_int32, C.c_int32, vp, vp, C.c_uint32]
        _LIB = L
    return _LIB


def _check(rc: int) -> None:
    if rc != 0:
        raise EngineError(f"tga error {rc}: {lib().tga_last_error().decode()}")


def _win(window) -> Window:
    ws, we = window if window is not None else FULL_WINDOW
    return Window(int(ws), int(we))


class _Out:
    """Resolves an output buffer: a fresh numpy array, or a caller's device tensor."""

    def __init__(self, n, dtype, out=None):
        if out is None:
            self.arr = np.empty(n, dtype)
            self.ptr = self.arr.ctypes.data_as(C.c_void_p)
            self.flags = 0
        else:
            self.arr = out
            self.ptr = C.c_void_p(out.data_ptr())
            self.flags = TGA_OUT_DEVICE


WORKLOADS = ("RA", "BFS", "CC", "PR", "HITS", "DC", "GC")


def parse_workload(value: str) -> str:
    """Workload spellings of BaselineRunner.Workload.parse (BaselineRunner.java:60-88)."""
    v = value.upper().replace("-", "").replace("_", "")
    table = {"RA": "RA", "RANDOMACCESS": "RA", "BFS": "BFS", "BREADTHFIRSTSEARCH": "BFS", "CC": "CC",
             "CONNECTEDCOMPONENTS": "CC", "PR": "PR", "PAGERANK": "PR", "HITS": "HITS", "DC": "DC",
             "DEGREECENTRALITY": "DC", "GC": "GC", "GRAPHCOLORING": "GC"}
    if v not in table:
        raise ValueError("Unknown workload: " + value)
    return table[v]


class TemporalGraph:
    """One compressed temporal graph resident on one GPU."""

    def __init__(self, directory: str, name: str, device: int = 0, max_label: int = 0, active=None,
                 min_timestamp: int = TGA_FROM_FILE):
        ab, ae = active if active is not None else (0, 0)
        desc = DatasetDesc(max_label, ab, ae, min_timestamp)
        h = C.c_void_p()
        _check(lib().tga_open_dir(directory.encode(), name.encode(), C.byref(desc), device, C.byref(h)))
        self.h = h
        self.device = device
        info = GraphInfo()
        _check(lib().tga_info(self.h, C.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in GraphInfo._fields_}
        self.n = info.n

    @classmethod
    def open(cls, root: str, dataset: str, **kw) -> "TemporalGraph":
        """Opens root/data/compressed/<dataset>/ (the reference's layout)."""
        return cls(os.path.join(root, "data", "compressed", dataset), dataset, **kw)

    def close(self):
        if getattr(self, "h", None):
            lib().tga_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def profile(self, enable: bool = True) -> None:
        """Times every decode-kernel launch with HIP events on the handle's stream."""
        _check(lib().tga_profile(self.h, int(enable)))

    def profile_read(self, reset: bool = True):
        """Returns (summed kernel milliseconds, launches) since the last reset."""
        ms, n = C.c_double(), C.c_int64()
        _check(lib().tga_profile_read(self.h, C.byref(ms), C.byref(n), int(reset)))
        return ms.value, n.value

    def device_bytes(self) -> int:
        """Device memory the handle holds now (grows with first-use state)."""
        info = GraphInfo()
        _check(lib().tga_info(self.h, C.byref(info)))
        return int(info.device_bytes)

    def autotune(self, passes: int = 3) -> dict:
        """Times the decode walk's burst-length candidates over full
        neighbor-access passes and keeps the fastest (tga_autotune); returns
        the report."""
        rep = TuneReport()
        _check(lib().tga_autotune(self.h, passes, C.byref(rep)))
        return {"burst": rep.burst, "flat_stride": rep.flat_stride,
                "pass_ms": {int(rep.setting[i]): round(rep.pass_ms[i], 4) for i in range(rep.candidates)}}

    def debug_fail_alloc(self, nth: int) -> None:
        """Error-path test hook: the nth device allocation from now fails with
        TGA_ENOMEM (0 turns it off)."""
        _check(lib().tga_debug_fail_alloc(self.h, int(nth)))

    # -- the seven entry points -------------------------------------------
    def ts_decode_bench(self, method: int, passes: int = 5):
        """(ms per pass, sum of the decoded times) of a timestamp decode
        method over the flat records (tga_ts_decode_bench): 0 one lane per
        record from its column copy, 1 one wave per record, speculative."""
        ms, sm = C.c_double(), C.c_uint64()
        _check(lib().tga_ts_decode_bench(self.h, method, passes, C.byref(ms), C.byref(sm)))
        return ms.value, sm.value

    def neighbor_access(self, verts=None):
        """Returns (checksum, contacts) over the listed vertices (all active when None)."""
        cs, cnt = C.c_uint64(), C.c_int64()
        if verts is None:
            _check(lib().tga_neighbor_access(self.h, None, 0, C.byref(cs), C.byref(cnt)))
        else:
            v = np.ascontiguousarray(verts, np.int32)
            _check(lib().tga_neighbor_access(self.h, v.ctypes.data_as(C.POINTER(C.c_int32)), len(v), C.byref(cs),
                                             C.byref(cnt)))
        return cs.value, cnt.value

    def neighbor_access_noalloc(self, verts=None):
        """neighbor_access through the entry point that allocates nothing (a
        fixed per-GPU device array; at most 65536 listed vertices)."""
        cs, cnt = C.c_uint64(), C.c_int64()
        if verts is None:
            _check(lib().tga_neighbor_access_noalloc(self.h, None, 0, C.byref(cs), C.byref(cnt)))
        else:
            v = np.ascontiguousarray(verts, np.int32)
            _check(lib().tga_neighbor_access_noalloc(self.h, v.ctypes.data_as(C.POINTER(C.c_int32)), len(v),
                                                     C.byref(cs), C.byref(cnt)))
        return cs.value, cnt.value

    def degree_centrality(self, window=None, centrality=False, out=None, out_centrality=None):
        o = _Out(self.n, np.int32, out)
        c = _Out(self.n, np.int32, out_centrality) if (centrality or out_centrality is not None) else None
        _check(lib().tga_degree_centrality(self.h, _win(window), o.ptr, c.ptr if c else None, o.flags))
        return (o.arr, c.arr) if c else o.arr

    def degree_centrality_batch(self, windows, centrality=True, out=None, out_centrality=None):
        """Batched degree centrality: up to 64 windows in one call.  Returns
        (nq, n) int32 counts (and centralities) whose row q equals
        degree_centrality(windows[q], centrality=...); `out` / `out_centrality`
        may be a caller's (nq, n) device tensors instead."""
        nq = len(windows)
        wins = (Window * nq)(*[_win(w) for w in windows])
        if out is None:
            cnt = np.empty((nq, self.n), np.int32)
            cen = np.empty((nq, self.n), np.int32) if centrality else None
            pc, pe, flags = cnt.ctypes.data_as(C.c_void_p), cen.ctypes.data_as(C.c_void_p) if centrality else None, 0
        else:
            cnt, cen = out, (out_centrality if centrality else None)
            pc, pe, flags = C.c_void_p(out.data_ptr()), C.c_void_p(cen.data_ptr()) if centrality else None, TGA_OUT_DEVICE
        _check(lib().tga_degree_centrality_batch(self.h, wins, nq, pc, pe, flags))
        return (cnt, cen) if centrality else cnt

    def bfs(self, src: int, window=None, out=None):
        o = _Out(self.n, np.int32, out)
        _check(lib().tga_bfs(self.h, _win(window), src, o.ptr, o.flags))
        return o.arr

    def bfs_batch(self, srcs, windows=None, out=None):
        """Batched BFS: up to 64 (source, window) queries in one call.  Returns
        an (nq, n) int32 array whose row q equals bfs(srcs[q], windows[q]); `out`
        may be a caller's (nq, n) device tensor instead."""
        srcs = np.ascontiguousarray(srcs, np.int32)
        nq = len(srcs)
        wins = (Window * nq)(*[_win(w) for w in (windows if windows is not None else [None] * nq)])
        if out is None:
            arr = np.empty((nq, self.n), np.int32)
            ptr, flags = arr.ctypes.data_as(C.c_void_p), 0
        else:
            arr, ptr, flags = out, C.c_void_p(out.data_ptr()), TGA_OUT_DEVICE
        _check(lib().tga_bfs_batch(self.h, wins, srcs.ctypes.data_as(C.POINTER(C.c_int32)), nq, ptr, flags))
        return arr

    def bfs_server(self):
        """Opens a BFS server (continuous batching of up to 64 queries)."""
        return BfsServer(self)

    def cc_server(self):
        """Opens a connected-components server (continuous admission of up to
        64 window queries into the batched label columns)."""
        return CcServer(self)

    def pagerank_server(self, damping: float = 0.85):
        """Opens a PageRank server (continuous admission of up to 64 window
        queries, one iteration of every running query per step)."""
        return PrServer(self, damping)

    def earliest_arrival(self, src: int, window=None, out=None):
        o = _Out(self.n, np.int32, out)
        _check(lib().tga_earliest_arrival(self.h, _win(window), src, o.ptr, o.flags))
        return o.arr

    def latest_departure(self, target: int, window=None, out=None, rounds=False):
        """Latest departure to `target`: per vertex, the latest first-contact
        time of a time-respecting path to the target inside the window
        (-1073741823 = -(INT32_MAX // 2) where none exists, the value temporal_ligra's
        LatestDeparture-Compute.C:116-119 leaves; the window end at the target
        itself).
        With rounds=True also returns the number of pull passes run."""
        o = _Out(self.n, np.int32, out)
        nr = C.c_int32(0)
        _check(lib().tga_latest_departure(self.h, _win(window), target, o.ptr, C.byref(nr), o.flags))
        return (o.arr, nr.value) if rounds else o.arr

    def connected_components(self, window=None, out=None):
        o = _Out(self.n, np.int32, out)
        _check(lib().tga_connected_components(self.h, _win(window), o.ptr, o.flags))
        return o.arr

    def connected_components_batch(self, windows, out=None):
        """Batched connected components: up to 64 windows in one call.  Returns
        an (nq, n) int32 array whose row q equals connected_components(windows[q]);
        `out` may be a caller's (nq, n) device tensor instead."""
        nq = len(windows)
        wins = (Window * nq)(*[_win(w) for w in windows])
        if out is None:
            arr = np.empty((nq, self.n), np.int32)
            ptr, flags = arr.ctypes.data_as(C.c_void_p), 0
        else:
            arr, ptr, flags = out, C.c_void_p(out.data_ptr()), TGA_OUT_DEVICE
        _check(lib().tga_connected_components_batch(self.h, wins, nq, ptr, flags))
        return arr

    def graph_coloring_batch(self, windows, rounds: int = 512, priority=None, out=None):
        """Batched graph coloring: up to 64 windows in one call (one priority
        array for all).  Returns an (nq, n) int32 array whose row q equals
        graph_coloring(windows[q], rounds, priority); `out` may be a caller's
        (nq, n) device tensor instead (then `priority`, if given, is a device
        tensor too)."""
        nq = len(windows)
        wins = (Window * nq)(*[_win(w) for w in windows])
        if out is None:
            arr = np.empty((nq, self.n), np.int32)
            ptr, flags = arr.ctypes.data_as(C.c_void_p), 0
        else:
            arr, ptr, flags = out, C.c_void_p(out.data_ptr()), TGA_OUT_DEVICE
        pr = None
        if priority is not None:
            if flags & TGA_OUT_DEVICE:
                pr = C.c_void_p(priority.data_ptr())
            else:
                priority = np.ascontiguousarray(priority, np.int32)
                pr = priority.ctypes.data_as(C.c_void_p)
        _check(lib().tga_graph_coloring_batch(self.h, wins, nq, rounds, pr, ptr, flags))
        return arr

    def graph_coloring(self, window=None, rounds: int = 512, priority=None, out=None):
        o = _Out(self.n, np.int32, out)
        pr = None
        if priority is not None:
            if o.flags & TGA_OUT_DEVICE:
                pr = C.c_void_p(priority.data_ptr())
            else:
                priority = np.ascontiguousarray(priority, np.int32)
                pr = priority.ctypes.data_as(C.c_void_p)
        _check(lib().tga_graph_coloring(self.h, _win(window), rounds, pr, o.ptr, o.flags))
        return o.arr

    def pagerank(self, window=None, iters: int = 20, damping: float = 0.85, tol: float = 0.0, out=None):
        o = _Out(self.n, np.float32, out)
        it = C.c_int32()
        _check(lib().tga_pagerank(self.h, _win(window), iters, damping, tol, o.ptr, C.byref(it), o.flags))
        self.last_iterations = it.value
        return o.arr

    def pagerank_batch(self, windows, iters: int = 20, damping: float = 0.85, out=None):
        """Batched PageRank: up to 64 windows advanced together (FP32 columns,
        fixed iterations).  Returns an (nq, n) float32 array whose row q equals
        pagerank(windows[q], iters, damping) up to FP64 summation order; `out`
        may be a caller's (nq, n) device tensor instead."""
        nq = len(windows)
        wins = (Window * nq)(*[_win(w) for w in windows])
        if out is None:
            arr = np.empty((nq, self.n), np.float32)
            ptr, flags = arr.ctypes.data_as(C.c_void_p), 0
        else:
            arr, ptr, flags = out, C.c_void_p(out.data_ptr()), TGA_OUT_DEVICE
        _check(lib().tga_pagerank_batch(self.h, wins, nq, iters, damping, ptr, flags))
        return arr

    def hits_batch(self, windows, iters: int = 20, out_hub=None, out_auth=None):
        """Batched HITS: up to 64 windows together (FP32 columns).  Returns
        (hub, auth), each (nq, n) float32, row q equal to hits(windows[q], iters)
        up to FP64 summation order; out_hub / out_auth may be (nq, n) device
        tensors instead."""
        nq = len(windows)
        wins = (Window * nq)(*[_win(w) for w in windows])
        if out_hub is None:
            hub, auth = np.empty((nq, self.n), np.float32), np.empty((nq, self.n), np.float32)
            ph, pa, flags = hub.ctypes.data_as(C.c_void_p), auth.ctypes.data_as(C.c_void_p), 0
        else:
            hub, auth = out_hub, out_auth
            ph, pa, flags = C.c_void_p(hub.data_ptr()), C.c_void_p(auth.data_ptr()), TGA_OUT_DEVICE
        _check(lib().tga_hits_batch(self.h, wins, nq, iters, ph, pa, flags))
        return hub, auth

    def hits(self, window=None, iters: int = 20, out_hub=None, out_auth=None):
        h = _Out(self.n, np.float32, out_hub)
        a = _Out(self.n, np.float32, out_auth)
        _check(lib().tga_hits(self.h, _win(window), iters, h.ptr, a.ptr, h.flags))
        return h.arr, a.arr

    def run(self, workload: str, **kw):
        """Runs a workload by its BaselineRunner name (RA, BFS, CC, PR, HITS, DC, GC)."""
        w = parse_workload(workload)
        if w == "RA":
            return self.neighbor_access(kw.get("verts"))
        if w == "BFS":
            return self.bfs(kw.get("src", 0), kw.get("window"))
        if w == "CC":
            return self.connected_components(kw.get("window"))
        if w == "PR":
            return self.pagerank(kw.get("window"), kw.get("iters", 20), kw.get("damping", 0.85), kw.get("tol", 0.0))
        if w == "HITS":
            return self.hits(kw.get("window"), kw.get("iters", 20))
        if w == "DC":
            return self.degree_centrality(kw.get("window"))
        return self.graph_coloring(kw.get("window"), kw.get("rounds", 512))


class _LaneServer:
    """Continuous batching on one TemporalGraph: admit() puts a query into a
    free lane, step() advances every running query and returns the lanes
    that finished, result(lane) returns a finished query's output (equal to
    the plain call's) and frees the lane."""

    LANES = 64
    _kind = ""
    _dtype = np.int32

    def __init__(self, graph: "TemporalGraph", *open_args):
        self.graph = graph
        self.h = C.c_void_p()
        _check(getattr(lib(), f"tga_{self._kind}_server_open")(graph.h, *open_args, C.byref(self.h)))

    def _admit(self, *args) -> int:
        lane = C.c_int32()
        _check(getattr(lib(), f"tga_{self._kind}_server_admit")(self.h, *args, C.byref(lane)))
        return lane.value

    def step(self) -> list:
        fin = C.c_uint64()
        _check(getattr(lib(), f"tga_{self._kind}_server_step")(self.h, C.byref(fin)))
        return [q for q in range(self.LANES) if fin.value >> q & 1]

    def result(self, lane: int, out=None):
        o = _Out(self.graph.n, self._dtype, out)
        _check(getattr(lib(), f"tga_{self._kind}_server_result")(self.h, int(lane), o.ptr, o.flags))
        return o.arr

    def close(self) -> None:
        if self.h:
            getattr(lib(), f"tga_{self._kind}_server_close")(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class BfsServer(_LaneServer):
    """BFS queries: each step advances every running query one level in a
    single decode pass; result() equals TemporalGraph.bfs."""

    _kind = "bfs"

    def admit(self, src: int, window=None) -> int:
        return self._admit(int(src), _win(window))


class CcServer(_LaneServer):
    """Connected-components queries: each step runs a group of label rounds
    over the batched run structure; a query finishes at its fixpoint and
    result() equals TemporalGraph.connected_components."""

    _kind = "cc"

    def admit(self, window=None) -> int:
        return self._admit(_win(window))


class PrServer(_LaneServer):
    """PageRank queries (the server's damping, each query its iteration
    count): each step runs one iteration of every running query; result()
    equals the query's row of TemporalGraph.pagerank_batch."""

    _kind = "pr"
    _dtype = np.float32

    def __init__(self, graph: "TemporalGraph", damping: float = 0.85):
        super().__init__(graph, C.c_float(damping))

    def admit(self, window=None, iters: int = 20) -> int:
        return self._admit(_win(window), int(iters))


# ------------------------------------------------------------ GPU encoder --

def _take_stream(h) -> bytes:
    """Reads an encoded device stream into bytes and frees it."""
    try:
        nbytes = lib().tga_stream_bytes(h)
        buf = np.empty(max(int(nbytes), 1), np.uint8)
        _check(lib().tga_stream_read(h, buf.ctypes.data_as(C.POINTER(C.c_uint8)), nbytes))
        return buf[:nbytes].tobytes()
    finally:
        lib().tga_stream_free(h)


def encode_timestamps(src, t, min_ts: int, n: int, device: int = 0):
    """The .timestamps stream of contacts sorted by (src, dst, t), written on
    the GPU (tga_encode_timestamps).  Returns (bytes, offsets): the stream and
    the n + 1 bit offsets of every vertex's first code and of the end."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    t = np.ascontiguousarray(t, dtype=np.int64)
    off = np.empty(n + 1, np.uint64)
    h = C.c_void_p()
    _check(lib().tga_encode_timestamps(src.ctypes.data_as(C.POINTER(C.c_int32)), t.ctypes.data_as(C.POINTER(C.c_int64)),
                                       len(src), int(min_ts), int(n), device, C.byref(h),
                                       off.ctypes.data_as(C.POINTER(C.c_uint64))))
    return _take_stream(h), off


def encode_gamma_gaps(values, device: int = 0) -> bytes:
    """gamma codes of values[0] and of every later difference (the .offsets
    stream of record bit offsets), written on the GPU."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    h = C.c_void_p()
    _check(lib().tga_encode_gamma_gaps(v.ctypes.data_as(C.POINTER(C.c_uint64)), len(v), device, C.byref(h)))
    return _take_stream(h)


def encode_ef_index(values, l: int, device: int = 0):
    """The GPU index files (lower, upper) of non-decreasing offsets with
    lower-bit width l, written on the GPU."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    lo, up = C.c_void_p(), C.c_void_p()
    _check(lib().tga_encode_ef_index(v.ctypes.data_as(C.POINTER(C.c_uint64)), len(v), int(l), device, C.byref(lo),
                                     C.byref(up)))
    return _take_stream(lo), _take_stream(up)


def encode_gpu(src, dst, t, directory: str, name: str, n_min: int = 0, device: int = 0):
    """Writes the dataset files ``directory/name.*`` as fmt.encode does, byte
    for byte: the graph stream and .properties on the host (the record
    planner, fmt.encode_graph), then .offsets, .timestamps and the four GPU
    index files on the GPU.  Returns the format kit's statistics."""
    from dataclasses import replace
    from . import fmt
    t = np.ascontiguousarray(t, dtype=np.int64)
    goff, st = fmt.encode_graph(src, dst, directory, name, n_min=n_min)
    n = len(goff) - 1
    tmin = int(t.min())
    ts, toff = encode_timestamps(src, t, tmin, n, device)
    offs = encode_gamma_gaps(goff, device)
    graph_bytes = os.path.getsize(os.path.join(directory, name + ".graph"))
    lg = fmt.ef_low_width(graph_bytes, n)
    lt = fmt.ef_low_width(len(ts), n)
    glo, gup = encode_ef_index(goff, lg, device)
    tlo, tup = encode_ef_index(toff, lt, device)
    for fname, data in ((name + ".timestamps", ts), (name + ".offsets", offs), ("graph_lower_bits.bin", glo),
                        ("graph_upper_bits.bin", gup), ("timestamp_lower_bits.bin", tlo),
                        ("timestamp_upper_bits.bin", tup)):
        with open(os.path.join(directory, fname), "wb") as f:
            f.write(data)
    return replace(st, min_timestamp=tmin, timestamps_bits=int(toff[-1]), l_graph=lg, l_timestamps=lt)
