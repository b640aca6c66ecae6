"""Benchmark: decoded temporal edges/s of the plain neighbor-access pass.

One step = one full neighbor-access (RA) pass of the GPU engine over every
contact of the workload, with the compressed graph resident in HBM.  The
headline workload is the large single-GPU configuration (``netflow25``):
netflow-like, 2^25 vertices, 2^30 contacts, streams of about 2.7 GB (beyond
the 256 MB Infinity Cache, so a pass reads HBM), generated and encoded by the
format kit before the timed region (about a minute); its line also carries
plain BFS, CC and PR over the full and a 1% window.  Beside it, under the key
``chunglu21``, the plain at-scale configuration (Chung-Lu, exponent 2.1,
2^21 vertices, 2^25 contacts, diurnal timestamps over 2^24 ticks): its RA
pass and the seven-workload table, and the batched and served query legs run
on that graph and on rmat16.  ``--workload chunglu21`` makes that graph the
headline instead (``--no-side`` drops the side measurement).

Prints one JSON line with the benchmark record's fields plus:
  roofline      the decode pass against MI355X HBM bandwidth (8 TB/s):
                algorithmic bytes per pass (|.graph| + |.timestamps| + the
                Elias-Fano index bits (n+1)(l_g+2 + l_t+2)/8) divided by the
                average pass duration measured with HIP events on the engine's
                stream (the pass is the group kernels launched concurrently);
                `traffic` is the rocprof FETCH_SIZE + WRITE_SIZE per pass, scaled
                by a copy-kernel calibration, from profiles/ when that
                measurement was taken on the current engine sources (else null).
  checksum      the decoded neighbor-access checksum, and whether it equals
                the checksum of the raw contact list the dataset was encoded
                from (a mismatch ends the run with an error).
  cpu_baseline  the CPU oracle's neighbor access (single thread, the build's
                C restatement of the reference reader, kind "port") timed on
                this host over a bounded vertex sample.
  workloads_ms  the seven entry points (plus earliest arrival and latest
                departure) over windows covering 100%, 50% and 10% of the
                timestamp span.
  batched_mix   256 window queries of the batching mix (40% BFS, 20% CC, 30%
                PR, 10% DC) as batched calls of up to 64 per kind against 256
                plain calls, on this graph and on rmat16 (results compared).
  batched_bfs   64 windowed BFS queries in one batched call against 64 plain
                calls (results compared, outputs in HBM).
  batched_dc    64 windowed degree-centrality queries in one batched call
                against 64 plain calls (results compared, outputs in HBM).
  batched_cc    64 connected-components queries in one batched call against
                64 plain calls, windowed and full windows.
  batched_pr    64 PageRank queries (20 iterations) in one batched call
                against 64 plain calls, windowed and full windows.
  batched_hits  64 HITS queries (20 iterations) in one batched call against
                64 plain calls, windowed and full windows.
  batched_gc    64 graph-coloring queries (512 rounds) in one batched call
                against 64 plain calls, windowed and full windows, on the
                at-scale graph and on rmat16.
  served_mix    Poisson arrivals of 40% BFS, 20% CC, 30% PR, 10% DC at 500/s
                and 2000/s: continuous batching against sequential plain
                execution.
  served_mix_rmat16  the same mix on rmat16 at 100/s, 1k/s and 10k/s.
  served_bfs    Poisson BFS arrivals at 1k/s and 10k/s: continuous batching
                against sequential plain BFS, queries/s and p50/p99 latency.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_file(workload: str) -> str:
    return os.path.join(ROOT, "profiles", f"k_ra_{workload}.traffic.json")


WORKLOAD_DESC = {
    "rmat16": "R-MAT (0.57, 0.19, 0.19), 2^16 vertices, 2^20 contacts, Geometric contacts per pair, "
              "uniform timestamps over 2^20 ticks",
    "chunglu21": "Chung-Lu exponent 2.1, 2^21 vertices, 2^25 contacts, diurnal timestamps over 2^24 ticks",
    "netflow25": "netflow-like, 2^25 vertices, 2^30 contacts, power-law endpoints (exponent 2.5), Pareto "
                 "contacts per pair (mean ~20), bursty Pareto(1.5) inter-arrivals over 2^30 ticks",
}


def engine_source_digest() -> str:
    """SHA-256 over the engine's HIP sources (identifies the measured build)."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "compresstemporalgraph_amd", "csrc", "engine")
    for name in sorted(os.listdir(d)):
        with open(os.path.join(d, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()


def measured_traffic(workload: str):
    """HBM bytes per pass from the committed rocprof FETCH_SIZE / WRITE_SIZE
    measurement, when it was taken on this workload with these engine sources."""
    try:
        with open(traffic_file(workload)) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("workload") != workload or rec.get("engine_sha256") != engine_source_digest():
        return None
    return rec.get("bytes_per_pass")


def workload_times(g, t_lo: int, t_hi: int, sources: int = 8) -> dict:
    """Milliseconds per call of the seven workloads over windows covering
    10%, 50% and 100% of the timestamp span (the at-scale plain
    configuration): BFS from `sources` seeded sources (mean), CC, PR (20
    iterations, FP32), HITS (20 iterations, FP32), DC (counts and the Java
    centrality), GC and RA; plus the earliest-arrival rule and latest
    departure (to the same first source).  Each call is
    timed once after one warm call."""
    import numpy as np
    rng = np.random.default_rng(11)
    span = t_hi - t_lo
    windows = {"100%": (t_lo, t_hi), "50%": (t_lo + span // 4, t_lo + span // 4 + span // 2),
               "10%": (t_lo + span // 2, t_lo + span // 2 + span // 10)}
    srcs = [int(v) for v in rng.integers(0, g.info["n"], sources)]

    def timed(f):
        f()
        t0 = time.perf_counter()
        f()
        return round((time.perf_counter() - t0) * 1e3, 3)

    out = {}
    for name, w in windows.items():
        row = {"RA": timed(lambda: g.neighbor_access()),
               "DC": timed(lambda: g.degree_centrality(w, centrality=True)),
               "BFS": round(sum(timed(lambda s=s: g.bfs(s, w)) for s in srcs) / len(srcs), 3),
               "EA": timed(lambda: g.earliest_arrival(srcs[0], w)),
               "LD": timed(lambda: g.latest_departure(srcs[0], w)),
               "CC": timed(lambda: g.connected_components(w)),
               "GC": timed(lambda: g.graph_coloring(w, rounds=512)),
               "PR": timed(lambda: g.pagerank(w, iters=20)),
               "HITS": timed(lambda: g.hits(w, iters=20))}
        out[name] = row
    return out


def large_workload_times(g, t_lo: int, t_hi: int, sources: int = 4) -> dict:
    """Milliseconds per call of plain BFS (mean over `sources` seeded
    sources), CC and PR (20 iterations, FP32) over the full window and a 1%
    window in the middle of the span, plus the RA pass (the large single-GPU
    configuration).  Results go to device buffers (HBM), as the reference's
    kernels leave them: a host copy of one 2^25-entry result over PCIe would
    take longer than most of these calls.  Each call is timed once after one
    warm call, synchronised on the device."""
    import numpy as np
    import torch
    lvl = torch.empty(g.info["n"], dtype=torch.int32, device="cuda:0")
    rank = torch.empty(g.info["n"], dtype=torch.float32, device="cuda:0")
    rng = np.random.default_rng(11)
    span = t_hi - t_lo
    windows = {"100%": (t_lo, t_hi), "1%": (t_lo + span // 2, t_lo + span // 2 + span // 100)}
    srcs = [int(v) for v in rng.integers(0, g.info["n"], sources)]

    def timed(f):
        f()
        t0 = time.perf_counter()
        f()
        return round((time.perf_counter() - t0) * 1e3, 3)

    out = {}
    for name, w in windows.items():
        out[name] = {"RA": timed(lambda: g.neighbor_access()),
                     "BFS": round(sum(timed(lambda s=s: (g.bfs(s, w, out=lvl), torch.cuda.synchronize()))
                                      for s in srcs) / len(srcs), 3),
                     "CC": timed(lambda: (g.connected_components(w, out=lvl), torch.cuda.synchronize())),
                     "PR": timed(lambda: (g.pagerank(w, iters=20, out=rank), torch.cuda.synchronize()))}
        log(f"[bench] {name} window: {out[name]}")
    return out


def large_legs(args) -> int:
    """The large configuration's plain workloads with results in HBM (child
    process: torch's HIP runtime must initialise before the engine's)."""
    import torch
    torch.zeros(1, device="cuda:0")
    from compresstemporalgraph_amd import datasets
    from compresstemporalgraph_amd.engine import TemporalGraph
    root = args.data_root or os.path.join(tempfile.gettempdir(), "tga_bench")
    desc = datasets.ensure(root, args.workload)
    g = TemporalGraph(datasets.dataset_dir(root, args.workload), args.workload, device=0)
    g.autotune(passes=2)
    out = large_workload_times(g, desc.get("t_min", 0), desc.get("t_max", 1 << 24))
    g.close()
    print(json.dumps(out), flush=True)
    return 0


def ts_variants(args) -> int:
    """Times the two timestamp decode methods of tga_ts_decode_bench over the
    flat records of --workload (one lane per record from its column copy,
    and one wave per record decoding speculatively at 64 bit offsets and
    chaining the codes by shuffles) and prints their JSON; both must decode
    the same times."""
    import torch
    torch.zeros(1, device="cuda:0")
    from compresstemporalgraph_amd import datasets
    from compresstemporalgraph_amd.engine import TemporalGraph
    root = args.data_root or os.path.join(tempfile.gettempdir(), "tga_bench")
    datasets.ensure(root, args.workload, gpu=True)
    g = TemporalGraph(datasets.dataset_dir(root, args.workload), args.workload, device=0)
    out = {"workload": args.workload}
    sums = set()
    for method, name in ((0, "lane_per_record_columns"), (1, "wave_per_record_speculative")):
        ms, sm = g.ts_decode_bench(method, passes=args.steps)
        out[name] = {"ms": round(ms, 4)}
        sums.add(sm)
        log(f"[bench] timestamp decode, {name}: {ms:.3f} ms per pass (sum {sm:#x})")
    g.close()
    out["same_times"] = len(sums) == 1
    print(json.dumps(out), flush=True)
    return 0 if len(sums) == 1 else 1


def encode_leg(args) -> int:
    """Times writing a workload's dataset files with the host encoder alone
    (fmt.encode) and with the split encoder (the graph stream on the host,
    .offsets, .timestamps and the GPU index files on the GPU:
    engine.encode_gpu), checks that the two write the same bytes, and prints
    the times as JSON."""
    import hashlib
    import shutil
    from compresstemporalgraph_amd import datasets, fmt
    from compresstemporalgraph_amd.engine import encode_gpu
    root = args.data_root or os.path.join(tempfile.gettempdir(), "tga_bench")
    w = datasets.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    src, dst, t = datasets.generate(w)
    gen_s = time.perf_counter() - t0
    out = {"workload": args.workload, "contacts": int(len(src)), "generate_s": round(gen_s, 2)}
    dirs = {}
    for how in ("host", "gpu"):
        d = os.path.join(root, "encode_" + how)
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        t0 = time.perf_counter()
        if how == "host":
            fmt.encode(src, dst, t, d, args.workload, n_min=w.n)
        else:
            encode_gpu(src, dst, t, d, args.workload, n_min=w.n)
        out[how + "_s"] = round(time.perf_counter() - t0, 2)
        log(f"[bench] {how} encoder: {out[how + '_s']} s")
        dirs[how] = d
    # the graph half alone, as the split encoder runs it on the host
    d = os.path.join(root, "encode_graph_only")
    shutil.rmtree(d, ignore_errors=True)
    t0 = time.perf_counter()
    fmt.encode_graph(src, dst, d, args.workload, n_min=w.n)
    out["host_graph_half_s"] = round(time.perf_counter() - t0, 2)
    same = True
    for f in sorted(os.listdir(dirs["host"])):
        digests = []
        for how in ("host", "gpu"):
            h = hashlib.sha256()
            with open(os.path.join(dirs[how], f), "rb") as fh:
                for chunk in iter(lambda: fh.read(1 << 24), b""):
                    h.update(chunk)
            digests.append(h.hexdigest())
        same &= digests[0] == digests[1]
    out["identical_files"] = same
    for d in list(dirs.values()) + [d]:
        shutil.rmtree(d, ignore_errors=True)
    print(json.dumps(out), flush=True)
    return 0 if same else 1


def batched_bfs(g, t_lo: int, t_hi: int, nq: int = 64) -> dict:
    """64 BFS queries with random sources and windows whose lengths are
    log-uniform in [2^10, 2^24] and starts uniform (the batching mix's BFS
    share): one tga_bfs_batch call against nq plain tga_bfs calls, outputs on
    the device in both cases; every batched row is compared with its plain
    result."""
    import numpy as np
    import torch
    rng = np.random.default_rng(5)
    n = g.info["n"]
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), nq)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), nq)
    wins = [(int(a), int(a + b)) for a, b in zip(starts, lens)]
    srcs = rng.integers(0, n, nq).astype(np.int32)
    out = torch.empty((nq, n), dtype=torch.int32, device="cuda:0")
    plain = torch.empty((nq, n), dtype=torch.int32, device="cuda:0")
    g.bfs_batch(srcs, wins, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.bfs_batch(srcs, wins, out=out)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    g.bfs(int(srcs[0]), wins[0], out=plain[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for q in range(nq):
        g.bfs(int(srcs[q]), wins[q], out=plain[q])
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    same = bool(torch.equal(out, plain))
    return {"queries": nq, "batched_ms": round(tb * 1e3, 3), "plain_ms": round(tp * 1e3, 3),
            "batched_qps": round(nq / tb, 1), "plain_qps": round(nq / tp, 1), "speedup": round(tp / tb, 2),
            "identical": same,
            "mix": "BFS, random sources, windows log-uniform in [2^10, 2^24] with uniform starts"}


def batched_dc(g, t_lo: int, t_hi: int, nq: int = 64) -> dict:
    """64 degree-centrality queries (counts and the Java centrality) over
    windows log-uniform in [2^10, 2^24] with uniform starts (the batching
    mix's DC share): one tga_degree_centrality_batch call against nq plain
    calls, outputs on the device in both cases, every row compared."""
    import numpy as np
    import torch
    rng = np.random.default_rng(7)
    n = g.info["n"]
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), nq)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), nq)
    wins = [(int(a), int(a + b)) for a, b in zip(starts, lens)]
    cnt, cen = (torch.empty((nq, n), dtype=torch.int32, device="cuda:0") for _ in range(2))
    pcnt, pcen = (torch.empty((nq, n), dtype=torch.int32, device="cuda:0") for _ in range(2))
    g.degree_centrality_batch(wins, out=cnt, out_centrality=cen)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.degree_centrality_batch(wins, out=cnt, out_centrality=cen)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    g.degree_centrality(wins[0], out=pcnt[0], out_centrality=pcen[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for q in range(nq):
        g.degree_centrality(wins[q], out=pcnt[q], out_centrality=pcen[q])
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    same = bool(torch.equal(cnt, pcnt) and torch.equal(cen, pcen))
    return {"queries": nq, "batched_ms": round(tb * 1e3, 3), "plain_ms": round(tp * 1e3, 3),
            "batched_qps": round(nq / tb, 1), "plain_qps": round(nq / tp, 1), "speedup": round(tp / tb, 2),
            "identical": same,
            "mix": "DC with centrality, windows log-uniform in [2^10, 2^24] with uniform starts"}


def batched_pr(g, t_lo: int, t_hi: int, nq: int = 64) -> dict:
    """64 PageRank queries (20 iterations, FP32) in one tga_pagerank_batch
    call against 64 plain calls, outputs in HBM: windows log-uniform in
    [2^10, 2^24] (the batching mix's PR share) and full windows.  The first
    batched call builds the per-handle run structure and is reported apart;
    rows are compared with the plain results (max relative L-inf)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(9)
    n = g.info["n"]
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), nq)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), nq)
    out = {}
    for name, wins in (("windowed", [(int(a), int(a + b)) for a, b in zip(starts, lens)]), ("full", [None] * nq)):
        bat = torch.empty((nq, n), dtype=torch.float32, device="cuda:0")
        pla = torch.empty((nq, n), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.pagerank_batch(wins, iters=20, out=bat)
        torch.cuda.synchronize()
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        g.pagerank_batch(wins, iters=20, out=bat)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        g.pagerank(wins[0], iters=20, out=pla[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q in range(nq):
            g.pagerank(wins[q], iters=20, out=pla[q])
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        rel = float(((bat.double() - pla.double()).abs().amax(dim=1) /
                     pla.double().abs().amax(dim=1).clamp_min(1e-300)).max())
        out[name] = {"queries": nq, "batched_ms": round(tb * 1e3, 3), "first_call_ms": round(first * 1e3, 3),
                     "plain_ms": round(tp * 1e3, 3), "batched_qps": round(nq / tb, 1),
                     "plain_qps": round(nq / tp, 1), "speedup": round(tp / tb, 2), "max_rel_diff": rel}
    return out


def batched_cc(g, t_lo: int, t_hi: int, nq: int = 64) -> dict:
    """64 connected-components queries in one tga_connected_components_batch
    call against 64 plain calls, outputs in HBM, windows log-uniform in
    [2^10, 2^24] (the batching mix's CC share) and full windows; every row
    compared with its plain result."""
    import numpy as np
    import torch
    rng = np.random.default_rng(13)
    n = g.info["n"]
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), nq)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), nq)
    out = {}
    for name, wins in (("windowed", [(int(a), int(a + b)) for a, b in zip(starts, lens)]), ("full", [None] * nq)):
        bat = torch.empty((nq, n), dtype=torch.int32, device="cuda:0")
        pla = torch.empty((nq, n), dtype=torch.int32, device="cuda:0")
        g.connected_components_batch(wins, out=bat)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.connected_components_batch(wins, out=bat)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        g.connected_components(wins[0], out=pla[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q in range(nq):
            g.connected_components(wins[q], out=pla[q])
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        out[name] = {"queries": nq, "batched_ms": round(tb * 1e3, 3), "plain_ms": round(tp * 1e3, 3),
                     "batched_qps": round(nq / tb, 1), "plain_qps": round(nq / tp, 1),
                     "speedup": round(tp / tb, 2), "identical": bool(torch.equal(bat, pla))}
    return out


def batched_gc(g, t_lo: int, t_hi: int, nq: int = 64) -> dict:
    """64 graph-coloring queries (default priorities, 512 rounds) in one
    tga_graph_coloring_batch call against 64 plain calls, outputs in HBM,
    windowed and full windows; every row compared with its plain result."""
    import numpy as np
    import torch
    rng = np.random.default_rng(17)
    n = g.info["n"]
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), nq)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), nq)
    out = {}
    for name, wins in (("windowed", [(int(a), int(a + b)) for a, b in zip(starts, lens)]), ("full", [None] * nq)):
        bat = torch.empty((nq, n), dtype=torch.int32, device="cuda:0")
        pla = torch.empty((nq, n), dtype=torch.int32, device="cuda:0")
        g.graph_coloring_batch(wins, out=bat)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.graph_coloring_batch(wins, out=bat)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        g.graph_coloring(wins[0], out=pla[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q in range(nq):
            g.graph_coloring(wins[q], out=pla[q])
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        out[name] = {"queries": nq, "batched_ms": round(tb * 1e3, 3), "plain_ms": round(tp * 1e3, 3),
                     "batched_qps": round(nq / tb, 1), "plain_qps": round(nq / tp, 1),
                     "speedup": round(tp / tb, 2), "identical": bool(torch.equal(bat, pla))}
    return out


def batched_hits(g, t_lo: int, t_hi: int, nq: int = 64) -> dict:
    """64 HITS queries (20 iterations, FP32) in one tga_hits_batch call
    against 64 plain calls, outputs in HBM, windowed and full windows; rows
    compared with the plain results (max relative L-inf over hub and auth)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(21)
    n = g.info["n"]
    lens = np.exp(rng.uniform(np.log(2.0 ** 10), np.log(2.0 ** 24), nq)).astype(np.int64)
    starts = rng.integers(t_lo, max(t_hi, t_lo + 1), nq)
    out = {}
    for name, wins in (("windowed", [(int(a), int(a + b)) for a, b in zip(starts, lens)]), ("full", [None] * nq)):
        hb, ab, hp, ap = (torch.empty((nq, n), dtype=torch.float32, device="cuda:0") for _ in range(4))
        g.hits_batch(wins, iters=20, out_hub=hb, out_auth=ab)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.hits_batch(wins, iters=20, out_hub=hb, out_auth=ab)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        g.hits(wins[0], iters=20, out_hub=hp[0], out_auth=ap[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q in range(nq):
            g.hits(wins[q], iters=20, out_hub=hp[q], out_auth=ap[q])
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        rel = max(float(((x.double() - y.double()).abs().amax(dim=1) /
                         y.double().abs().amax(dim=1).clamp_min(1e-300)).max()) for x, y in ((hb, hp), (ab, ap)))
        out[name] = {"queries": nq, "batched_ms": round(tb * 1e3, 3), "plain_ms": round(tp * 1e3, 3),
                     "batched_qps": round(nq / tb, 1), "plain_qps": round(nq / tp, 1),
                     "speedup": round(tp / tb, 2), "max_rel_diff": rel}
    return out


def served_bfs(g, t_lo: int, t_hi: int) -> dict:
    """Serving: BFS window queries at Poisson rates of 1k and 10k per second
    (log-uniform window lengths in [2^10, 2^24], random sources), continuous
    batching against sequential plain BFS on the same arrivals; results stay
    in HBM.  The timed runs drop each result once answered; at 1k/s a further
    batched and plain run keep every result and compare them."""
    import torch
    from compresstemporalgraph_amd import serving
    out = {}
    for rate, count in ((1000, 500), (10000, 2000)):
        req = serving.poisson_requests(rate, count, g.info["n"], t_lo, t_hi, seed=rate)
        sb = serving.serve_batched(g, req, device_out=True)
        sp = serving.serve_plain(g, req, device_out=True)
        row = {"batched": sb, "plain": sp}
        if rate == 1000:
            kb, kp = {}, {}
            serving.serve_batched(g, req, keep=kb, device_out=True)
            serving.serve_plain(g, req, keep=kp, device_out=True)
            row["identical"] = bool(all(torch.equal(kb[i], kp[i]) for i in kb)) and len(kb) == count
            if not row["identical"]:
                raise SystemExit("served BFS results differ from the plain results")
        out[f"{rate}/s"] = row
    return out


MIX = {"DC": 0.1, "PR": 0.3, "CC": 0.2}  # the rest BFS: the batching mix of the at-scale queries


def served_mix(g, t_lo: int, t_hi: int, rates=((500, 400), (2000, 800))) -> dict:
    """Serving a mixed stream of window queries (40% BFS, 20% CC, 30% PR with
    20 iterations, 10% DC) at Poisson rates (default 500 and 2000 per second,
    `rates` holds (rate, queries) pairs) on one handle: BFS, CC and PR by
    continuous admission into their servers' lanes, DC in batched calls, with
    every kind's state charged against the 8 GB budget; against sequential
    plain execution.  Results in HBM.  The timed runs drop each result once
    answered (as a server handing results on would); at the first rate a
    further batched and plain run keep every result, and each is compared
    with the plain one (PageRank rows within 1e-6 relative L-inf, the rest
    exact)."""
    import torch
    from compresstemporalgraph_amd import serving
    # Untimed: the handle builds its caches (run structure, contact-time
    # cache) on a first small batched call, as a long-running server would
    # have done before its first request.
    wins = [(t_lo, t_hi), (t_lo, (t_lo + t_hi) // 2)]
    g.pagerank_batch(wins, iters=1)
    g.connected_components_batch(wins)
    g.degree_centrality_batch(wins)
    out = {}
    for rate, count in rates:
        req = serving.mixed_requests(rate, count, g.info["n"], t_lo, t_hi, seed=rate + 7, shares=MIX)
        sb = serving.serve_batched(g, req, device_out=True)
        sp = serving.serve_plain(g, req, device_out=True)
        row = {"batched": sb, "plain": sp}
        if rate == rates[0][0]:
            kb, kp = {}, {}
            serving.serve_batched(g, req, keep=kb, device_out=True)
            serving.serve_plain(g, req, keep=kp, device_out=True)
            def same(a, b):
                if isinstance(a, tuple):
                    return all(torch.equal(x, y) for x, y in zip(a, b))
                if a.dtype == torch.float32:
                    d = (a.double() - b.double()).abs().max() / b.double().abs().max().clamp_min(1e-300)
                    return bool(d <= 1e-6)
                return torch.equal(a, b)
            row["identical"] = bool(len(kb) == count and all(same(kb[i], kp[i]) for i in kb))
            if not row["identical"]:
                raise SystemExit("served mix results differ from the plain results")
        out[f"{rate}/s"] = row
    return out


def batched_mix(g, t_lo: int, t_hi: int, nq: int = 256) -> dict:
    """The batching configuration's query set: nq window queries (lengths
    log-uniform in [2^10, 2^24], uniform starts; 40% BFS from random sources,
    20% CC, 30% PR with 20 iterations, 10% DC with centrality), answered as
    batched calls of up to 64 queries of a kind (a remainder below
    serving.MIN_BATCH of its kind, 8 for BFS, as plain calls: the serving
    loop's rule, since a batched call costs about the same at any width),
    against nq plain calls.
    Results in HBM; every batched result is compared with its plain one
    (PageRank within 1e-6 relative L-inf, the rest exact).  One untimed
    batched pass and one plain call per kind come first (caches and list
    buffers)."""
    import torch
    from compresstemporalgraph_amd import serving
    n = g.info["n"]
    req = serving.mixed_requests(1.0, nq, n, t_lo, t_hi, seed=77, shares=MIX)
    calls = {"CC": serving._cc_batch, "PR": serving._pr_batch, "DC": serving._dc_batch}

    def run_batched():
        res = {}
        for kind in ("BFS", "CC", "PR", "DC"):
            idx = [i for i, r in enumerate(req) if r.kind == kind]
            for b in range(0, len(idx), 64):
                part = idx[b:b + 64]
                if len(part) < serving.MIN_BATCH.get(kind, 8):
                    res.update((i, serving._plain_one(g, req[i], True)) for i in part)
                    continue
                if kind == "BFS":
                    out = torch.empty((len(part), n), dtype=torch.int32, device="cuda:0")
                    g.bfs_batch([req[i].src for i in part], [req[i].window for i in part], out=out)
                    rows = [out[j] for j in range(len(part))]
                else:
                    rows = calls[kind](g, req, part, True)
                res.update(zip(part, rows))
        torch.cuda.synchronize()
        return res

    def run_plain():
        res = {i: serving._plain_one(g, r, True) for i, r in enumerate(req)}
        torch.cuda.synchronize()
        return res

    run_batched()
    for kind in ("BFS", "CC", "PR", "DC"):
        first = [r for r in req if r.kind == kind]
        if first:
            serving._plain_one(g, first[0], True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rb = run_batched()
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    rp = run_plain()
    tp = time.perf_counter() - t0

    def same(a, b):
        if isinstance(a, tuple):
            return all(torch.equal(x, y) for x, y in zip(a, b))
        if a.dtype == torch.float32:
            d = (a.double() - b.double()).abs().max() / b.double().abs().max().clamp_min(1e-300)
            return bool(d <= 1e-6)
        return torch.equal(a, b)

    identical = len(rb) == nq and all(same(rb[i], rp[i]) for i in range(nq))
    kinds = {k: sum(r.kind == k for r in req) for k in ("BFS", "CC", "PR", "DC")}
    return {"queries": nq, "kinds": kinds, "batched_ms": round(tb * 1e3, 3), "plain_ms": round(tp * 1e3, 3),
            "batched_qps": round(nq / tb, 1), "plain_qps": round(nq / tp, 1), "speedup": round(tp / tb, 2),
            "identical": bool(identical)}


def query_legs(args) -> int:
    """Batched and served BFS legs (child process of the default run)."""
    import torch
    torch.zeros(1, device="cuda:0")
    from compresstemporalgraph_amd import datasets
    from compresstemporalgraph_amd.engine import TemporalGraph
    root = args.data_root or os.path.join(tempfile.gettempdir(), "tga_bench")
    desc = datasets.ensure(root, args.workload)
    g = TemporalGraph(datasets.dataset_dir(root, args.workload), args.workload, device=0)
    g.autotune(passes=2)
    t_lo, t_hi = desc.get("t_min", 0), desc.get("t_max", 1 << 24)
    bmix = batched_mix(g, t_lo, t_hi)
    log(f"[bench] batched 256-query mix on {args.workload}: {bmix}")
    if not bmix["identical"]:
        raise SystemExit("batched mix results differ from the plain results")
    if args.mix_only:
        # The serving configuration on this graph: the same mix at Poisson
        # 100/s, 1k/s and 10k/s.
        smix = served_mix(g, t_lo, t_hi, rates=((100, 200), (1000, 1000), (10000, 4000)))
        log(f"[bench] served mix on {args.workload}: {smix}")
        bgc = batched_gc(g, t_lo, t_hi)
        log(f"[bench] batched GC on {args.workload}: {bgc}")
        if not all(v["identical"] for v in bgc.values()):
            raise SystemExit("batched GC results differ from the plain results")
        g.close()
        print(json.dumps({"batched_mix": bmix, "served_mix": smix, "batched_gc": bgc}), flush=True)
        return 0
    batched = batched_bfs(g, t_lo, t_hi)
    log(f"[bench] batched BFS: {batched}")
    if not batched["identical"]:
        raise SystemExit("batched BFS results differ from the plain results")
    bdc = batched_dc(g, t_lo, t_hi)
    log(f"[bench] batched DC: {bdc}")
    if not bdc["identical"]:
        raise SystemExit("batched DC results differ from the plain results")
    bcc = batched_cc(g, t_lo, t_hi)
    log(f"[bench] batched CC: {bcc}")
    if not all(v["identical"] for v in bcc.values()):
        raise SystemExit("batched CC results differ from the plain results")
    bpr = batched_pr(g, t_lo, t_hi)
    log(f"[bench] batched PR: {bpr}")
    if max(v["max_rel_diff"] for v in bpr.values()) > 1e-5:
        raise SystemExit("batched PageRank differs from the plain results beyond 1e-5")
    bhits = batched_hits(g, t_lo, t_hi)
    log(f"[bench] batched HITS: {bhits}")
    if max(v["max_rel_diff"] for v in bhits.values()) > 1e-5:
        raise SystemExit("batched HITS differs from the plain results beyond 1e-5")
    bgc = batched_gc(g, t_lo, t_hi)
    log(f"[bench] batched GC: {bgc}")
    if not all(v["identical"] for v in bgc.values()):
        raise SystemExit("batched GC results differ from the plain results")
    served = smix = None
    if not args.no_served:
        served = served_bfs(g, t_lo, t_hi)
        log(f"[bench] served BFS: {served}")
        smix = served_mix(g, t_lo, t_hi)
        log(f"[bench] served BFS/CC/PR/DC mix: {smix}")
    g.close()
    print(json.dumps({"batched_mix": bmix, "batched_bfs": batched, "batched_dc": bdc, "batched_cc": bcc,
                      "batched_pr": bpr, "batched_hits": bhits, "batched_gc": bgc, "served_bfs": served,
                      "served_mix": smix}),
          flush=True)
    return 0


def cpu_baseline(ds_dir: str, name: str, n: int, budget_s: float) -> dict:
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    orc = O.Oracle(ds_dir, name, n)
    # Bounded sample: consecutive vertex blocks spread over the id range,
    # grown until about budget_s of single-thread decode work.
    rng = np.random.default_rng(7)
    verts = rng.permutation(n).astype(np.int32)
    done, contacts, elapsed = 0, 0, 0.0
    step = 4096
    while elapsed < budget_s and done < n:
        chunk = verts[done:done + step]
        t0 = time.perf_counter()
        _, cnt = orc.neighbor_access(chunk)
        elapsed += time.perf_counter() - t0
        contacts += cnt
        done += len(chunk)
        step = min(step * 2, 1 << 18)
    # The same sample over all the host cores this job may use (16 per GPU).
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    _, cnt_mt = orc.neighbor_access_mt(verts[:done], threads)
    el_mt = time.perf_counter() - t0
    orc.close()
    return {"value": contacts / elapsed, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"{done} random vertices ({contacts} contacts) of the same graph, "
                      f"{elapsed:.1f} s single-thread oracle decode",
            "all_cores": {"value": cnt_mt / el_mt, "unit": "edges/s", "cores": threads, "kind": "port",
                          "sample": f"the same {done} vertices over {threads} threads, {el_mt:.2f} s"}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="netflow25")
    ap.add_argument("--no-side", action="store_true",
                    help="with the netflow25 headline: skip the chunglu21 pass, workload table and query legs")
    ap.add_argument("--data-root", default=None)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the seven-workload table and the batched / served query legs")
    ap.add_argument("--no-query-legs", action="store_true",
                    help="time the seven workloads but skip the batched / served BFS legs")
    ap.add_argument("--no-served", action="store_true",
                    help="with --query-legs: the batched legs only (no served streams)")
    ap.add_argument("--large-legs", action="store_true",
                    help="only time the large configuration's plain workloads and print their JSON (child process)")
    ap.add_argument("--mix-only", action="store_true",
                    help="with --query-legs: only the 256-query batched mix and the served mix at 100/s, 1k/s "
                         "and 10k/s")
    ap.add_argument("--encode-leg", action="store_true",
                    help="only time the host and the split (host graph + GPU streams) encoders on --workload, "
                         "check they write the same bytes, and print their JSON")
    ap.add_argument("--ts-variants", action="store_true",
                    help="only time the timestamp decode methods (one lane per record from the column copy, one "
                         "wave per record decoding speculatively) over --workload's flat records")
    ap.add_argument("--query-legs", action="store_true",
                    help="only run the batched and served BFS legs and print their JSON (run as a child process)")
    ap.add_argument("--traffic-bytes", type=float, default=None,
                    help="HBM bytes per pass from rocprof FETCH_SIZE + WRITE_SIZE (default: the committed "
                         "measurement in profiles/ when it matches the engine sources)")
    ap.add_argument("--calibrate-copy", type=int, default=None, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    if args.calibrate_copy is not None:
        import torch
        nbytes = args.calibrate_copy << 20
        a = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {nbytes} bytes three times")
        return 0

    if args.query_legs:
        return query_legs(args)
    if args.encode_leg:
        return encode_leg(args)
    if args.ts_variants:
        return ts_variants(args)
    if args.large_legs:
        return large_legs(args)

    from compresstemporalgraph_amd import datasets
    from compresstemporalgraph_amd.engine import TemporalGraph

    root = args.data_root or os.path.join(tempfile.gettempdir(), "tga_bench")
    os.environ.setdefault("TGAF_PROGRESS", "1")  # the large graph encodes for minutes
    # The at-scale plain graph (chunglu21) carries the seven-workload table
    # and the batched / served query legs; with the large graph as the
    # headline it is measured first, as extra keys of the same line.
    side = "chunglu21" if args.workload == "netflow25" and not args.no_side else None
    legs_on = args.workload if args.workload != "netflow25" else side
    batched = bdc = bcc = bpr = bhits = bgc = served = smix = bmix = smix_c1 = None
    if not args.no_workloads and not args.no_query_legs and legs_on:
        datasets.ensure(root, legs_on)
        # The query legs keep results in torch device tensors.  torch brings
        # its own HIP runtime (which must initialise before the engine's in a
        # process), so they run in a child process, before this process
        # touches the GPU.  The served mix runs on one handle (one stream),
        # so HIP's default of four hardware queues is enough.
        import subprocess
        cmd = [sys.executable, os.path.abspath(__file__), "--query-legs", "--workload", legs_on]
        if args.data_root:
            cmd += ["--data-root", args.data_root]
        env = dict(os.environ)
        child = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, env=env)
        if child.returncode != 0:
            raise SystemExit(f"query legs failed (exit {child.returncode})")
        legs = json.loads(child.stdout.strip().splitlines()[-1])
        bmix = {legs_on: legs["batched_mix"]}
        # The same 256-query mix on the small R-MAT graph (2^16 vertices,
        # 2^20 contacts), the graph the batching configuration names.
        cmd_c1 = [sys.executable, os.path.abspath(__file__), "--query-legs", "--mix-only", "--workload", "rmat16"]
        if args.data_root:
            cmd_c1 += ["--data-root", args.data_root]
        child = subprocess.run(cmd_c1, stdout=subprocess.PIPE, text=True, env=env)
        if child.returncode != 0:
            raise SystemExit(f"rmat16 mix leg failed (exit {child.returncode})")
        c1 = json.loads(child.stdout.strip().splitlines()[-1])
        bmix["rmat16"], smix_c1 = c1["batched_mix"], c1["served_mix"]
        batched, bdc, bcc, bpr = legs["batched_bfs"], legs["batched_dc"], legs["batched_cc"], legs["batched_pr"]
        bhits = legs["batched_hits"]
        bgc = {legs_on: legs["batched_gc"], "rmat16": c1["batched_gc"]}
        served, smix = legs["served_bfs"], legs["served_mix"]

    side_rec = None
    if side:
        sdesc = datasets.ensure(root, side)
        g = TemporalGraph(datasets.dataset_dir(root, side), side, device=0)
        rec = ra_pass(g, sdesc, args.steps, args.warmup, None)
        if not args.no_workloads:
            t0 = time.perf_counter()
            rec["workloads_ms"] = workload_times(g, sdesc.get("t_min", 0), sdesc.get("t_max", 1 << 24))
            log(f"[bench] {side} workloads timed in {time.perf_counter() - t0:.1f} s: {rec['workloads_ms']}")
        g.close()
        rec["workload"] = f"{side}: {WORKLOAD_DESC[side]}"
        side_rec = rec

    t0 = time.perf_counter()
    desc = datasets.ensure(root, args.workload, gpu=args.workload == "netflow25")  # GPU writers for the 1B graph
    log(f"[bench] dataset {args.workload}: n={desc['max_vertex']} contacts={desc['contacts']} "
        f"ready in {time.perf_counter() - t0:.1f} s")
    ds_dir = datasets.dataset_dir(root, args.workload)
    large = args.workload == "netflow25"

    t0 = time.perf_counter()
    g = TemporalGraph(ds_dir, args.workload, device=0)
    log(f"[bench] opened on GPU in {time.perf_counter() - t0:.2f} s: {g.info}")
    info = g.info
    n = info["n"]
    contacts = info["contacts"]
    rec = ra_pass(g, desc, args.steps, args.warmup, args.traffic_bytes)
    workloads = None
    if not args.no_workloads and not large:
        t0 = time.perf_counter()
        workloads = workload_times(g, desc.get("t_min", 0), desc.get("t_max", 1 << 24))
        log(f"[bench] workloads timed in {time.perf_counter() - t0:.1f} s: {workloads}")
    g.close()
    if not args.no_workloads and large:
        import subprocess
        cmd = [sys.executable, os.path.abspath(__file__), "--large-legs", "--workload", args.workload]
        if args.data_root:
            cmd += ["--data-root", args.data_root]
        child = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
        if child.returncode != 0:
            raise SystemExit(f"large workload legs failed (exit {child.returncode})")
        workloads = json.loads(child.stdout.strip().splitlines()[-1])
        log(f"[bench] workloads (results in HBM): {workloads}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(ds_dir, args.workload, n, args.cpu_budget)
        log(f"[bench] cpu baseline {cpu['value']:.3e} edges/s ({cpu['sample']})")

    line = {
        "metric": "decoded temporal edges/s",
        "value": rec["value"],
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: neighbor-access (RA) full decode pass; "
                               f"{WORKLOAD_DESC.get(args.workload, args.workload)}",
                   "vertices": n, "contacts": contacts,
                   "graph_bytes": info["graph_bytes"], "timestamps_bytes": info["timestamps_bytes"],
                   "parallelism": "single GPU"},
        "roofline": rec["roofline"],
        "autotune": rec["autotune"],
        "checksum": rec["checksum"],
        "cpu_baseline": cpu,
        "workloads_ms": workloads,
        "chunglu21": side_rec,
        "batched_mix": bmix,
        "batched_bfs": batched,
        "batched_dc": bdc,
        "batched_cc": bcc,
        "batched_pr": bpr,
        "batched_hits": bhits,
        "batched_gc": bgc,
        "served_mix": smix,
        "served_mix_rmat16": smix_c1,
        "served_bfs": served,
    }
    print(json.dumps(line), flush=True)
    return 0


def ra_pass(g, desc: dict, steps: int, warmup: int, traffic_bytes) -> dict:
    """Times `steps` full neighbor-access passes (after `warmup` untimed ones)
    with the graph resident in HBM; the pass's kernels are also timed with HIP
    events on the engine's stream.  The decoded checksum must equal the raw
    contact list's (size-independent parity at full size)."""
    info = g.info
    n, contacts, name = info["n"], info["contacts"], desc["name"]
    tune = g.autotune(passes=3)  # the walk's burst length, kept for the handle (outside the timed region)
    log(f"[bench] {name}: autotune {tune}")
    for _ in range(warmup):
        g.neighbor_access()
    g.profile(True)
    g.profile_read(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        cs, cnt = g.neighbor_access()
    elapsed = time.perf_counter() - t0
    kernel_ms, launches = g.profile_read(reset=True)
    g.profile(False)
    assert cnt == contacts
    log(f"[bench] {name}: {steps} steps in {elapsed:.3f} s; decode kernels {kernel_ms / steps:.3f} ms per pass "
        f"({launches} launches); checksum {cs:#018x}")
    raw_cs = desc.get("checksum")
    if raw_cs is not None and cs != raw_cs:
        raise SystemExit(f"{name}: decoded checksum {cs:#018x} differs from the raw list's {raw_cs:#018x}")
    # One pass = the concurrent launches of the item groups, timed as one
    # interval.
    avg_kernel_s = kernel_ms / steps / 1e3
    index_bits = (n + 1) * (info["l_graph"] + 2 + info["l_timestamps"] + 2)
    alg_bytes = info["graph_bytes"] + info["timestamps_bytes"] + index_bits / 8.0
    achieved = alg_bytes / avg_kernel_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic_bytes if traffic_bytes is not None else measured_traffic(name),
                "kernel": "k_ra (the concurrent group launches of one pass)",
                "kernel_ms": round(avg_kernel_s * 1e3, 4), "launches_per_pass": launches // steps,
                "algorithmic_bytes": int(alg_bytes)}
    return {"value": contacts * steps / elapsed, "ms_per_step": elapsed * 1e3 / steps, "roofline": roofline,
            "autotune": tune,
            "checksum": {"value": f"{cs:#018x}", "equals_raw_list": raw_cs is not None and cs == raw_cs}}


if __name__ == "__main__":
    sys.exit(main())
