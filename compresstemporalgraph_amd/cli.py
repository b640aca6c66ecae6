"""Command line mirroring the reference's CPU driver.

    python -m compresstemporalgraph_amd.cli <dataset> <workload> [--root R] [--window S E] ...

Usage, workload names and the RESULT line follow
cpu_java/.../BaselineRunner.java:639-681 (``BaselineRunner <dataset>
<workload>``); the dataset is read from ``<root>/data/compressed/<dataset>/``
(root from --root, else TGA_COMPRESS_ROOT, else the current directory, as
gpu_cuda/shared/path_config.cuh:11-20 resolves it).  ``--generate`` first
writes one of the named synthetic workloads (compresstemporalgraph_amd.datasets)
under the root.  The timed region is the GPU call only; the file load and the
decode-schedule build are outside it, as in the reference.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import datasets
from .engine import FULL_WINDOW, TemporalGraph, parse_workload


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="tga_run", description=__doc__.splitlines()[0])
    ap.add_argument("dataset")
    ap.add_argument("workload")
    ap.add_argument("--root", default=os.environ.get("TGA_COMPRESS_ROOT", "."))
    ap.add_argument("--generate", action="store_true", help="write the named synthetic workload first")
    ap.add_argument("--window", type=int, nargs=2, metavar=("START", "END"), default=None)
    ap.add_argument("--source", type=int, default=0, help="BFS source vertex")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--rounds", type=int, default=512)
    ap.add_argument("--active", type=int, nargs=2, metavar=("BEGIN", "END"), default=None)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--out", default=None, help="save the per-vertex result as .npy")
    a = ap.parse_args(argv)

    workload = parse_workload(a.workload)
    if a.generate:
        datasets.ensure(a.root, a.dataset)
    g = TemporalGraph.open(a.root, a.dataset, device=a.device, active=a.active)
    window = tuple(a.window) if a.window else FULL_WINDOW
    t0 = time.perf_counter_ns()
    if workload == "RA":
        res = g.neighbor_access()
    elif workload == "BFS":
        res = g.bfs(a.source, window)
    elif workload == "CC":
        res = g.connected_components(window)
    elif workload == "PR":
        res = g.pagerank(window, iters=a.iters)
    elif workload == "HITS":
        res = g.hits(window, iters=a.iters)
    elif workload == "DC":
        res = g.degree_centrality(window)
    else:
        res = g.graph_coloring(window, rounds=a.rounds)
    elapsed = time.perf_counter_ns() - t0
    summary = ""
    if workload == "RA":
        summary = f" checksum={res[0]:#018x} contacts={res[1]}"
    elif workload in ("BFS", "CC", "DC", "GC"):
        arr = np.asarray(res)
        summary = f" reached={int((arr != 0x7FFFFFFF).sum())}" if workload in ("BFS", "GC") else ""
    if a.out is not None:
        np.save(a.out, np.asarray(res[0] if workload == "HITS" else res))
    print(f"RESULT dataset={a.dataset} workload={workload} elapsed_ns={elapsed} elapsed_ms={elapsed / 1e6:.3f} "
          f"basename={os.path.join(a.root, 'data', 'compressed', a.dataset, a.dataset)} "
          f"active_from={g.info['active_begin']} active_to={g.info['active_end']}{summary}")
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
