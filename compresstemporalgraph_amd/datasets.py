"""Synthetic workload datasets in the reference's directory layout.

A dataset named ``D`` under ``root`` is stored exactly where the reference
looks for it (gpu_cuda/shared/dataset_config_base.cuh:44-68):

    root/data/compressed/D/D.graph, D.offsets, D.properties, D.timestamps
    root/data/compressed/D/graph_{lower,upper}_bits.bin
    root/data/compressed/D/timestamp_{lower,upper}_bits.bin
    root/data/raw/D.raw.txt.gz          (optional 4-column contact list)

Workloads (seeded, deterministic):

* ``rmat16`` — R-MAT (0.57, 0.19, 0.19, 0.05) over 2^16 vertices, 2^20
  contacts, Geometric(1/3) contacts per sampled pair, t uniform in [0, 2^20),
  seed 1.  The reference's CPU-runnable plain case.
* ``chunglu21`` — Chung-Lu, exponent 2.1, 2^21 vertices, 2^25 contacts,
  timestamps i.i.d. from the diurnal intensity 1 + 0.5 sin(2 pi t / 2^18) over
  2^24 ticks, seed 2.  The plain at-scale case.
* ``netflow25`` — netflow-like, 2^25 vertices, 2^30 contacts: pairs from
  power-law weights (exponent 2.5 on both ends, so the largest out-degree is
  about 10^6 contacts), a Pareto (alpha 2) number of contacts per pair with
  mean about 20, bursty Pareto (alpha 1.5) inter-arrival gaps of at least 256
  ticks over 2^30 ticks, seed 4.  The large single-GPU case: its streams
  (about 2.5 GB) exceed the 256 MB Infinity Cache, so a pass reads HBM.
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass

import numpy as np

from . import fmt


@dataclass
class Workload:
    name: str
    kind: str
    n: int
    m: int
    seed: int
    t_span: int
    write_raw: bool


WORKLOADS = {
    "rmat16": Workload("rmat16", "rmat", 1 << 16, 1 << 20, 1, 1 << 20, True),
    "chunglu21": Workload("chunglu21", "chung_lu", 1 << 21, 1 << 25, 2, 1 << 24, False),
    "netflow25": Workload("netflow25", "netflow", 1 << 25, 1 << 30, 4, 1 << 30, False),
}


def dataset_dir(root: str, name: str) -> str:
    return os.path.join(root, "data", "compressed", name)


def raw_path(root: str, name: str) -> str:
    return os.path.join(root, "data", "raw", name + ".raw.txt.gz")


def generate(w: Workload):
    if w.kind == "rmat":
        scale = int(w.n).bit_length() - 1
        return fmt.gen_rmat(scale, w.m, t_span=w.t_span, seed=w.seed)
    if w.kind == "chung_lu":
        return fmt.gen_chung_lu(w.n, w.m, exponent=2.1, t_span=w.t_span, period=w.t_span // 64, seed=w.seed)
    if w.kind == "netflow":
        return fmt.gen_netflow(w.n, w.m, exponent=2.5, count_min=10.0, gap_min=256.0, t_span=w.t_span, seed=w.seed)
    raise ValueError(w.kind)


def write_dataset(root: str, name: str, src, dst, t, n_min: int = 0, raw: bool = False, gpu: bool = False) -> dict:
    """Encodes a contact list and writes the dataset plus a descriptor JSON;
    with gpu, the per-contact streams and the index files are written on the
    GPU (engine.encode_gpu, the same bytes)."""
    d = dataset_dir(root, name)
    os.makedirs(d, exist_ok=True)
    if gpu:
        from .engine import encode_gpu
        st = encode_gpu(src, dst, t, d, name, n_min=n_min)
    else:
        st = fmt.encode(src, dst, t, d, name, n_min=n_min)
    if raw:
        os.makedirs(os.path.dirname(raw_path(root, name)), exist_ok=True)
        fmt.write_raw_gz(raw_path(root, name), src, dst, t)
    desc = {
        "name": name,
        "max_label": st.nodes - 1,
        "max_vertex": st.nodes,
        "active_begin": 0,
        "active_end": st.nodes,
        "min_timestamp": st.min_timestamp,
        "contacts": st.arcs,
        "t_min": int(np.min(t)) if len(t) else 0,
        "t_max": int(np.max(t)) if len(t) else 0,
        "checksum": fmt.contact_checksum(src, dst, t),
        "stats": asdict(st),
    }
    with open(os.path.join(d, name + ".dataset.json"), "w") as f:
        json.dump(desc, f, indent=1)
    return desc


def ensure(root: str, name: str, gpu: bool = False) -> dict:
    """Generates workload ``name`` under ``root`` unless it is already there
    (gpu: encode with the GPU writers, as write_dataset)."""
    d = dataset_dir(root, name)
    meta = os.path.join(d, name + ".dataset.json")
    if os.path.exists(meta):
        with open(meta) as f:
            return json.load(f)
    w = WORKLOADS[name]
    src, dst, t = generate(w)
    return write_dataset(root, name, src, dst, t, n_min=w.n, raw=w.write_raw, gpu=gpu)


def load_descriptor(root: str, name: str) -> dict:
    with open(os.path.join(dataset_dir(root, name), name + ".dataset.json")) as f:
        return json.load(f)


def contacts_of(root: str, name: str):
    """Regenerates the contact list of a named workload (for brute-force checks)."""
    src, dst, t = generate(WORKLOADS[name])
    return np.asarray(src), np.asarray(dst), np.asarray(t)
