"""Shared fixtures: small synthetic datasets written by the format kit.

GPU tests are marked ``gpu``; everything else runs on the CPU.  Datasets are
generated into a session temp directory with the reference's layout
(root/data/compressed/<name>/...), so nothing here reads the reference tree.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from compresstemporalgraph_amd import datasets, fmt  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def sort_contacts(src, dst, t):
    order = np.lexsort((t, dst, src))
    return np.asarray(src)[order].astype(np.int32), np.asarray(dst)[order].astype(np.int32), \
        np.asarray(t)[order].astype(np.int64)


def adversarial_graph(seed=7, n=3000):
    """Lists built to exercise every record feature: references down to depth 3
    (consecutive vertices share most successors), copy blocks with skips,
    intervals, multiples, self loops, vertex 0 as first residual, neighbor
    MAX_LABEL, empty vertices and a hub."""
    rng = np.random.default_rng(seed)
    S, D, T = [], [], []
    base = np.sort(rng.choice(n, 60, replace=False))
    for x in range(n - 5):
        kind = x % 11
        if kind == 10:
            continue  # empty vertex
        if kind < 6:  # near-copies of a shared base: references and blocks
            keep = base[rng.random(len(base)) < 0.85]
            extra = rng.choice(n, rng.integers(0, 6), replace=True)
            lst = np.concatenate([keep, extra])
            if rng.random() < 0.3:
                base = np.sort(np.unique(np.concatenate([keep, rng.choice(n, 8)])))
        elif kind == 6:  # intervals
            a = int(rng.integers(0, n - 40))
            lst = np.concatenate([np.arange(a, a + int(rng.integers(4, 30))), rng.choice(n, 3)])
        elif kind == 7:  # vertex 0 as the first residual, and the largest id
            lst = np.array([0, 5, 9, n - 1, n - 1])
        elif kind == 8:  # self loop and multiples
            lst = np.array([x] * 3 + list(rng.choice(n, 4)))
        else:
            lst = rng.choice(n, rng.integers(1, 12))
        reps = rng.geometric(0.5, len(lst))
        for v, r in zip(lst, reps):
            for _ in range(int(r)):
                S.append(x)
                D.append(int(v))
                T.append(int(rng.integers(0, 5000)))
    # one hub
    hub = n - 5
    for v in rng.choice(n, 4000):
        S.append(hub)
        D.append(int(v))
        T.append(int(rng.integers(0, 5000)))
    return sort_contacts(S, D, T), n


def sample_contacts(dense=True):
    """The 1000-contact ChronoGraph fixture (4 columns: src dst 1 t).

    With ids up to 9992 and mostly zero timestamps, the raw fixture's
    .timestamps stream has fewer bits than MAX_LABEL, so the reference's
    index width log2(bytes * 8 / MAX_LABEL) would be log2(0) (bfs.cu:1786);
    ``dense=True`` relabels the ids to 0..k-1 in increasing order, which keeps
    every list's order and makes the dataset loadable."""
    a = np.loadtxt(os.path.join(GOLDEN, "cbtComm-sample.txt"), dtype=np.int64)
    s, d, t = a[:, 0], a[:, 1], a[:, 3]
    if dense:
        ids, inv = np.unique(np.concatenate([s, d]), return_inverse=True)
        s, d = inv[:len(s)], inv[len(s):]
    return sort_contacts(s, d, t)


class DS:
    """One dataset on disk plus its raw contacts."""

    def __init__(self, root, name, src, dst, t, n_min=0):
        self.root, self.name = root, name
        self.src, self.dst, self.t = src, dst, t
        self.desc = datasets.write_dataset(root, name, src, dst, t, n_min=n_min)
        self.n = self.desc["max_vertex"]
        self.dir = datasets.dataset_dir(root, name)

    def oracle(self):
        import oracle as O
        if not hasattr(self, "_oracle"):
            self._oracle = O.Oracle(self.dir, self.name, self.n)
        return self._oracle


@pytest.fixture(scope="session")
def data_root(tmp_path_factory):
    return str(tmp_path_factory.mktemp("tga_data"))


@pytest.fixture(scope="session")
def ds_sample(data_root):
    s, d, t = sample_contacts()
    return DS(data_root, "sample", s, d, t)


@pytest.fixture(scope="session")
def ds_adv(data_root):
    (s, d, t), n = adversarial_graph()
    return DS(data_root, "adv", s, d, t, n_min=n)


@pytest.fixture(scope="session")
def ds_rmat(data_root):
    s, d, t = fmt.gen_rmat(16, 1 << 20, seed=1)
    return DS(data_root, "rmat16", s, d, t, n_min=1 << 16)


@pytest.fixture(scope="session")
def ds_nf(data_root):
    """Netflow-like: every contact in a run of one repeated neighbor, bursty
    times whose deltas turn negative between pairs, hub records of ~10^4
    contacts (chunked)."""
    s, d, t = fmt.gen_netflow(1 << 14, 1 << 19, t_span=1 << 20, gap_min=16.0, seed=6)
    return DS(data_root, "nf14", s, d, t, n_min=1 << 14)


@pytest.fixture(scope="session")
def ds_cl(data_root):
    s, d, t = fmt.gen_chung_lu(1 << 14, 1 << 18, exponent=2.1, t_span=1 << 20, period=(1 << 20) // 64, seed=5)
    return DS(data_root, "cl14", s, d, t, n_min=1 << 14)


@pytest.fixture(scope="session")
def ds_ext(data_root):
    """Timestamps across the whole int32 range: consecutive contacts of a
    record jump between INT32_MIN, INT32_MAX and values near 0, so deltas
    reach +-(2^32 - 1) (33-bit int2nat values, the longest zeta_2 codes);
    one record long enough to be chunked."""
    rng = np.random.default_rng(11)
    n = 3000
    ext = np.array([-2**31, 2**31 - 1, -1, 0, 1, -2**31 + 1, 2**31 - 2], np.int64)
    src = np.concatenate([rng.integers(0, n, 40000), np.full(3000, 7)])
    dst = rng.integers(0, n, len(src))
    t = np.where(rng.random(len(src)) < 0.7, ext[rng.integers(0, len(ext), len(src))],
                 rng.integers(-2**31, 2**31, len(src)))
    s, d, tt = sort_contacts(src, dst, t)
    return DS(data_root, "ext32", s, d, tt, n_min=n)
