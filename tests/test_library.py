"""The C-ABI libraries load and export every declared entry point; host-side
error paths return status codes instead of exiting (CPU only: nothing here
launches a kernel)."""
import ctypes as C
import os
import re
import shutil

import numpy as np
import pytest

from compresstemporalgraph_amd import build, datasets, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tgaf?_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("tga.h", build.ENGINE_LIB), ("tga_encode.h", build.ENGINE_LIB),
                                        ("tga_format.h", build.FORMAT_LIB)])
def test_exports(header, lib):
    assert os.path.exists(lib), f"{lib} not built"
    L = C.CDLL(lib)
    names = declared(header)
    assert len(names) >= 5
    for name in names:
        assert hasattr(L, name), f"{os.path.basename(lib)} does not export {name}"


def test_header_declares_seven_workloads():
    names = declared("tga.h")
    for w in ("tga_bfs", "tga_connected_components", "tga_degree_centrality", "tga_graph_coloring", "tga_hits",
              "tga_neighbor_access", "tga_pagerank"):
        assert w in names


def _open_rc(directory, name, max_label=0):
    L = engine.lib()
    h = C.c_void_p()
    desc = engine.DatasetDesc(max_label, 0, 0, engine.TGA_FROM_FILE)
    rc = L.tga_open_dir(directory.encode(), name.encode(), C.byref(desc), 0, C.byref(h))
    return rc, L.tga_last_error().decode()


def test_open_missing_file_is_an_error(tmp_path):
    rc, msg = _open_rc(str(tmp_path), "nothing")
    assert rc == -2 and "nothing" in msg  # TGA_EIO


def test_open_truncated_and_log2_zero(tmp_path, ds_sample):
    d = str(tmp_path / "t")
    shutil.copytree(ds_sample.dir, d)
    # header-only timestamps: fewer stream bits than MAX_LABEL -> log2(0)
    with open(os.path.join(d, "sample.timestamps"), "r+b") as f:
        f.truncate(8)
    rc, msg = _open_rc(d, "sample")
    assert rc == -4 and "log2" in msg  # TGA_ERANGE
    with open(os.path.join(d, "sample.timestamps"), "r+b") as f:
        f.truncate(4)
    rc, msg = _open_rc(d, "sample")
    assert rc == -3  # TGA_EFORMAT: no 64-bit header


def test_open_short_index_is_an_error(tmp_path, ds_sample):
    d = str(tmp_path / "i")
    shutil.copytree(ds_sample.dir, d)
    with open(os.path.join(d, "graph_upper_bits.bin"), "r+b") as f:
        f.truncate(16)
    rc, msg = _open_rc(d, "sample")
    assert rc == -3 and "index" in msg


def test_parse_workload_names():
    assert engine.parse_workload("pagerank") == "PR"
    assert engine.parse_workload("Breadth-First-Search") == "BFS"
    assert engine.parse_workload("graph_coloring") == "GC"
    with pytest.raises(ValueError):
        engine.parse_workload("sssp")


def test_workload_generators_are_deterministic():
    a = datasets.generate(datasets.Workload("x", "rmat", 1 << 10, 5000, 3, 1 << 10, False))
    b = datasets.generate(datasets.Workload("x", "rmat", 1 << 10, 5000, 3, 1 << 10, False))
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    c = datasets.generate(datasets.Workload("y", "chung_lu", 1 << 12, 20000, 4, 1 << 16, False))
    e = datasets.generate(datasets.Workload("y", "chung_lu", 1 << 12, 20000, 4, 1 << 16, False))
    for u, v in zip(c, e):
        np.testing.assert_array_equal(u, v)
    s, d, t = c
    assert len(s) == 20000 and (np.diff(s) >= 0).all()


def test_netflow_generator(tmp_path):
    """The netflow-like generator: exactly m contacts sorted by (src, dst, t),
    deterministic, contacts per pair with mean near 20, times inside the span;
    the raw checksum stored with the dataset equals the oracle's decoded
    neighbor-access checksum of its encoding."""
    import sys
    w = datasets.Workload("nf", "netflow", 1 << 12, 200000, 4, 1 << 20, False)
    s, d, t = datasets.generate(w)
    s2, d2, t2 = datasets.generate(w)
    for u, v in ((s, s2), (d, d2), (t, t2)):
        np.testing.assert_array_equal(u, v)
    assert len(s) == 200000 and t.min() >= 0 and t.max() < 1 << 20
    key = s.astype(np.int64) << 32 | d
    assert (np.diff(key) >= 0).all()
    same = np.diff(key) == 0
    assert (np.diff(t)[same] >= 0).all()
    per_pair = len(s) / (1 + (~same).sum())
    assert 15 < per_pair < 25, per_pair
    desc = datasets.write_dataset(str(tmp_path), "nf", s, d, t, n_min=w.n)
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
    import oracle as O
    o = O.Oracle(datasets.dataset_dir(str(tmp_path), "nf"), "nf", w.n)
    cs, cnt = o.neighbor_access()
    o.close()
    assert cnt == len(s) and cs == desc["checksum"]


def test_entry_points_reject_null_handles():
    """Every compute entry point returns TGA_EINVAL for a NULL handle (or
    server) and names itself in tga_last_error(); none exits or touches a
    device."""
    L = engine.lib()
    W = engine.Window
    w = W(0, 10)
    buf = (C.c_int32 * 4)()
    fbuf = (C.c_float * 4)()
    wins = (W * 1)(w)
    src = (C.c_int32 * 1)(0)
    u64 = C.c_uint64()
    i64 = C.c_int64()
    i32 = C.c_int32()
    calls = {
        "tga_neighbor_access": lambda: L.tga_neighbor_access(None, None, 0, C.byref(u64), C.byref(i64)),
        "tga_degree_centrality": lambda: L.tga_degree_centrality(None, w, buf, None, 0),
        "tga_degree_centrality_batch": lambda: L.tga_degree_centrality_batch(None, wins, 1, buf, None, 0),
        "tga_bfs": lambda: L.tga_bfs(None, w, 0, buf, 0),
        "tga_bfs_batch": lambda: L.tga_bfs_batch(None, wins, src, 1, buf, 0),
        "tga_bfs_server_open": lambda: L.tga_bfs_server_open(None, C.byref(C.c_void_p())),
        "tga_bfs_server_admit": lambda: L.tga_bfs_server_admit(None, 0, w, C.byref(i32)),
        "tga_bfs_server_step": lambda: L.tga_bfs_server_step(None, C.byref(u64)),
        "tga_bfs_server_result": lambda: L.tga_bfs_server_result(None, 0, buf, 0),
        "tga_cc_server_open": lambda: L.tga_cc_server_open(None, C.byref(C.c_void_p())),
        "tga_cc_server_admit": lambda: L.tga_cc_server_admit(None, w, C.byref(i32)),
        "tga_cc_server_step": lambda: L.tga_cc_server_step(None, C.byref(u64)),
        "tga_cc_server_result": lambda: L.tga_cc_server_result(None, 0, buf, 0),
        "tga_pr_server_open": lambda: L.tga_pr_server_open(None, 0.85, C.byref(C.c_void_p())),
        "tga_pr_server_admit": lambda: L.tga_pr_server_admit(None, w, 1, C.byref(i32)),
        "tga_pr_server_step": lambda: L.tga_pr_server_step(None, C.byref(u64)),
        "tga_pr_server_result": lambda: L.tga_pr_server_result(None, 0, fbuf, 0),
        "tga_earliest_arrival": lambda: L.tga_earliest_arrival(None, w, 0, buf, 0),
        "tga_latest_departure": lambda: L.tga_latest_departure(None, w, 0, buf, None, 0),
        "tga_connected_components": lambda: L.tga_connected_components(None, w, buf, 0),
        "tga_connected_components_batch": lambda: L.tga_connected_components_batch(None, wins, 1, buf, 0),
        "tga_graph_coloring": lambda: L.tga_graph_coloring(None, w, 4, None, buf, 0),
        "tga_pagerank": lambda: L.tga_pagerank(None, w, 1, 0.85, 0.0, fbuf, None, 0),
        "tga_pagerank_batch": lambda: L.tga_pagerank_batch(None, wins, 1, 1, 0.85, fbuf, 0),
        "tga_hits": lambda: L.tga_hits(None, w, 1, fbuf, fbuf, 0),
        "tga_hits_batch": lambda: L.tga_hits_batch(None, wins, 1, 1, fbuf, fbuf, 0),
        "tga_profile": lambda: L.tga_profile(None, 1),
        "tga_profile_read": lambda: L.tga_profile_read(None, None, None, 0),
        "tga_info": lambda: L.tga_info(None, None),
    }
    for name, call in calls.items():
        assert call() == -1, name  # TGA_EINVAL
        assert name in L.tga_last_error().decode(), name
