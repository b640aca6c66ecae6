"""Captured iteration loops (GPU): the batched PageRank, HITS and connected
components calls replay their iteration / round loops from hipGraphs cached
per batch shape.  Every replay -- several in a row, with other work and
synchronisations between them, and after the shape changed and back --
must equal the same call launched one kernel at a time (a handle opened with
TGA_GRAPHS=0), bit for bit, and the oracle.
"""
import numpy as np
import pytest

from compresstemporalgraph_amd.engine import TemporalGraph

pytestmark = pytest.mark.gpu


def _wins(ds, k, seed):
    rng = np.random.default_rng(seed)
    lo, hi = int(ds.t.min()), int(ds.t.max())
    out = []
    for q in range(k):
        if q % 5 == 0:
            out.append(None)
        else:
            a, b = sorted(int(v) for v in rng.integers(lo, hi + 1, 2))
            out.append((a, b))
    return out


@pytest.mark.parametrize("which", ["ds_rmat", "ds_nf"])
def test_graph_replays_equal_eager(which, request, monkeypatch):
    ds = request.getfixturevalue(which)
    print(f"\n[gpu] graph replays on {ds.name}", flush=True)
    g = TemporalGraph(ds.dir, ds.name)
    monkeypatch.setenv("TGA_GRAPHS", "0")
    e = TemporalGraph(ds.dir, ds.name)
    monkeypatch.delenv("TGA_GRAPHS")
    o = ds.oracle()
    w1, w2 = _wins(ds, 64, 1), _wins(ds, 20, 2)
    try:
        want = {}
        for tag, w in (("w1", w1), ("w2", w2)):
            want[tag] = (e.pagerank_batch(w, iters=20), e.hits_batch(w, iters=10), e.connected_components_batch(w))
        src = int(ds.src[0])
        for rep, tag in enumerate(["w1", "w1", "w1", "w2", "w1"]):
            w = w1 if tag == "w1" else w2
            pr = g.pagerank_batch(w, iters=20)
            g.neighbor_access()  # other work between replays
            hub, auth = g.hits_batch(w, iters=10)
            lv = g.bfs(src)
            cc = g.connected_components_batch(w)
            np.testing.assert_array_equal(pr, want[tag][0], err_msg=f"pagerank replay {rep}")
            np.testing.assert_array_equal(hub, want[tag][1][0], err_msg=f"hits replay {rep}")
            np.testing.assert_array_equal(auth, want[tag][1][1], err_msg=f"hits replay {rep}")
            np.testing.assert_array_equal(cc, want[tag][2], err_msg=f"cc replay {rep}")
            np.testing.assert_array_equal(lv, o.bfs(src))
        for q in (0, 1, 7):
            np.testing.assert_array_equal(want["w1"][2][q], o.cc(window=w1[q]))
            ref, _ = o.pagerank(20, window=w1[q], dtype=np.float32)
            err = np.abs(want["w1"][0][q] - ref).max() / np.abs(ref).max()
            assert err <= 1e-5
    finally:
        g.close()
        e.close()
