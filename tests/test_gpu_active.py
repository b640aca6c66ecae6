"""GPU parity with an active vertex sub-range (the reference's
ACTIVE_VERTEX_BEGIN / isTraversableNode): every entry point, the batched
forms and both the list and decode paths, on a handle opened with
active = [n/10, n - n/7), against the oracle with the same range.  Inactive
vertices are neither traversed nor followed, so this exercises the activity
filters of every kernel, the contact-time cache's inactive-neighbor times and
the run structure's dropped runs."""
import numpy as np
import pytest

from compresstemporalgraph_amd import fmt
from compresstemporalgraph_amd.engine import TemporalGraph

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def rel_linf(a, b):
    b = np.asarray(b, np.float64)
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module", params=["adv", "cl14", "nf14"])
def act(request):
    ds = request.getfixturevalue({"adv": "ds_adv", "cl14": "ds_cl", "nf14": "ds_nf"}[request.param])
    rng = (ds.n // 10, ds.n - ds.n // 7)
    print(f"\n[gpu] dataset {ds.name} n={ds.n} active={rng}", flush=True)
    g = TemporalGraph(ds.dir, ds.name, active=rng)
    yield ds, g, rng
    g.close()


def windows(ds):
    lo, hi = int(ds.t.min()), int(ds.t.max())
    span = hi - lo
    return [None, (lo + span // 4, lo + span // 2), (lo + span // 3, lo + span // 3)]


def sources(ds, rng):
    inside = ds.src[(ds.src >= rng[0]) & (ds.src < rng[1])]
    return [rng[0], int(inside[len(inside) // 2])]


def test_active_neighbor_access(act):
    ds, g, rng = act
    assert g.neighbor_access() == ds.oracle().neighbor_access(active=rng)
    verts = np.random.default_rng(3).integers(0, ds.n, 300).astype(np.int32)
    assert g.neighbor_access(verts) == ds.oracle().neighbor_access(verts, active=rng)


@pytest.mark.parametrize("path", ["lists", "decode"])
def test_active_dc_gc_pr_hits(act, path, monkeypatch):
    ds, g, rng = act
    if path == "decode":
        monkeypatch.setenv("TGA_LIST_MAX_BYTES", "0")
    o = ds.oracle()
    prio = fmt.java_random_ints(123, 256, ds.n)
    for w in windows(ds):
        cnt, cen = g.degree_centrality(w, centrality=True)
        ocnt, ocen = o.degree(window=w, active=rng, centrality=True)
        np.testing.assert_array_equal(cnt, ocnt)
        np.testing.assert_array_equal(cen, ocen)
        np.testing.assert_array_equal(g.graph_coloring(w), o.gc(prio, 512, window=w, active=rng))
        ref, _ = o.pagerank(20, window=w, active=rng, dtype=np.float32)
        assert rel_linf(g.pagerank(w, iters=20), ref) <= REL_TOL
        hub, auth = g.hits(w, iters=10)
        h32, a32 = o.hits(10, window=w, active=rng, dtype=np.float32)
        assert rel_linf(hub, h32) <= REL_TOL and rel_linf(auth, a32) <= REL_TOL


def test_active_frontier_algorithms(act):
    ds, g, rng = act
    o = ds.oracle()
    for w in windows(ds):
        np.testing.assert_array_equal(g.connected_components(w), o.cc(window=w, active=rng))
        for s in sources(ds, rng):
            np.testing.assert_array_equal(g.bfs(s, w), o.bfs(s, window=w, active=rng))
            np.testing.assert_array_equal(g.earliest_arrival(s, w), o.earliest_arrival(s, window=w, active=rng))
            np.testing.assert_array_equal(g.latest_departure(s, w), o.latest_departure(s, window=w, active=rng))


def test_active_batched(act):
    ds, g, rng = act
    o = ds.oracle()
    wins = windows(ds) * 3
    srcs = sources(ds, rng) * 5
    srcs = srcs[:len(wins)]
    lv = g.bfs_batch(srcs, wins)
    cnt, cen = g.degree_centrality_batch(wins)
    lab = g.connected_components_batch(wins)
    pr = g.pagerank_batch(wins, iters=20)
    hub, auth = g.hits_batch(wins, iters=10)
    for q, w in enumerate(wins):
        h32, a32 = o.hits(10, window=w, active=rng, dtype=np.float32)
        assert rel_linf(hub[q], h32) <= REL_TOL and rel_linf(auth[q], a32) <= REL_TOL, f"hits {q}"
        np.testing.assert_array_equal(lv[q], o.bfs(srcs[q], window=w, active=rng), err_msg=f"bfs {q}")
        ocnt, ocen = o.degree(window=w, active=rng, centrality=True)
        np.testing.assert_array_equal(cnt[q], ocnt, err_msg=f"dc {q}")
        np.testing.assert_array_equal(cen[q], ocen, err_msg=f"dc {q}")
        np.testing.assert_array_equal(lab[q], o.cc(window=w, active=rng), err_msg=f"cc {q}")
        ref, _ = o.pagerank(20, window=w, active=rng, dtype=np.float32)
        assert rel_linf(pr[q], ref) <= REL_TOL, f"pr {q}"
