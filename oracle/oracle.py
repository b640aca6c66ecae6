"""CPU oracle — TEST INFRASTRUCTURE ONLY.

Python binding of ``oracle/liboracle.so`` (the C restatement in ``oracle.c``)
plus the raw-contact brute force of ``oracle/bruteforce.py``.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this module, and only as the checker or the timed CPU baseline;
the product library never calls it.

How each oracle function is pinned (checked by something other than itself):

* decode tables — regenerated from the closed-form gamma / zeta_3 definitions
  and compared entry by entry (and by SHA-256 of the text) with
  gpu_cuda/decode_tables/{gammaInA,zetaInA,gamma8192}.txt
  (tests/test_format_codes.py, fixture tests/golden/decode_tables.json).
* gamma / zeta_k / unary / int2nat — known-answer vectors (WriteLongZetaTest
  values 1, 2, 1, 198765 after a 64-bit header 124; hand-derived codewords)
  and encode->decode round trips over random values.
* record grammar (``decode_list``) — encode->decode round trips of random and
  adversarial multigraphs against the raw contact list (every successor, every
  timestamp, every vertex), the intervalizeMultiples KAT of
  BVMultiGraphTest.java:98-122, and the IntMultiplesSequenceIterator KAT of
  IntMultiplesSequenceIteratorTest.java:9-31 (via a record built from it).
* Elias-Fano getOffset — compared with the offsets the encoder wrote, for every
  vertex (brute-force select over the whole upper-bit file).
* algorithms — compared with ``bruteforce.py``, which works on the raw
  contact list (never on the decoder), and which itself is checked against
  closed forms: BFS levels against scipy's unweighted shortest paths, CC
  labels against a transitive closure on small graphs, PageRank against the
  stationary vector of the explicit Google matrix, HITS against the principal
  eigenvectors of A^T A and A A^T, GC against the greedy-rounds invariant
  (no colored vertex is blocked by an uncolored higher-priority neighbor).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
INF32 = 0x7FFFFFFF

_lib = None


class _Q(C.Structure):
    _fields_ = [("ws", C.c_int64), ("we", C.c_int64), ("active_begin", C.c_int32), ("active_end", C.c_int32)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        p32, p64, pu64 = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)
        L.orc_open.restype = C.c_void_p
        L.orc_open.argtypes = [C.c_char_p, C.c_char_p, C.c_int64]
        L.orc_close.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_offsets.argtypes = [C.c_void_p, C.c_int64, pu64, pu64]
        L.orc_widths.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), p64]
        L.orc_decode_all.argtypes = [C.c_void_p]
        L.orc_num_contacts.argtypes = [C.c_void_p]
        L.orc_num_contacts.restype = C.c_int64
        L.orc_q1_records.argtypes = [C.c_void_p]
        L.orc_q1_records.restype = C.c_int64
        L.orc_copy_csr.argtypes = [C.c_void_p, p64, p32, p32]
        L.orc_contact_hash.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        L.orc_contact_hash.restype = C.c_uint64
        L.orc_neighbor_access.argtypes = [C.c_void_p, p32, C.c_int64, C.c_int32, C.c_int32, pu64, p64]
        L.orc_neighbor_access_mt.argtypes = [C.c_void_p, p32, C.c_int64, C.c_int32, C.c_int32, C.c_int, pu64, p64]
        Q = C.POINTER(_Q)
        L.orc_degree.argtypes = [C.c_void_p, Q, p32, p32]
        L.orc_bfs.argtypes = [C.c_void_p, Q, C.c_int32, p32]
        L.orc_earliest_arrival.argtypes = [C.c_void_p, Q, C.c_int32, p32]
        L.orc_latest_departure.argtypes = [C.c_void_p, Q, C.c_int32, p32]
        L.orc_cc.argtypes = [C.c_void_p, Q, p32]
        L.orc_gc.argtypes = [C.c_void_p, Q, p32, C.c_int32, p32]
        L.orc_pagerank_f64.argtypes = [C.c_void_p, Q, C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.orc_pagerank_f32.argtypes = [C.c_void_p, Q, C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_float)]
        L.orc_hits_f64.argtypes = [C.c_void_p, Q, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_hits_f32.argtypes = [C.c_void_p, Q, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_decode_codes.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, p32, C.c_int64, pu64, pu64]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def decode_codes(buf: bytes, pos: int, kinds):
    """Reads codes (0 unary, 1 gamma, k zeta_k, 64 a 64-bit integer) from bit `pos`."""
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    out = np.zeros(len(k), np.uint64)
    end = C.c_uint64()
    if lib().orc_decode_codes(buf, len(buf), pos, _p(k, C.c_int32), len(k), _p(out, C.c_uint64), C.byref(end)):
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())
    return [int(v) for v in out], end.value


def contact_hash(x: int, r: int, t: int) -> int:
    return lib().orc_contact_hash(x, r, t)


FULL_WINDOW = (-(1 << 31), (1 << 31) - 1)


class Oracle:
    """CPU restatement of the reference path over one dataset directory."""

    def __init__(self, directory: str, name: str, n: int):
        self.n = n
        self.h = lib().orc_open(directory.encode(), name.encode(), n)
        if not self.h:
            raise RuntimeError("oracle open failed: " + lib().orc_last_error().decode())
        self._csr = None

    def close(self):
        if self.h:
            lib().orc_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc < 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        return rc

    def widths(self):
        lg, lt, mt = C.c_int(), C.c_int(), C.c_int64()
        lib().orc_widths(self.h, C.byref(lg), C.byref(lt), C.byref(mt))
        return lg.value, lt.value, mt.value

    def offsets(self, x: int):
        g, t = C.c_uint64(), C.c_uint64()
        self._chk(lib().orc_offsets(self.h, x, C.byref(g), C.byref(t)))
        return g.value, t.value

    def csr(self):
        """Decodes every vertex; returns (start[n+1], adj[m], tim[m])."""
        if self._csr is None:
            self._chk(lib().orc_decode_all(self.h))
            m = lib().orc_num_contacts(self.h)
            start = np.empty(self.n + 1, np.int64)
            adj = np.empty(m, np.int32)
            tim = np.empty(m, np.int32)
            lib().orc_copy_csr(self.h, _p(start, C.c_int64), _p(adj, C.c_int32), _p(tim, C.c_int32))
            self._csr = (start, adj, tim)
        return self._csr

    def q1_records(self) -> int:
        self.csr()
        return lib().orc_q1_records(self.h)

    def _q(self, window, active):
        ws, we = window if window is not None else FULL_WINDOW
        ab, ae = active if active is not None else (0, self.n)
        return C.byref(_Q(ws, we, ab, ae))

    def neighbor_access(self, verts=None, active=None):
        ab, ae = active if active is not None else (0, self.n)
        cs, cnt = C.c_uint64(), C.c_int64()
        if verts is None:
            v = np.zeros(1, np.int32)
            nv = 0
        else:
            v = np.ascontiguousarray(verts, np.int32)
            nv = len(v)
        self._chk(lib().orc_neighbor_access(self.h, _p(v, C.c_int32), nv, ab, ae, C.byref(cs), C.byref(cnt)))
        return cs.value, cnt.value

    def neighbor_access_mt(self, verts, threads, active=None):
        """neighbor_access over `threads` threads (for timing only)."""
        ab, ae = active if active is not None else (0, self.n)
        cs, cnt = C.c_uint64(), C.c_int64()
        v = np.ascontiguousarray(verts, np.int32)
        self._chk(lib().orc_neighbor_access_mt(self.h, _p(v, C.c_int32), len(v), ab, ae, threads, C.byref(cs),
                                               C.byref(cnt)))
        return cs.value, cnt.value

    def degree(self, window=None, active=None, centrality=False):
        self.csr()
        cnt = np.empty(self.n, np.int32)
        cen = np.empty(self.n, np.int32) if centrality else None
        self._chk(lib().orc_degree(self.h, self._q(window, active), _p(cnt, C.c_int32),
                                   _p(cen, C.c_int32) if cen is not None else C.POINTER(C.c_int32)()))
        return (cnt, cen) if centrality else cnt

    def bfs(self, src, window=None, active=None):
        self.csr()
        out = np.empty(self.n, np.int32)
        self._chk(lib().orc_bfs(self.h, self._q(window, active), src, _p(out, C.c_int32)))
        return out

    def earliest_arrival(self, src, window=None, active=None):
        self.csr()
        out = np.empty(self.n, np.int32)
        self._chk(lib().orc_earliest_arrival(self.h, self._q(window, active), src, _p(out, C.c_int32)))
        return out

    def latest_departure(self, z, window=None, active=None):
        self.csr()
        out = np.empty(self.n, np.int32)
        self._chk(lib().orc_latest_departure(self.h, self._q(window, active), z, _p(out, C.c_int32)))
        return out

    def cc(self, window=None, active=None):
        self.csr()
        out = np.empty(self.n, np.int32)
        self._chk(lib().orc_cc(self.h, self._q(window, active), _p(out, C.c_int32)))
        return out

    def gc(self, prio, rounds=512, window=None, active=None):
        self.csr()
        prio = np.ascontiguousarray(prio, np.int32)
        out = np.empty(self.n, np.int32)
        self._chk(lib().orc_gc(self.h, self._q(window, active), _p(prio, C.c_int32), rounds, _p(out, C.c_int32)))
        return out

    def pagerank(self, iters=20, damping=0.85, tol=0.0, window=None, active=None, dtype=np.float64):
        self.csr()
        out = np.empty(self.n, dtype)
        fn = lib().orc_pagerank_f64 if dtype == np.float64 else lib().orc_pagerank_f32
        ct = C.c_double if dtype == np.float64 else C.c_float
        it = self._chk(fn(self.h, self._q(window, active), iters, damping, tol, _p(out, ct)))
        return out, it

    def hits(self, iters=20, window=None, active=None, dtype=np.float64):
        self.csr()
        hub = np.empty(self.n, dtype)
        auth = np.empty(self.n, dtype)
        fn = lib().orc_hits_f64 if dtype == np.float64 else lib().orc_hits_f32
        ct = C.c_double if dtype == np.float64 else C.c_float
        self._chk(fn(self.h, self._q(window, active), iters, _p(hub, ct), _p(auth, ct)))
        return hub, auth
