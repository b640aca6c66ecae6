"""ctypes binding of the host-side format kit (``libtga_format.so``).

The format kit writes the ChronoGraph/WebGraph dataset files the engine reads
unchanged (``.graph``, ``.offsets``, ``.properties``, ``.timestamps``) and the
four GPU Elias-Fano index files, generates the synthetic workloads, and reads
and writes raw 4-column ``.gz`` contact lists.  Function names and argument
meaning follow ``include/tga_format.h``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import build as _build

_LIB = None


class FormatError(RuntimeError):
    """Raised when a format-kit call returns a non-zero status."""


class _Stats(C.Structure):
    _fields_ = [
        ("nodes", C.c_int64), ("arcs", C.c_int64), ("min_timestamp", C.c_int64),
        ("graph_bits", C.c_uint64), ("timestamps_bits", C.c_uint64),
        ("copied_arcs", C.c_int64), ("multiple_arcs", C.c_int64),
        ("intervalised_arcs", C.c_int64), ("residual_arcs", C.c_int64),
        ("records_with_ref", C.c_int64), ("max_ref_depth", C.c_int32),
        ("l_graph", C.c_int32), ("l_timestamps", C.c_int32), ("pad_", C.c_int32),
    ]


@dataclass
class EncodeStats:
    nodes: int
    arcs: int
    min_timestamp: int
    graph_bits: int
    timestamps_bits: int
    copied_arcs: int
    multiple_arcs: int
    intervalised_arcs: int
    residual_arcs: int
    records_with_ref: int
    max_ref_depth: int
    l_graph: int
    l_timestamps: int


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_build.FORMAT_LIB):
            _build.build_format()
        lib = C.CDLL(_build.FORMAT_LIB)
        p32, p64 = C.POINTER(C.c_int32), C.POINTER(C.c_int64)
        lib.tgaf_last_error.restype = C.c_char_p
        lib.tgaf_encode.argtypes = [p32, p32, p64, C.c_int64, C.c_int64, C.c_char_p, C.c_char_p, C.POINTER(_Stats)]
        lib.tgaf_encode_graph.argtypes = [p32, p32, C.c_int64, C.c_int64, C.c_char_p, C.c_char_p,
                                          C.POINTER(C.c_uint64), C.POINTER(_Stats)]
        lib.tgaf_build_gpu_index.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(_Stats)]
        lib.tgaf_gen_rmat.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64,
                                      C.c_uint64, C.c_int64, p32, p32, p64]
        lib.tgaf_gen_chung_lu.argtypes = [C.c_int64, C.c_double, C.c_int64, C.c_int64, C.c_uint64, C.c_int64,
                                          p32, p32, p64]
        lib.tgaf_gen_netflow.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_uint64,
                                         C.c_int64, p32, p32, p64]
        lib.tgaf_contact_checksum.argtypes = [p32, p32, p64, C.c_int64, C.POINTER(C.c_uint64)]
        lib.tgaf_write_raw_gz.argtypes = [C.c_char_p, p32, p32, p64, C.c_int64]
        lib.tgaf_read_raw_gz.argtypes = [C.c_char_p, C.c_int, p64, p32, p32, p64]
        lib.tgaf_decode_table.argtypes = [C.c_int, p32]
        lib.tgaf_java_random_ints.argtypes = [C.c_int64, C.c_int32, C.c_int64, p32]
        lib.tgaf_encode_codes.argtypes = [p32, C.POINTER(C.c_uint64), C.c_int64, C.POINTER(C.c_uint8), C.c_int64,
                                          C.POINTER(C.c_uint64)]
        _LIB = lib
    return _LIB


def _check(rc: int) -> None:
    if rc != 0:
        raise FormatError(f"format kit error {rc}: {_lib().tgaf_last_error().decode()}")


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _stats(s: _Stats) -> EncodeStats:
    return EncodeStats(**{k: getattr(s, k) for k in EncodeStats.__dataclass_fields__})


def encode(src, dst, t, directory: str, name: str, n_min: int = 0) -> EncodeStats:
    """Encodes contacts sorted by (src, dst) into ``directory/name.*`` plus the GPU index files."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    t = np.ascontiguousarray(t, dtype=np.int64)
    if not (len(src) == len(dst) == len(t)):
        raise ValueError("src, dst and t must have equal length")
    os.makedirs(directory, exist_ok=True)
    st = _Stats()
    _check(_lib().tgaf_encode(_p(src, C.c_int32), _p(dst, C.c_int32), _p(t, C.c_int64), len(src), n_min,
                              directory.encode(), name.encode(), C.byref(st)))
    return _stats(st)


def encode_graph(src, dst, directory: str, name: str, n_min: int = 0):
    """The graph half of ``encode``: ``directory/name.graph`` and
    ``.properties`` only.  Returns (goff, stats): the n + 1 record bit
    offsets of the graph stream (uint64) and the statistics (timestamp fields
    zero)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    if len(src) != len(dst):
        raise ValueError("src and dst must have equal length")
    n = max(int(n_min), int(max(src.max(), dst.max())) + 1 if len(src) else 0)
    os.makedirs(directory, exist_ok=True)
    goff = np.empty(n + 1, np.uint64)
    st = _Stats()
    _check(_lib().tgaf_encode_graph(_p(src, C.c_int32), _p(dst, C.c_int32), len(src), n_min, directory.encode(),
                                    name.encode(), _p(goff, C.c_uint64), C.byref(st)))
    return goff, _stats(st)


def ef_low_width(stream_bytes: int, n: int) -> int:
    """Lower-bit width of a GPU index: floor(log2(bits / MAX_LABEL)), MAX_LABEL
    = n - 1 (bfs.cu:1785-1786); raises for a log2(0)."""
    q = stream_bytes * 8 // (n - 1)
    if q == 0:
        raise FormatError("stream has fewer bits than MAX_LABEL: log2(0) in the index width")
    return q.bit_length() - 1


def build_gpu_index(directory: str, name: str) -> EncodeStats:
    st = _Stats()
    _check(_lib().tgaf_build_gpu_index(directory.encode(), name.encode(), C.byref(st)))
    return _stats(st)


def gen_rmat(scale: int, m: int, a=0.57, b=0.19, c=0.19, p_geo=1.0 / 3.0, t_span=1 << 20, seed=1):
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    t = np.empty(m, np.int64)
    _check(_lib().tgaf_gen_rmat(scale, a, b, c, p_geo, t_span, seed, m, _p(src, C.c_int32), _p(dst, C.c_int32),
                                _p(t, C.c_int64)))
    return src, dst, t


def gen_chung_lu(n: int, m: int, exponent=2.1, t_span=1 << 24, period=(1 << 24) // 64, seed=2):
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    t = np.empty(m, np.int64)
    _check(_lib().tgaf_gen_chung_lu(n, exponent, t_span, period, seed, m, _p(src, C.c_int32), _p(dst, C.c_int32),
                                    _p(t, C.c_int64)))
    return src, dst, t


def gen_netflow(n: int, m: int, exponent=2.5, count_min=10.0, gap_min=256.0, t_span=1 << 30, seed=4):
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    t = np.empty(m, np.int64)
    _check(_lib().tgaf_gen_netflow(n, exponent, count_min, gap_min, t_span, seed, m, _p(src, C.c_int32),
                                   _p(dst, C.c_int32), _p(t, C.c_int64)))
    return src, dst, t


def contact_checksum(src, dst, t) -> int:
    """The neighbor-access checksum of a raw contact list (what the decoding
    pass returns for its encoding)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    t = np.ascontiguousarray(t, dtype=np.int64)
    out = C.c_uint64()
    _check(_lib().tgaf_contact_checksum(_p(src, C.c_int32), _p(dst, C.c_int32), _p(t, C.c_int64), len(src),
                                        C.byref(out)))
    return out.value


def write_raw_gz(path: str, src, dst, t) -> None:
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    t = np.ascontiguousarray(t, dtype=np.int64)
    _check(_lib().tgaf_write_raw_gz(path.encode(), _p(src, C.c_int32), _p(dst, C.c_int32), _p(t, C.c_int64),
                                    len(src)))


def read_raw_gz(path: str, headers: bool = False):
    m = C.c_int64(0)
    null32, null64 = C.POINTER(C.c_int32)(), C.POINTER(C.c_int64)()
    _check(_lib().tgaf_read_raw_gz(path.encode(), int(headers), C.byref(m), null32, null32, null64))
    src = np.empty(m.value, np.int32)
    dst = np.empty(m.value, np.int32)
    t = np.empty(m.value, np.int64)
    _check(_lib().tgaf_read_raw_gz(path.encode(), int(headers), C.byref(m), _p(src, C.c_int32),
                                   _p(dst, C.c_int32), _p(t, C.c_int64)))
    return src, dst, t


def decode_table(kind: str) -> np.ndarray:
    out = np.empty(65536, np.int32)
    _check(_lib().tgaf_decode_table({"gamma": 0, "zeta3": 1}[kind], _p(out, C.c_int32)))
    return out


def java_random_ints(seed: int, bound: int, count: int) -> np.ndarray:
    out = np.empty(count, np.int32)
    _check(_lib().tgaf_java_random_ints(seed, bound, count, _p(out, C.c_int32)))
    return out


def encode_codes(kinds, values):
    """Writes codes (0 unary, 1 gamma, k zeta_k, 64 a 64-bit integer); returns (bytes, bits)."""
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    v = np.ascontiguousarray(np.array(values, dtype=np.uint64))
    cap = int(sum(64 if kk == 64 else 200 for kk in kinds) // 8 + 16)
    out = np.zeros(cap, np.uint8)
    nb = C.c_uint64()
    _check(_lib().tgaf_encode_codes(_p(k, C.c_int32), _p(v, C.c_uint64), len(k), _p(out, C.c_uint8), cap,
                                    C.byref(nb)))
    return bytes(out[:(nb.value + 7) // 8]), nb.value
