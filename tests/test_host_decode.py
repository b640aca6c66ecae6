"""The engine's device decoder built for the CPU under sanitizers.

tests/host_decode/harness.cpp runs the kernels' own decode code (headers and
Elias-Fano offsets, the checkpoint build of long records, every work item
with its launch group's decoder depth) on the host, compiled with clang under
AddressSanitizer + UndefinedBehaviorSanitizer with every local initialised to
a pattern (-ftrivial-auto-var-init=pattern), and again optimised with every
local initialised to zero.  A decoder that read an uninitialised field,
indexed a private array out of range or read past a stream would abort or
give different outputs in the two builds; both must reproduce, per vertex,
the contact count and the sum of the checksum terms of the raw contact list.
(MemorySanitizer is not used: the system libstdc++ is not instrumented for
it and reports its own string copies.)
"""
import os
import subprocess

import numpy as np
import pytest

import bruteforce as B
from conftest import ROOT

CLANG = "/opt/rocm/llvm/bin/clang++"
SRC = os.path.join(ROOT, "tests", "host_decode", "harness.cpp")
ENGINE = os.path.join(ROOT, "compresstemporalgraph_amd", "csrc", "engine")
VARIANTS = {
    # locals start as a 0xAA.. pattern; reads past buffers and UB abort
    "asan_pattern": ["-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                     "-ftrivial-auto-var-init=pattern"],
    # locals start as zero, optimised as the kernels are: a read of an
    # uninitialised field gives a different value here than under the pattern
    "o3_zero": ["-O3", "-ftrivial-auto-var-init=zero"],
}


@pytest.fixture(scope="module", params=sorted(VARIANTS))
def harness(request, tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("ROCm clang not found")
    out = str(tmp_path_factory.mktemp("harness") / f"harness_{request.param}")
    cmd = [CLANG, "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", ENGINE,
           "-I", os.path.join(ROOT, "include"), "-g", *VARIANTS[request.param], SRC, "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def _run(harness, ds, tmp_path, stride=None, burst=None, col=False):
    out = str(tmp_path / f"{ds.name}.bin")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    cmd = [harness, ds.dir, ds.name, out]
    if stride or burst or col:
        cmd += [str(stride or 0), str(burst or 4)] + (["1"] if col else [])
    p = subprocess.run(cmd, capture_output=True, text=True, env=env)
    assert p.returncode == 0, p.stderr[-4000:]
    raw = np.fromfile(out, np.uint64)
    n = ds.n
    cnt = raw[:n].astype(np.int64)
    sums = raw[n:2 * n]
    err = int(raw[2 * n])
    return cnt, sums, err


def _raw_sums(ds):
    h = B.contact_hashes(ds.src, ds.dst, ds.t)
    sums = np.zeros(ds.n, np.uint64)
    np.add.at(sums, ds.src.astype(np.int64), h)
    return np.bincount(ds.src, minlength=ds.n).astype(np.int64), sums


@pytest.mark.parametrize("which", ["ds_sample", "ds_adv", "ds_cl", "ds_nf", "ds_rmat", "ds_ext"])
def test_host_decode_matches_raw(harness, which, request, tmp_path):
    ds = request.getfixturevalue(which)
    cnt, sums, err = _run(harness, ds, tmp_path)
    assert err == 0
    want_cnt, want_sums = _raw_sums(ds)
    np.testing.assert_array_equal(cnt, want_cnt)
    np.testing.assert_array_equal(sums, want_sums)


@pytest.mark.parametrize("stride,burst", [(16, 4), (16, 32), (1024, 16), (128, 1)])
@pytest.mark.parametrize("which", ["ds_adv", "ds_nf"])
def test_host_decode_flat_strides(harness, which, stride, burst, request, tmp_path):
    """The same with the flat-record checkpoint stride forced to 16 (most
    records chunked, many split runs) and to 1024 (the large-graph stride),
    and the walk's burst length from 1 to 32 (the autotuner's range)."""
    ds = request.getfixturevalue(which)
    cnt, sums, err = _run(harness, ds, tmp_path, stride, burst)
    assert err == 0
    want_cnt, want_sums = _raw_sums(ds)
    np.testing.assert_array_equal(cnt, want_cnt)
    np.testing.assert_array_equal(sums, want_sums)


@pytest.mark.parametrize("stride,burst", [(None, None), (16, 4), (1024, 16)])
@pytest.mark.parametrize("which", ["ds_sample", "ds_adv", "ds_nf", "ds_ext"])
def test_host_decode_columns(harness, which, stride, burst, request, tmp_path):
    """Timestamps read from the column copy the engine builds at open
    (groups of 64 list positions, interleaved 16-byte units, unused units
    filled with ones) instead of the linear stream: the same per-vertex
    counts and checksums."""
    ds = request.getfixturevalue(which)
    cnt, sums, err = _run(harness, ds, tmp_path, stride, burst, col=True)
    assert err == 0
    want_cnt, want_sums = _raw_sums(ds)
    np.testing.assert_array_equal(cnt, want_cnt)
    np.testing.assert_array_equal(sums, want_sums)


@pytest.mark.parametrize("col", [False, True])
def test_host_decode_flags_corrupt_timestamps(harness, ds_nf, col, tmp_path):
    """Zeroing a few 16-byte stretches of the .timestamps stream (codes with
    an over-long prefix, and every later code of the record misframed) must
    raise the code error (bit 16 of the error word): whole records end their
    codes exactly where the next record starts.  The flat stride is raised
    above every degree, so no record is chunked and the walk, not the
    checkpoint build, meets the corrupt codes."""
    import shutil
    d = tmp_path / "corrupt"
    shutil.copytree(ds_nf.dir, d)
    p = d / (ds_nf.name + ".timestamps")
    b = bytearray(p.read_bytes())
    for k in range(1, 9):
        at = len(b) * k // 9
        b[at:at + 16] = bytes(16)
    p.write_bytes(bytes(b))

    class Corrupt:
        pass
    c = Corrupt()
    c.dir, c.name, c.n = str(d), ds_nf.name, ds_nf.n
    assert np.bincount(ds_nf.src).max() < 1 << 14
    cnt, sums, err = _run(harness, c, tmp_path, 1 << 14, None, col=col)
    assert err & 16
