"""Builds the native libraries of the package in-tree.

* ``libtga_format.so`` — host C++ format kit (encoder, index writer,
  generators), compiled with g++.
* ``libtga.so`` — the gfx950 HIP engine behind the C ABI of ``include/tga.h``,
  compiled with hipcc for ``--offload-arch=gfx950``.

Both land next to this file so that they travel with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

FORMAT_LIB = os.path.join(HERE, "libtga_format.so")
ENGINE_LIB = os.path.join(HERE, "libtga.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = "gfx950"


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build_format() -> str:
    src = os.path.join(CSRC, "format", "encoder.cpp")
    tmp = FORMAT_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-I", INCLUDE, src, "-o", tmp, "-lz", "-pthread"])
    os.replace(tmp, FORMAT_LIB)
    return FORMAT_LIB


def engine_sources() -> list[str]:
    eng = os.path.join(CSRC, "engine")
    return sorted(os.path.join(eng, f) for f in os.listdir(eng) if f.endswith(".hip") or f.endswith(".cpp"))


def build_engine() -> str:
    """Compiles every engine source to an object in parallel (rocPRIM's
    translation unit dominates), then links libtga.so."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    flags = [HIPCC, "-O3", "-std=c++17", "-fPIC", f"--offload-arch={OFFLOAD_ARCH}",
             # Packed-FP32 VALU forms are kept out of every kernel.
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument",
             "-I", INCLUDE, "-I", os.path.join(CSRC, "engine"), "-I", os.path.join(CSRC, "format")]
    tmp = ENGINE_LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="tga_build_") as d:
        srcs = engine_sources()
        objs = [os.path.join(d, os.path.basename(s) + ".o") for s in srcs]
        with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
            for f in [ex.submit(_run, flags + ["-c", s, "-o", o]) for s, o in zip(srcs, objs)]:
                f.result()
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={OFFLOAD_ARCH}", *objs, "-o", tmp])
    os.replace(tmp, ENGINE_LIB)
    return ENGINE_LIB


def build_all() -> None:
    build_format()
    build_engine()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"format": build_format, "engine": build_engine, "all": build_all}[w]()
        print("built", w)
