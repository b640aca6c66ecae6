"""MI355X-native engine for temporal analytics on WebGraph/ChronoGraph-compressed graphs.

* ``engine``   — ``TemporalGraph``: the seven reference workloads on one GPU
  through the C ABI of ``include/tga.h`` (``libtga.so``, gfx950 HIP).
* ``fmt``      — the host format kit (``libtga_format.so``): encoder, GPU
  index writer, generators, raw ``.gz`` I/O.
* ``datasets`` — named synthetic workloads in the reference's directory layout.
* ``build``    — in-tree build of both libraries.
"""
__all__ = ["engine", "fmt", "datasets", "build"]
