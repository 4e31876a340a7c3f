"""CPU checkers for RendererHIP -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (softglrender_amd/) imports, links or runs anything in
this package. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker or as the timed CPU baseline.

* cpu_ref   -- the project's own C++ restatement of RendererSoft's triangle path
               (oracle/cpu_ref.cpp), built into oracle/_build/libcpu_ref.so.
* refsoft   -- SoftGLRender's RendererSoft itself, compiled from the reference
               sources by oracle/Makefile into oracle/_ref/ (serial = the parity
               oracle; threaded = the shipped configuration, a timing baseline).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
BUILD_DIR = os.path.join(HERE, "_build")
_libs = {}


def build(with_reference: bool = True) -> None:
    """Compiles the restatement, and the reference when its sources are present."""
    subprocess.run(["make", "-s", "-C", HERE, "cpu_ref"], check=True)
    if with_reference and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-j4", "-C", HERE, "ref"], check=True)


def _load(path):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"checker library not built: {path}")
        lib = ctypes.CDLL(path)
        _libs[path] = lib
    return _libs[path]


def have_reference(threaded: bool = False) -> bool:
    name = "librefsoft_threaded.so" if threaded else "librefsoft_serial.so"
    return os.path.exists(os.path.join(REF_DIR, name))


def _replay(fn, stream_bytes: bytes, out_nbytes: int, repeats: int):
    buf = ctypes.create_string_buffer(max(out_nbytes, 1))
    times = (ctypes.c_double * max(repeats, 1))()
    fn.restype = ctypes.c_longlong
    fn.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_double)]
    n = fn(stream_bytes, len(stream_bytes), buf, out_nbytes, repeats, times)
    if n < 0:
        raise RuntimeError(f"replay failed with status {n}")
    return buf.raw[:n], list(times)


def ref_replay(stream_bytes: bytes, out_nbytes: int, repeats: int = 1, threaded: bool = False):
    """Runs the stream through SoftGLRender's RendererSoft. Returns (output bytes, per-frame ms)."""
    name = "librefsoft_threaded.so" if threaded else "librefsoft_serial.so"
    lib = _load(os.path.join(REF_DIR, name))
    return _replay(lib.ref_replay, stream_bytes, out_nbytes, repeats)


def cpu_ref_replay(stream_bytes: bytes, out_nbytes: int, repeats: int = 1, threads: int = 1, stats=None,
                   exact_math: bool = False, texel_bytes=None):
    """Runs the stream through the restatement. Returns (output bytes, per-frame ms).

    exact_math=False calls the host libm's powf/log2f/exp2f like the reference;
    exact_math=True uses correctly rounded versions, the arithmetic of the GPU path.

    `threads` > 1 splits fragment work over rows of 32-pixel blocks (same result).
    `stats`, if a dict, receives counters of the last frame (fragments shaded etc.).
    `texel_bytes`, if a list, receives per draw of the last frame the bytes of the distinct texels its shaders
    read (4 per texel); counting runs the raster on one thread."""
    lib = _load(os.path.join(BUILD_DIR, "libcpu_ref.so"))
    lib.cpu_ref_set_threads(ctypes.c_int(threads))
    lib.cpu_ref_set_math(ctypes.c_int(1 if exact_math else 0))
    lib.cpu_ref_count_texels(ctypes.c_int(1 if texel_bytes is not None else 0))
    try:
        out = _replay(lib.cpu_ref_replay, stream_bytes, out_nbytes, repeats)
    finally:
        lib.cpu_ref_count_texels(ctypes.c_int(0))
    if texel_bytes is not None:
        arr = (ctypes.c_longlong * 4096)()
        n = lib.cpu_ref_draw_texel_bytes(arr, 4096)
        texel_bytes[:] = [int(arr[i]) for i in range(max(n, 0))]
    if stats is not None:
        arr = (ctypes.c_longlong * 8)()
        lib.cpu_ref_stats(arr)
        keys = ["fs_invocations", "covered_samples", "triangles_rasterized", "clipped_triangles",
                "culled_triangles", "quads", "texel_fetches", "draws"]
        stats.update({k: int(arr[i]) for i, k in enumerate(keys)})
    return out
