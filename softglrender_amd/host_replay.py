"""Runs a command stream through the C++ RendererHIP class (librenderer_hip.so:
rhip_host_replay), i.e. the Renderer interface a SoftGLRender Viewer calls."""
from __future__ import annotations

import ctypes
import os

from ._lib import LIB_DIR

LIB_PATH = os.path.join(LIB_DIR, "librenderer_hip.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"librenderer_hip is not built ({LIB_PATH} missing)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.rhip_host_replay.restype = ctypes.c_longlong
        _lib.rhip_host_replay.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    return _lib


def host_replay(data: bytes, out_nbytes: int, repeats: int = 1, device: int = 0, layouts: dict | None = None):
    """Returns (output bytes, per-frame ms)."""
    lib = load()
    pairs = [v for kv in (layouts or {}).items() for v in kv]
    arr = (ctypes.c_int * max(len(pairs), 1))(*pairs)
    buf = ctypes.create_string_buffer(max(out_nbytes, 1))
    times = (ctypes.c_double * max(repeats, 1))()
    n = lib.rhip_host_replay(data, len(data), buf, out_nbytes, repeats, times, device, arr, len(pairs) // 2)
    if n < 0:
        raise RuntimeError(f"rhip_host_replay failed ({n})")
    return buf.raw[:n], list(times)
