"""ctypes binding of librhip (include/rhip.h).

The library is built in-tree (softglrender_amd/lib/librhip.so) by
softglrender_amd.build. There is no fallback: if the library is missing or a
call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "librhip.so")

RHIP_ERRORS = {-1: "RHIP_E_INVALID", -2: "RHIP_E_HIP", -3: "RHIP_E_UNSUPPORTED", -4: "RHIP_E_OOM",
               -5: "RHIP_E_DEVICE"}


class RhipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{RHIP_ERRORS.get(code, code)}: {msg}")
        self.code = code


class RhsStates(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "blend", "blendFuncRgb", "blendSrcRgb", "blendDstRgb", "blendFuncAlpha", "blendSrcAlpha", "blendDstAlpha",
        "depthTest", "depthMask", "depthFunc", "cullFace", "primitiveType", "polygonMode")] + \
        [("lineWidth", ctypes.c_float)]


class TextureDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("type", ctypes.c_int), ("format", ctypes.c_int),
                ("usage", ctypes.c_uint32), ("use_mipmaps", ctypes.c_int), ("multi_sample", ctypes.c_int),
                ("layout", ctypes.c_int)]


class SamplerDesc(ctypes.Structure):
    _fields_ = [("filter_min", ctypes.c_int), ("filter_mag", ctypes.c_int), ("wrap_s", ctypes.c_int),
                ("wrap_t", ctypes.c_int), ("wrap_r", ctypes.c_int), ("border_color", ctypes.c_int)]


class ClearStates(ctypes.Structure):
    _fields_ = [("depth_flag", ctypes.c_int), ("color_flag", ctypes.c_int), ("clear_color", ctypes.c_float * 4),
                ("clear_depth", ctypes.c_float)]


class DrawDesc(ctypes.Structure):
    _fields_ = [("vao", ctypes.c_int), ("program", ctypes.c_int), ("states", RhsStates),
                ("uniform_model", ctypes.c_void_p), ("uniform_scene", ctypes.c_void_p),
                ("uniform_material", ctypes.c_void_p), ("sampler_tex", ctypes.c_int * 8)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_longlong) for n in ("fs_invocations", "samples_written", "triangles_binned",
                                                  "tile_pairs", "clipped_triangles", "draws",
                                                  "hiz_culled")]


class NnError(ctypes.Structure):
    _fields_ = [("mean", ctypes.c_double), ("p99", ctypes.c_double), ("max", ctypes.c_double),
                ("samples", ctypes.c_longlong)]


class BatchView(ctypes.Structure):
    _fields_ = [("fbo", ctypes.c_int), ("viewport", ctypes.c_int * 4), ("clear", ClearStates), ("draw", DrawDesc)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int), ("waves", ctypes.c_int)]


class BatchTiming(ctypes.Structure):
    _fields_ = [("config", BatchConfig), ("ms", ctypes.c_double)]


# every exported symbol of rhip.h, with its argument types
SIGNATURES = {
    "rhip_device_create": [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    "rhip_device_destroy": [ctypes.c_void_p],
    "rhip_last_error": [],
    "rhip_texture_create": [ctypes.c_void_p, ctypes.POINTER(TextureDesc), ctypes.POINTER(ctypes.c_int)],
    "rhip_texture_set_sampler": [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(SamplerDesc)],
    "rhip_texture_init": [ctypes.c_void_p, ctypes.c_int],
    "rhip_texture_upload": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t],
    "rhip_texture_upload_level": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_size_t],
    "rhip_texture_read": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                          ctypes.c_size_t],
    "rhip_texture_destroy": [ctypes.c_void_p, ctypes.c_int],
    "rhip_vao_create": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                        ctypes.POINTER(ctypes.c_int)],
    "rhip_vao_update": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t],
    "rhip_vao_destroy": [ctypes.c_void_p, ctypes.c_int],
    "rhip_fbo_create": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    "rhip_fbo_attach_color": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "rhip_fbo_attach_depth": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int],
    "rhip_fbo_destroy": [ctypes.c_void_p, ctypes.c_int],
    "rhip_program_create": [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)],
    "rhip_begin_render_pass": [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ClearStates)],
    "rhip_set_viewport": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "rhip_draw": [ctypes.c_void_p, ctypes.POINTER(DrawDesc)],
    "rhip_draw_noalloc": [ctypes.c_void_p, ctypes.POINTER(DrawDesc)],
    "rhip_draw_noalloc_capacity": [],
    "rhip_end_render_pass": [ctypes.c_void_p],
    "rhip_wait_idle": [ctypes.c_void_p],
    "rhip_debug_raise_device_error": [ctypes.c_void_p, ctypes.c_int],
    "rhip_set_hierarchical_z": [ctypes.c_void_p, ctypes.c_int],
    "rhip_libm_info": [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double)],
    "rhip_libm_eval": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p],
    "rhip_libm_eval_host": [ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p],
    "rhip_stats_get": [ctypes.c_void_p, ctypes.POINTER(Stats)],
    "rhip_stats_reset": [ctypes.c_void_p],
    "rhip_device_stream": [ctypes.c_void_p],
    "rhip_profile_enable": [ctypes.c_void_p, ctypes.c_int],
    "rhip_profile_query": [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)],
    "rhip_neural_param_count": [],
    "rhip_neural_init": [ctypes.c_void_p, ctypes.c_uint],
    "rhip_neural_set_weights": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
    "rhip_neural_get_weights": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
    "rhip_neural_records": [ctypes.c_void_p, ctypes.c_longlong],
    "rhip_neural_capture": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "rhip_neural_material": [ctypes.c_void_p, ctypes.c_int],
    "rhip_neural_fallback": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int],
    "rhip_neural_record_count": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)],
    "rhip_neural_read_records": [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_void_p],
    "rhip_neural_write_records": [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_void_p],
    "rhip_neural_gradient": [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p,
                             ctypes.POINTER(ctypes.c_double)],
    "rhip_neural_train": [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                          ctypes.c_float, ctypes.c_float, ctypes.c_uint, ctypes.c_void_p],
    "rhip_neural_evaluate": [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int,
                             ctypes.POINTER(NnError), ctypes.POINTER(NnError), ctypes.c_void_p],
    "rhip_batch_create": [ctypes.c_void_p, ctypes.POINTER(BatchView), ctypes.c_int, ctypes.POINTER(BatchConfig),
                          ctypes.POINTER(ctypes.c_int)],
    "rhip_batch_run": [ctypes.c_void_p, ctypes.c_int],
    "rhip_batch_tune": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(BatchTiming), ctypes.c_int,
                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(BatchConfig)],
    "rhip_batch_get_config": [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(BatchConfig)],
    "rhip_batch_set_config": [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(BatchConfig)],
    "rhip_batch_view_times": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int],
    "rhip_batch_destroy": [ctypes.c_void_p, ctypes.c_int],
}

_lib = None


def load(path: str = LIB_PATH):
    """Loads librhip. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"librhip is not built ({path} missing): run softglrender_amd.build.build()")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.rhip_last_error.restype = ctypes.c_char_p
    lib.rhip_device_stream.restype = ctypes.c_void_p
    lib.rhip_draw_noalloc_capacity.restype = ctypes.c_size_t
    _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        raise RhipError(rc, _lib.rhip_last_error().decode(errors="replace"))
    return rc
