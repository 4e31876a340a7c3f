"""librhip's C-ABI: the built library loads and exports every entry point
include/rhip.h declares; without a GPU the device call fails cleanly."""
import ctypes
import os
import re

import pytest

from softglrender_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "rhip.h")).read()
    return sorted(set(re.findall(r"^(?:int|size_t|const char \*|void \*)\s*(rhip_\w+)\(", text, re.M)))


def test_header_declares_the_renderer_interface():
    names = _declared()
    for n in ["rhip_device_create", "rhip_texture_create", "rhip_vao_create", "rhip_fbo_create", "rhip_program_create",
              "rhip_begin_render_pass", "rhip_set_viewport", "rhip_draw", "rhip_end_render_pass", "rhip_wait_idle"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("librhip not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared()) <= set(_lib.SIGNATURES)


def test_invalid_arguments_fail_cleanly_without_a_device():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("librhip not built")
    lib = _lib.load()
    assert lib.rhip_device_create(0, None) == -1
    assert b"null" in lib.rhip_last_error()
    assert lib.rhip_draw(None, None) == -1
    assert lib.rhip_draw_noalloc(None, None) == -1
    assert lib.rhip_debug_raise_device_error(None, 0) == -1
    assert lib.rhip_draw_noalloc_capacity() == 64 << 20
    assert lib.rhip_texture_create(None, None, None) == -1
