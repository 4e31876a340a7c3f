"""Work a kernel drops reaches the caller: rhip_wait_idle and rhip_texture_read
return RHIP_E_DEVICE naming the stage once, then succeed again. The condition
is raised from the host by a kernel that sets the clipping stage's error word
(rhip_debug_raise_device_error), as the clipping kernel does on an overflow."""
import numpy as np
import pytest

from softglrender_amd import _lib
from softglrender_amd.renderer import TextureDesc

pytestmark = pytest.mark.gpu


def test_wait_idle_reports_a_dropped_triangle_once(renderer):
    lib = _lib.load()
    renderer.waitIdle()
    _lib.check(lib.rhip_debug_raise_device_error(renderer.dev, 0))
    with pytest.raises(_lib.RhipError) as e:
        renderer.waitIdle()
    assert e.value.code == -5 and "clipping" in str(e.value)
    renderer.waitIdle()  # reported once, then cleared
    assert lib.rhip_debug_raise_device_error(renderer.dev, 7) == -1


def test_readback_reports_a_dropped_triangle(renderer):
    lib = _lib.load()
    t = renderer.createTexture(TextureDesc(width=8, height=8))
    t.initImageData()
    _lib.check(lib.rhip_debug_raise_device_error(renderer.dev, 0))
    with pytest.raises(_lib.RhipError) as e:
        t.readImage()
    assert e.value.code == -5
    img = t.readImage()
    assert img.shape == (8, 8, 4) and not np.any(img)
