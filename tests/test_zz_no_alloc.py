"""rhip_draw_noalloc (the draw whose arena is the library's static scratch)
against rhip_draw, through the C-ABI: identical outputs on two no-allocation
calls in a row, the second with a smaller input, and a clear refusal of a
draw that does not fit the scratch."""
import numpy as np
import pytest

from softglrender_amd import _lib, scenes
from softglrender_amd.replay import replay

pytestmark = pytest.mark.gpu


def _run(renderer, sc, no_alloc):
    renderer.no_alloc = no_alloc
    try:
        outs, _ = replay(renderer, sc.stream.bytes(), layouts=sc.meta.get("layouts"))
    finally:
        renderer.no_alloc = False
    return outs


def _assert_identical(got, want, label):
    assert len(got) == len(want), label
    for i, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, (label, i, g.shape, w.shape)
        bad = int((g.view(np.uint8) != w.view(np.uint8)).sum())
        assert bad == 0, f"{label}: output {i} differs in {bad} bytes"


def test_no_alloc_draw_equals_draw_on_two_calls(renderer):
    # first a 640x480 4x-MSAA frame of 4,096 triangles, then a smaller 160x120
    # frame whose triangles clip: the second runs on what the first left behind
    big = scenes.scene_c1(msaa=True, read_msaa=True)
    small = scenes.scene_variant("msaa_clip")
    print("[gpu] no-alloc: reference draws", flush=True)
    want_big = _run(renderer, big, False)
    want_small = _run(renderer, small, False)
    print("[gpu] no-alloc: c1_msaa through rhip_draw_noalloc", flush=True)
    got_big = _run(renderer, big, True)
    print("[gpu] no-alloc: msaa_clip through rhip_draw_noalloc", flush=True)
    got_small = _run(renderer, small, True)
    _assert_identical(got_big, want_big, "c1_msaa")
    _assert_identical(got_small, want_small, "msaa_clip")
    assert int((want_big[0][..., 3] > 0).sum()) > 60000


def test_no_alloc_draw_refuses_what_does_not_fit(renderer):
    # a 1080p 4x-MSAA target needs 66 MB of fragment jobs alone, beyond the 64 MiB scratch
    sc = scenes.scene_c1(width=1920, height=1080, msaa=True, lon=8, lat=4)
    assert _lib.load().rhip_draw_noalloc_capacity() == 64 << 20
    with pytest.raises(_lib.RhipError) as e:
        _run(renderer, sc, True)
    assert e.value.code == -3 and "scratch" in str(e.value)
