"""The C++ RendererHIP class (softglrender_amd/host): it is written against
SoftGLRender's interface headers, so it must compile against the real ones, and
driven through that interface it must render what the checker renders."""
import os
import subprocess

import numpy as np
import pytest

from softglrender_amd import scenes
from softglrender_amd.host_replay import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "Render")), reason="SoftGLRender sources not present")
@pytest.mark.parametrize("src", ["RendererHIP.cpp", "replay.cpp"])
def test_compiles_against_softglrender_headers(src):
    cmd = ["g++", "-std=gnu++11", "-fsyntax-only", "-DSOFTGL_SIMD_OPT", "-I", os.path.join(REF, "src"), "-I",
           os.path.join(REF, "third_party", "glm"), "-I", os.path.join(ROOT, "softglrender_amd", "host"), "-I",
           os.path.join(ROOT, "include"), os.path.join(ROOT, "softglrender_amd", "host", src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_library_exports_host_replay():
    if not os.path.exists(LIB_PATH):
        pytest.skip("librenderer_hip not built")
    import ctypes
    lib = ctypes.CDLL(LIB_PATH)
    assert hasattr(lib, "rhip_host_replay")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1", "c1_msaa", "c2_small", "v_blend_alpha", "v_near_clip", "v_shadow_blinn",
                                  "v_layout_morton", "v_depth_always_nomask"])
def test_cpp_renderer_identical_to_checker(name):
    import oracle
    from softglrender_amd.host_replay import host_replay
    from tests.golden_scenes import GOLDEN
    sc = GOLDEN[name]()
    print(f"[gpu] C++ RendererHIP {name}", flush=True)
    out, _ = host_replay(sc.stream.bytes(), sc.stream.output_bytes(), layouts=sc.meta.get("layouts"))
    ref, _ = oracle.cpu_ref_replay(sc.stream.bytes(), sc.stream.output_bytes(), exact_math=False)
    assert out == ref
