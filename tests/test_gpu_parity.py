"""RendererHIP on the GPU against the CPU checker and the reference's golden
outputs, through librhip's C-ABI (softglrender_amd.renderer drives it).

The GPU reproduces the host libm's powf/log2f/exp2f/atan2f (rhip_libm.h) and
the multiply-adds the reference's compiler fused in its shaders, so every output
must be identical bit for bit: against the checker calling the host libm like
the reference, against the reference's own outputs (fixtures), and against the
full-size C2 frame rendered live where the reference is built."""
import os

import numpy as np
import pytest

import oracle
from softglrender_amd import scenes
from softglrender_amd.replay import replay
from tests.golden_scenes import GOLDEN
from tests.util import compare, differing_values, load_golden

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _hip(renderer, sc, repeats=1):
    outs, _ = replay(renderer, sc.stream.bytes(), repeats=repeats, layouts=sc.meta.get("layouts"))
    return outs


def _oracle(sc):
    out, _ = oracle.cpu_ref_replay(sc.stream.bytes(), sc.stream.output_bytes(), exact_math=False)
    return sc.stream.split_outputs(out)


def assert_identical(got, want, label):
    for i, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, (label, i, g.shape, w.shape)
        bad = int((g.view(np.uint8) != w.view(np.uint8)).sum())
        assert bad == 0, f"{label}: output {i} differs in {bad} bytes"


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_hip_matches_reference_fixture(renderer, name):
    sc = GOLDEN[name]()
    print(f"[gpu] fixture {name}", flush=True)
    got = _hip(renderer, sc)
    want = load_golden(os.path.join(GOLDEN_DIR, f"{name}.npz"))
    assert_identical(got, want, name)


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_hip_identical_to_checker(renderer, name):
    sc = GOLDEN[name]()
    assert_identical(_hip(renderer, sc), _oracle(sc), name)


def test_hip_empty_draw_resolves(renderer):
    # after a frame with a sphere, a clear and a draw with no indices: the resolved image is the clear colour
    sc = scenes.scene_msaa_empty_draw()
    got = _hip(renderer, sc)
    assert_identical(got, _oracle(sc), "msaa_empty_draw")
    assert np.all(got[0].reshape(-1, 4) == np.array([63, 127, 191, 255], np.uint8))


@pytest.mark.parametrize("seed", [5, 6])
@pytest.mark.parametrize("name", ["blend_add_one", "msaa_clip", "depth_lequal", "layout_morton", "wrap_mirror"])
def test_hip_matches_oracle_new_seeds(renderer, name, seed):
    sc = scenes.scene_variant(name, width=200, height=150, seed=seed)
    print(f"[gpu] variant {name} seed {seed}", flush=True)
    assert_identical(_hip(renderer, sc), _oracle(sc), name)


@pytest.mark.parametrize("name", ["msaa_blend", "blend_alpha", "msaa_clip"])
def test_resolved_image_read_between_draws(renderer, name):
    """rhip_draw leaves its multisample resolve pending until something may read the resolved image: reading
    the target between two draws of one pass must see the first draw's resolve, as RendererSoft resolves after
    every draw (RendererSoft.cpp:161-163)."""
    sc = scenes.scene_variant(name, width=120, height=90, seed=3, read_between=True)
    print(f"[gpu] read between draws: {name}", flush=True)
    assert_identical(_hip(renderer, sc), _oracle(sc), name)


def test_hip_deterministic_and_repeatable(renderer):
    sc = scenes.scene_c2(480, 270, freq=40, tex_size=256, shadow_size=512, read_shadow=True, read_samples=True)
    a = _hip(renderer, sc, repeats=3)
    b = _hip(renderer, sc, repeats=1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_hip_c1_full(renderer):
    sc = scenes.scene_c1()
    assert_identical(_hip(renderer, sc), _oracle(sc), "c1")
    s = renderer.stats()
    assert s["fs_invocations"] > 0


def test_hip_c2_full_size(renderer):
    """The 1080p MSAA-4x, 1M-triangle frame against the checker at full size."""
    print("[gpu] c2 full size", flush=True)
    sc = scenes.scene_c2(read_shadow=True, read_samples=True)
    got = _hip(renderer, sc)
    want = _oracle(sc)
    dmis, cmax, cfrac = compare(got, want, "c2")
    print(f"[gpu] c2: depth mismatches {dmis}, colour max {cmax}, colour diff fraction {cfrac:.2e}", flush=True)
    assert_identical(got, want, "c2")


@pytest.mark.skipif(not oracle.have_reference(), reason="the reference RendererSoft (oracle/_ref) is not built here")
def test_hip_c2_full_size_against_reference(renderer):
    """The full C2 frame against SoftGLRender's own RendererSoft (serial build) on the same inputs: colour,
    depth, every MSAA sample and the shadow map identical."""
    print("[gpu] c2 full size against the reference", flush=True)
    sc = scenes.scene_c2(read_shadow=True, read_samples=True)
    got = _hip(renderer, sc)
    ref, _ = oracle.ref_replay(sc.stream.bytes(), sc.stream.output_bytes())
    want = sc.stream.split_outputs(ref)
    dmis, cmax, cfrac = compare(got, want, "c2")
    n = differing_values(got, want)
    print(f"[gpu] c2 vs reference: depth mismatches {dmis}, colour max {cmax}, differing values {n}", flush=True)
    assert dmis == 0 and n == 0, (dmis, cmax, n)
    assert_identical(got, want, "c2 vs reference")


def test_random_triangles_with_screen_covering(renderer):
    """Many small random triangles plus full-screen ones (which need clipping)."""
    from softglrender_amd import stream as S
    w, h = 320, 240
    v, i = scenes.random_triangles(5000, w, h, 21, 0.25, 16.0, 0.0, 1.0, big=20)
    st = S.Stream()
    st.texture(2, w, h, usage=S.USAGE_COLOR, msaa=True)
    st.tex_init(2)
    st.texture(3, w, h, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH, msaa=True)
    st.tex_init(3)
    st.fbo(1, color=2, depth=3)
    st.vao(1, v, i)
    st.frame()
    st.begin(1, True, True, (0, 0, 0, 0), 1.0)
    st.viewport(0, 0, w, h)
    rs = S.RenderStates(depth_test=True, depth_func=S.DEPTH_LESS)
    st.draw(1, S.SHADING_VERTEX_COLOR, 0, rs, S.uniforms_model(np.eye(4), np.eye(4)), bytes(64), bytes(48))
    st.end()
    st.read(2)
    st.read(3, which=1)
    st.read(2, which=1)
    sc = scenes.Scene("rand", st, w, h, True, {})
    assert_identical(_hip(renderer, sc), _oracle(sc), "random")


def test_c3_shape_reduced_with_hierarchical_z(renderer):
    """The C3 workload's shape at 640x360 (random small triangles, then screen-covering ones at random depths,
    per-vertex colour, MSAA-4x, LESS, no culling): identical to the checker, and the per-tile hierarchical-Z
    test skips most (tile, triangle) pairs."""
    sc = scenes.scene_c3(640, 360, 40_000, 64, seed=7, read_samples=True)
    print("[gpu] c3 reduced 640x360", flush=True)
    renderer.resetStats()
    got = _hip(renderer, sc)
    st = renderer.stats()
    out, _ = oracle.cpu_ref_replay(sc.stream.bytes(), sc.stream.output_bytes(), threads=min(16, os.cpu_count()),
                                   exact_math=False)
    assert_identical(got, sc.stream.split_outputs(out), "c3_reduced")
    assert st["hiz_culled"] > st["tile_pairs"] // 2, st


def _translucent_stack_scene(msaa, seed=11, w=320, h=240):
    """An opaque layer of random triangles, then a blended draw that writes no depth: random small
    triangles everywhere and 14 large ones stacked over the left of the screen, more fragments per
    pixel there than a tile can list."""
    from softglrender_amd import stream as S
    v0, i0 = scenes.random_triangles(1500, w, h, seed, 4.0, 400.0, 0.3, 1.0)
    v1, _ = scenes.random_triangles(3000, w, h, seed + 1, 1.0, 300.0, 0.0, 0.8)
    rng = np.random.default_rng(seed + 2)
    big = np.zeros((14 * 3, 8), np.float32)
    for k in range(14):
        x1 = rng.uniform(-0.3, 0.1)
        big[3 * k:3 * k + 3, 0:2] = [[-1.0, -1.0], [x1, -1.0], [-1.0, 3.0]]
        big[3 * k:3 * k + 3, 2] = rng.uniform(0.0, 0.5)
        big[3 * k:3 * k + 3, 4:8] = rng.uniform(0, 1, 4)
    v1 = np.concatenate([v1, big])
    v1[:, 7] = rng.uniform(0.2, 0.8, v1.shape[0])
    i1 = np.arange(v1.shape[0], dtype=np.int32).reshape(-1, 3)
    st = S.Stream()
    st.texture(2, w, h, usage=S.USAGE_COLOR, msaa=msaa)
    st.tex_init(2)
    st.texture(3, w, h, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH, msaa=msaa)
    st.tex_init(3)
    st.fbo(1, color=2, depth=3)
    st.vao(1, v0, i0)
    st.vao(2, v1, i1)
    st.frame()
    st.begin(1, True, True, (0, 0, 0, 0), 1.0)
    st.viewport(0, 0, w, h)
    ident = S.uniforms_model(np.eye(4), np.eye(4))
    st.draw(1, S.SHADING_VERTEX_COLOR, 0, S.RenderStates(depth_test=True, depth_func=S.DEPTH_LESS), ident,
            bytes(64), bytes(48))
    rs = S.RenderStates(depth_test=True, depth_func=S.DEPTH_LESS)
    rs.depth_mask = False
    rs.blend = True
    rs.set_blend_factor(S.BF_SRC_ALPHA, S.BF_ONE_MINUS_SRC_ALPHA)
    st.draw(2, S.SHADING_VERTEX_COLOR, 0, rs, ident, bytes(64), bytes(48))
    st.end()
    st.read(2)
    st.read(3, which=1 if msaa else 0)
    if msaa:
        st.read(2, which=1)
    return scenes.Scene("translucent_stack", st, w, h, msaa, {})


@pytest.mark.parametrize("msaa", [False, True])
def test_translucent_stack_listed_and_overflowing_tiles(renderer, msaa):
    """A draw that writes no depth is listed, shaded and blended per pixel in order; tiles with more
    fragments than the list holds take the ordered raster. Both must match the checker exactly."""
    sc = _translucent_stack_scene(msaa)
    print(f"[gpu] translucent stack msaa={msaa}", flush=True)
    assert_identical(_hip(renderer, sc), _oracle(sc), "translucent_stack")
