"""The environment shaders on the checker: IBL irradiance (IBLIrradianceSoft.h:74-101), prefiltered
environment (IBLPrefilterSoft.h:88-168), skybox from a cube or an equirectangular map (SkyboxSoft.h:66-96)
and FXAA (FxaaSoft.h:81-266).

The checker reproduces the reference's arithmetic for these shaders, the compiler's fused multiply-adds and
the host libm's sinf/cosf/log2f/atan2f included, so against the reference's outputs every byte must be
identical: on the stored fixtures, and on other seeds and sizes where the reference is built. Independently of
the reference, a constant environment pins the shaders' mathematics: the prefiltered map of a constant is that
constant, the irradiance map is the constant up to the sum's discretisation, the skybox shows it everywhere
and FXAA leaves a flat image alone."""
import os

import numpy as np
import pytest

import oracle
from softglrender_amd import scenes
from tests.util import load_golden

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")
ENV_FIXTURES = {"ibl_cube": lambda: scenes.scene_ibl(),
                "ibl_equirect_fxaa": lambda: scenes.scene_ibl(equirect=True, fxaa=True)}


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(os.path.join(oracle.BUILD_DIR, "libcpu_ref.so")):
        oracle.build(with_reference=False)


def _checker(sc):
    out, _ = oracle.cpu_ref_replay(sc.stream.bytes(), sc.stream.output_bytes(), threads=4)
    return sc.stream.split_outputs(out)


def _assert_identical(got, want, label):
    assert len(got) == len(want), label
    for i, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, (label, i)
        bad = int((g.view(np.uint8) != w.view(np.uint8)).sum())
        assert bad == 0, f"{label}: output {i} differs in {bad} bytes"


@pytest.mark.parametrize("name", sorted(ENV_FIXTURES))
def test_checker_identical_to_reference_fixture(name):
    _assert_identical(_checker(ENV_FIXTURES[name]()), load_golden(os.path.join(GOLDEN_DIR, f"{name}.npz")), name)


@pytest.mark.skipif(not oracle.have_reference(), reason="the reference RendererSoft (oracle/_ref) is not built here")
@pytest.mark.parametrize("kw", [dict(src_size=64, irr_size=16, pref_size=32, seed=11, width=128, height=80),
                                dict(src_size=64, seed=12, width=160, height=96, equirect=True, fxaa=True),
                                dict(seed=13, width=72, height=136, fxaa=True)])
def test_checker_identical_to_live_reference(kw):
    sc = scenes.scene_ibl(**kw)
    ref, _ = oracle.ref_replay(sc.stream.bytes(), sc.stream.output_bytes())
    _assert_identical(_checker(sc), sc.stream.split_outputs(ref), str(kw))


@pytest.mark.parametrize("equirect", [False, True])
def test_constant_environment(equirect):
    c = (150, 90, 40)
    sc = scenes.scene_ibl(flat=c, equirect=equirect, fxaa=True)
    outs = _checker(sc)
    irr, pref, colour, fx = outs[:6], outs[6:36], outs[36], outs[38]
    for f in pref:  # sum(c NdotL) / sum(NdotL): c to within float rounding, then truncated to a byte
        d = f[..., :3].astype(int) - np.array(c)
        assert d.min() >= -1 and d.max() <= 0, d
    for f in irr:  # PI sum(c cos sin) / N over the 0.025-rad grid: c within a few percent
        r = f[..., :3].astype(float) / np.array(c)
        assert 0.95 < r.min() and r.max() < 1.05, r
    # the skybox shows the constant wherever the sphere does not cover it, and FXAA changes nothing there
    sky = np.all(colour[..., :3] == np.array(c), axis=-1)
    assert sky.mean() > 0.5
    edge = np.zeros_like(sky)
    edge[1:-1, 1:-1] = sky[:-2, 1:-1] & sky[2:, 1:-1] & sky[1:-1, :-2] & sky[1:-1, 2:] & sky[:-2, :-2] & sky[2:, 2:] \
        & sky[:-2, 2:] & sky[2:, :-2] & sky[1:-1, 1:-1]
    assert np.array_equal(fx[edge], colour[edge])


def _hip(renderer, sc):
    from softglrender_amd.replay import replay
    outs, _ = replay(renderer, sc.stream.bytes(), layouts=sc.meta.get("layouts"))
    return outs


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [dict(),
                                dict(equirect=True, fxaa=True),
                                dict(src_size=128, irr_size=32, pref_size=64, seed=21, width=320, height=180,
                                     fxaa=True),
                                dict(src_size=128, seed=22, width=256, height=144, equirect=True, fxaa=True),
                                dict(flat=(150, 90, 40), equirect=True, fxaa=True)])
def test_hip_identical_to_checker(renderer, kw):
    sc = scenes.scene_ibl(**kw)
    print(f"[gpu] environment shaders {kw}", flush=True)
    out, _ = oracle.cpu_ref_replay(sc.stream.bytes(), sc.stream.output_bytes(), threads=8)
    _assert_identical(_hip(renderer, sc), sc.stream.split_outputs(out), str(kw))
