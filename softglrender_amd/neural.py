"""The learned PBR shader's workflow on the C2 scene (no counterpart in SoftGLRender).

1. Render the C2 geometry with the plain image-based PBR shader from random
   cameras and with random materials, capturing a sample of the shaded
   fragments (shading inputs, plain GPU colour, FP64 plain colour).
2. Train the 24-64-64-3 network on the GPU (BF16 MFMA forward and backward,
   Adam on FP32 master weights).
3. Evaluate it on fragments from held-out cameras against the FP64 plain
   shader, per material; a material whose p99 error exceeds the threshold
   keeps the plain shader (per-material fallback: `apply_fallback` hands the
   list to the library, whose learned-shader draws of such a material, set by
   `renderer.neural_material`, then run the plain shader).

The point light of the C2 material is off for both shaders here: the
network's inputs (normal, view direction, N.V, roughness, metallic, albedo,
reflection encoding) carry no light position.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import scenes
from . import stream as S
from .replay import Replayer, parse

P99_LIMIT = 4.0  # 1/255 units: above it a material keeps the plain shader


def random_cameras(n, seed, radius=2.6):
    """Eye positions uniform on a sphere of `radius` around the object, looking at the origin."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * radius


def neural_materials(n, size, seed, base=None):
    """n (albedo, metal-rough) texture pairs: material 0 is `base` when given (the C2 maps), the others
    seeded noise with a random tint, roughness range and metallic coverage."""
    rng = np.random.default_rng(seed)
    mats = []
    for m in range(n):
        if m == 0 and base is not None:
            mats.append(base)
            continue
        tint = rng.uniform(0.3, 1.0, 3)
        r0, r1 = sorted(rng.uniform(0.05, 1.0, 2))
        metal_cut = rng.uniform(0.2, 0.9)
        s = int(rng.integers(1 << 20))
        alb = scenes.noise_texture(size, s, lambda q: np.concatenate(
            [scenes._to_u8((0.1 + 0.85 * q) * tint), np.full(q.shape[:2] + (1,), 255, np.uint8)], -1))
        mr = scenes.noise_texture(size, s + 1, lambda q: np.stack(
            [np.zeros(q.shape[:2], np.uint8), scenes._to_u8(r0 + (r1 - r0) * q[..., 0]),
             scenes._to_u8((q[..., 1] > metal_cut) * 0.9 + 0.05), np.full(q.shape[:2], 255, np.uint8)], -1))
        mats.append((alb, mr))
    return mats


@dataclass
class NeuralScene:
    scene: scenes.Scene
    draw_materials: list = field(default_factory=list)


def scene_views(cameras, materials, normal, width=1920, height=1080, freq=224, msaa=True, seed=3,
                shading=S.SHADING_PBR, verts=None, reads=False):
    """The C2 geometry, IBL cubes and normal map; one render pass per (camera, material) pair with the
    image-based PBR material (point light off). Textures of material m are ids 10 + 2m (albedo) and
    11 + 2m (metal-rough)."""
    st = S.Stream()
    if verts is None:
        verts = scenes.icosphere(freq, 1.0, displace=0.1, seed=seed)
    v, idx = verts
    irr, pref = scenes.ibl_maps(seed + 4)
    mip = dict(mipmaps=True, filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR,
               wrap=S.WRAP_REPEAT, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    st.texture(2, normal.shape[1], normal.shape[0], **mip)
    st.tex_data(2, [normal])
    for m, (alb, mr) in enumerate(materials):
        st.texture(10 + 2 * m, alb.shape[1], alb.shape[0], **mip)
        st.tex_data(10 + 2 * m, [alb])
        st.texture(11 + 2 * m, mr.shape[1], mr.shape[0], **mip)
        st.tex_data(11 + 2 * m, [mr])
    st.texture(4, 32, 32, type=S.TEX_CUBE, filter_min=S.FILTER_LINEAR, filter_mag=S.FILTER_LINEAR,
               usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    st.tex_data(4, irr)
    st.texture(5, 128, 128, type=S.TEX_CUBE, mipmaps=True, filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR,
               filter_mag=S.FILTER_LINEAR, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    st.tex_data(5, pref[0])
    for lev in range(1, len(pref)):
        for face in range(6):
            st.tex_level(5, face, lev, pref[lev][face])
    st.vao(1, v, idx)
    st.texture(6, width, height, usage=S.USAGE_COLOR | S.USAGE_SAMPLER, msaa=msaa)
    st.tex_init(6)
    st.texture(7, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH, msaa=msaa)
    st.tex_init(7)
    st.fbo(1, color=6, depth=7)
    st.frame()
    proj = scenes.projection(math.radians(60.0), width / height, 0.1, reverse_z=True)
    draws = []
    model = np.eye(4)
    rs = scenes._common_states(S.DEPTH_GREATER, cull=True)
    for eye in cameras:
        up = (0, 1, 0) if abs(eye[1]) < 0.99 * np.linalg.norm(eye) else (0, 0, 1)
        view = scenes.look_at(eye, (0, 0, 0), up)
        scene_u = S.uniforms_scene((0.1, 0.1, 0.1), tuple(eye), (1.6, 2.0, 2.4), (2.0, 2.0, 2.0))
        for m in range(len(materials)):
            st.begin(1, True, True, (0.02, 0.02, 0.05, 1.0), 0.0)
            st.viewport(0, 0, width, height)
            st.draw(1, shading, S.DEF_ALBEDO | S.DEF_NORMAL | S.DEF_METALROUGH, rs,
                    S.uniforms_model(model, proj @ view @ model, reverse_z=True), scene_u,
                    S.uniforms_material(0, 1, 0, 1.0, 0.5, (1, 1, 1, 1)),
                    samplers=[(S.SLOT_ALBEDO, 10 + 2 * m), (S.SLOT_NORMAL, 2), (S.SLOT_METALROUGH, 11 + 2 * m),
                              (S.SLOT_IRRADIANCE, 4), (S.SLOT_PREFILTER, 5)])
            st.end()
            draws.append(m)
    if reads:
        st.read(6)
    sc = scenes.Scene("neural_views", st, width, height, msaa, {"layouts": {2: 2}})
    return NeuralScene(sc, draws)


def capture(renderer, ns: NeuralScene, shift):
    """Replays the scene with fragment capture on (1 in 2^shift fragments of every draw); returns the
    (first, count) range of records it added."""
    first = renderer.neural_record_count()
    rp = Replayer(renderer, layouts=ns.scene.meta.get("layouts"))
    sink = []
    k = 0
    in_frame = False
    for op, p in parse(ns.scene.stream.bytes()):
        if op == S.OP_FRAME:
            in_frame = True
            continue
        if op == S.OP_READ:
            continue
        if in_frame and op == S.OP_DRAW:
            renderer.neural_capture(True, shift, ns.draw_materials[k])
            k += 1
        rp.exec(op, p, sink)
    renderer.neural_capture(False)
    renderer.waitIdle()
    return first, renderer.neural_record_count() - first


def apply_fallback(renderer, table):
    """Makes the learned-shader draws of every material the table flags use the plain shader; returns those
    materials."""
    bad = [t["material"] for t in table if t["fallback"]]
    renderer.neural_fallback(bad)
    return bad


def fallback_table(renderer, first, count, n_materials, limit=P99_LIMIT):
    """Per material: network and plain-shader errors on records [first, first + count) and whether the
    material keeps the plain shader (network p99 above `limit`)."""
    rows = []
    for m in range(n_materials):
        e_nn, e_pl, _ = renderer.neural_evaluate(first, count, material=m)
        rows.append({"material": m, "network": e_nn, "plain": e_pl, "fallback": e_nn["p99"] > limit})
    return rows
