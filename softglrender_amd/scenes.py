"""Seeded synthetic scenes, recorded as rhip_stream.h command streams.

Each builder returns a `Scene` holding the stream plus metadata. The call
sequence mirrors what SoftGLRender's Viewer issues for a frame
(src/Viewer/Viewer.cpp:100-168 drawFrame/drawShadowMap, :388-396 pipelineDraw,
:597-626 setupPipelineStates, :684-716 uniform updates); the camera matrices
follow src/Viewer/Camera.cpp:26-69 (infinite-far projection, optional reverse Z).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

from . import stream as S


# ----------------------------------------------------------------------------- camera (Camera.cpp:26-69)

def projection(fovy_rad, aspect, near, reverse_z=False):
    t = 1.0 / math.tan(fovy_rad * 0.5)
    p = np.zeros((4, 4), np.float64)
    p[0, 0] = t / aspect
    p[1, 1] = t
    # glm is column-major: projection[c][r]; here p[r, c]
    if reverse_z:
        p[2, 2] = 0.0
        p[3, 2] = -1.0
        p[2, 3] = near
    else:
        p[2, 2] = -1.0
        p[3, 2] = -1.0
        p[2, 3] = -near
    return p


def look_at(eye, center, up):
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    v = np.eye(4)
    v[0, :3], v[1, :3], v[2, :3] = s, u, -f
    v[0, 3], v[1, 3], v[2, 3] = -s @ eye, -u @ eye, f @ eye
    return v


# ----------------------------------------------------------------------------- geometry

def _pack_vertices(pos, uv, nrm, tan):
    """Vertex (src/Viewer/Model.h:20-25) with GLM_FORCE_ALIGNED: 64 bytes = 16 floats."""
    n = pos.shape[0]
    v = np.zeros((n, 16), np.float32)
    v[:, 0:3] = pos
    v[:, 4:6] = uv
    v[:, 8:11] = nrm
    v[:, 12:15] = tan
    return v


def uv_sphere(lon=64, lat=32, radius=1.0):
    """UV sphere with (lon+1)*(lat+1) vertices and lon*lat*2 CCW (outward) triangles."""
    j, i = np.meshgrid(np.arange(lat + 1), np.arange(lon + 1), indexing="ij")
    u = i / lon
    v = j / lat
    phi = u * 2 * np.pi
    theta = v * np.pi  # 0 at north pole
    nrm = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), -np.sin(theta) * np.sin(phi)], -1).reshape(-1, 3)
    tan = np.stack([-np.sin(phi), np.zeros_like(phi), -np.cos(phi)], -1).reshape(-1, 3)
    pos = nrm * radius
    uv = np.stack([u, 1.0 - v], -1).reshape(-1, 2)
    idx = []
    for jj in range(lat):
        for ii in range(lon):
            a = jj * (lon + 1) + ii
            b = a + 1
            c = a + (lon + 1)
            d = c + 1
            idx.append((a, c, b))
            idx.append((b, c, d))
    return _pack_vertices(pos, uv, nrm, tan), np.asarray(idx, np.int32)


def _value_noise3(p, seed, freq):
    """Trilinear value noise on an integer lattice with smoothstep weights."""
    rng = np.random.default_rng(seed)
    n = 64
    table = rng.random((n, n, n)).astype(np.float64)
    q = p * freq
    i0 = np.floor(q).astype(np.int64)
    f = q - i0
    w = f * f * (3 - 2 * f)
    acc = np.zeros(p.shape[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                ix = (i0[:, 0] + dx) % n
                iy = (i0[:, 1] + dy) % n
                iz = (i0[:, 2] + dz) % n
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                acc += table[ix, iy, iz] * wx * wy * wz
    return acc


def icosphere(freq, radius=1.0, displace=0.0, seed=0):
    """Class-I geodesic icosphere: 20*freq^2 triangles, 10*freq^2+2 vertices, CCW outward.

    Optional radial displacement by `displace` x 3-octave value noise (seeded)."""
    t = (1.0 + 5 ** 0.5) / 2
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                     [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    faces = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                      [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                      [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    # barycentric grid per face
    ii, jj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    keep = ii + jj <= freq
    gi, gj = ii[keep], jj[keep]
    local = {(a, b): k for k, (a, b) in enumerate(zip(gi.tolist(), gj.tolist()))}
    tri_local = []
    for a in range(freq):
        for b in range(freq - a):
            tri_local.append((local[(a, b)], local[(a + 1, b)], local[(a, b + 1)]))
            if a + b + 1 < freq:
                tri_local.append((local[(a + 1, b)], local[(a + 1, b + 1)], local[(a, b + 1)]))
    tri_local = np.asarray(tri_local, np.int64)
    npf = gi.size
    all_pos = []
    all_tri = []
    for fi, (A, B, C) in enumerate(faces):
        wa = (freq - gi - gj) / freq
        wb = gi / freq
        wc = gj / freq
        p = wa[:, None] * base[A] + wb[:, None] * base[B] + wc[:, None] * base[C]
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        all_pos.append(p)
        all_tri.append(tri_local + fi * npf)
    pos = np.concatenate(all_pos)
    tri = np.concatenate(all_tri)
    key = np.round(pos * 1e6).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    pos = pos[first]
    tri = inv[tri]
    # orientation: make every triangle CCW seen from outside
    a, b, c = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    unit = pos.copy()
    if displace:
        n = sum(_value_noise3(unit * 0.5 + 3.0, seed + o, 2.0 ** (o + 1)) * 0.5 ** o for o in range(3))
        n = n / (1 + 0.5 + 0.25)
        pos = unit * (radius * (1.0 + displace * (n - 0.5)))[:, None]
    else:
        pos = unit * radius
    # area-weighted vertex normals
    a, b, c = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    fn = np.cross(b - a, c - a)
    nrm = np.zeros_like(pos)
    for k in range(3):
        np.add.at(nrm, tri[:, k], fn)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    phi = np.arctan2(-unit[:, 2], unit[:, 0])
    uv = np.stack([phi / (2 * np.pi) + 0.5, np.arcsin(np.clip(unit[:, 1], -1, 1)) / np.pi + 0.5], -1)
    tan = np.stack([-np.sin(phi), np.zeros_like(phi), -np.cos(phi)], -1)
    return _pack_vertices(pos, uv, nrm, tan), tri.astype(np.int32)


# ----------------------------------------------------------------------------- textures

def checker_texture(size=512, cells=8, seed=1):
    rng = np.random.default_rng(seed)
    colors = rng.integers(0, 256, size=(cells, cells, 3), dtype=np.int64).astype(np.uint8)
    cell = size // cells
    img = np.empty((size, size, 4), np.uint8)
    img[..., :3] = np.repeat(np.repeat(colors, cell, 0), cell, 1)
    img[..., 3] = 255
    return img


def noise_texture(size, seed, channels_fn):
    """2D value noise (4 octaves) mapped through channels_fn(noise[h,w,3]) -> uint8 (h,w,4)."""
    rng = np.random.default_rng(seed)
    acc = np.zeros((size, size, 3))
    amp = 1.0
    for o in range(4):
        g = 4 * 2 ** o
        lat = rng.random((g + 1, g + 1, 3))
        xs = np.linspace(0, g, size, endpoint=False)
        i0 = np.floor(xs).astype(int)
        f = xs - i0
        f = f * f * (3 - 2 * f)
        top = lat[i0][:, i0] * (1 - f)[None, :, None] + lat[i0][:, i0 + 1] * f[None, :, None]
        bot = lat[i0 + 1][:, i0] * (1 - f)[None, :, None] + lat[i0 + 1][:, i0 + 1] * f[None, :, None]
        acc += amp * (top * (1 - f)[:, None, None] + bot * f[:, None, None])
        amp *= 0.5
    acc /= 1.875
    return channels_fn(acc)


def _to_u8(x):
    return np.clip(np.floor(x * 255.0 + 0.5), 0, 255).astype(np.uint8)


# ----------------------------------------------------------------------------- scenes

@dataclass
class Scene:
    name: str
    stream: S.Stream
    width: int
    height: int
    msaa: bool
    meta: dict = field(default_factory=dict)


def _common_states(depth_func=S.DEPTH_LESS, cull=True, blend=False):
    rs = S.RenderStates()
    rs.blend = blend
    rs.set_blend_factor(S.BF_SRC_ALPHA, S.BF_ONE_MINUS_SRC_ALPHA)
    rs.depth_test = True
    rs.depth_mask = not blend
    rs.depth_func = depth_func
    rs.cull_face = cull
    return rs


def scene_c1(width=640, height=480, seed=1, msaa=False, lon=64, lat=32, read_msaa=False):
    """C1: 640x480 RGBA8 + f32 depth, no MSAA; UV sphere (64x32) at the origin, camera (0,0,3), fovy 60,
    near 0.1; seeded 512^2 checker with full mips, trilinear, REPEAT, Linear layout; Blinn-Phong with one
    directional light; depth LESS; back-face culling."""
    st = S.Stream()
    verts, idx = uv_sphere(lon, lat)
    tex = checker_texture(512, 8, seed)
    st.texture(1, 512, 512, mipmaps=True, filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR,
               wrap=S.WRAP_REPEAT, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    st.tex_data(1, [tex])
    st.vao(1, verts, idx)
    st.texture(2, width, height, usage=S.USAGE_COLOR | S.USAGE_SAMPLER, msaa=msaa,
               filter_min=S.FILTER_LINEAR, filter_mag=S.FILTER_LINEAR)
    st.tex_init(2)
    st.texture(3, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH | S.USAGE_SAMPLER, msaa=msaa)
    st.tex_init(3)
    st.fbo(1, color=2, depth=3)
    st.frame()
    eye = (0.0, 0.0, 3.0)
    view = look_at(eye, (0, 0, 0), (0, 1, 0))
    proj = projection(math.radians(60.0), width / height, 0.1)
    model = np.eye(4)
    mvp = proj @ view @ model
    st.begin(1, True, True, (0, 0, 0, 0), 1.0)
    st.viewport(0, 0, width, height)
    st.draw(1, S.SHADING_BLINN_PHONG_DIR, S.DEF_ALBEDO, _common_states(S.DEPTH_LESS, cull=True),
            S.uniforms_model(model, mvp),
            S.uniforms_scene((0.2, 0.2, 0.2), eye, (1.0, 1.0, 1.0), (0.8, 0.8, 0.8)),
            S.uniforms_material(1, 0, 0, 1.0, 0.5, (1, 1, 1, 1)),
            samplers=[(S.SLOT_ALBEDO, 1)])
    st.end()
    st.read(2, which=0)
    st.read(3, which=1 if msaa else 0)
    if msaa and read_msaa:
        st.read(2, which=1)
    return Scene("c1", st, width, height, msaa, dict(triangles=int(idx.shape[0]), vertices=int(verts.shape[0])))


def scene_msaa_empty_draw(width=96, height=64):
    """A 4x-MSAA frame of a shaded sphere, then a second pass that clears to a colour and draws a vertex array
    with no indices: the reference resolves after every draw to a multisampled colour target, even one with no
    triangles (RendererSoft.cpp:161-163), so the resolved image read last is the clear colour."""
    st = S.Stream()
    verts, idx = uv_sphere(16, 8)
    st.vao(1, verts, idx)
    st.vao(2, verts, np.zeros(0, np.int32))
    st.texture(2, width, height, usage=S.USAGE_COLOR | S.USAGE_SAMPLER, msaa=True)
    st.tex_init(2)
    st.texture(3, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH | S.USAGE_SAMPLER, msaa=True)
    st.tex_init(3)
    st.fbo(1, color=2, depth=3)
    st.frame()
    eye = (0.0, 0.0, 3.0)
    mvp = projection(math.radians(60.0), width / height, 0.1) @ look_at(eye, (0, 0, 0), (0, 1, 0))
    args = (_common_states(S.DEPTH_LESS), S.uniforms_model(np.eye(4), mvp),
            S.uniforms_scene((0.2, 0.2, 0.2), eye, (1.0, 1.0, 1.0), (0.8, 0.8, 0.8)),
            S.uniforms_material(1, 0, 0, 1.0, 0.5, (1, 1, 1, 1)))
    for vao, clear in ((1, (0, 0, 0, 0)), (2, (0.25, 0.5, 0.75, 1.0))):
        st.begin(1, True, True, clear, 1.0)
        st.viewport(0, 0, width, height)
        st.draw(vao, S.SHADING_BLINN_PHONG_DIR, 0, *args)
        st.end()
    st.read(2, which=0)
    return Scene("msaa_empty_draw", st, width, height, True)


def scene_c4(n_scenes=64, n_views=256, sizes=(256, 512, 1024), tri_lo=1_000, tri_hi=500_000, tex_size=1024,
             seed=9, reads=False):
    """C4: `n_views` views of `n_scenes` seeded scenes. Scene i is a displaced icosphere with a
    triangle count log-uniform in [tri_lo, tri_hi] (frequency round(sqrt(T / 20))) and its own
    tex_size^2 RGBA8 checker with full mips (Tiled layout, trilinear). View v shows scene v mod
    n_scenes at a size drawn from `sizes` (square), from an eye uniform on a radius-3 sphere looking at
    the origin (fovy 60, near 0.1); Blinn-Phong with a point light, depth LESS, back-face culling, no
    MSAA. Every view renders into its own colour and depth textures."""
    rng = np.random.default_rng(seed)
    st = S.Stream()
    layouts = {}
    tris = np.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi), n_scenes))
    freqs = [max(1, int(round(math.sqrt(t / 20.0)))) for t in tris]
    mip = dict(mipmaps=True, filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR,
               wrap=S.WRAP_REPEAT, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    n_tris = []
    for i, f in enumerate(freqs):
        v, idx = icosphere(f, 1.0, displace=0.1, seed=seed * 1000 + i)
        st.vao(1 + i, v, idx)
        n_tris.append(int(idx.shape[0]))
        tid = 1 + i
        st.texture(tid, tex_size, tex_size, **mip)
        st.tex_data(tid, [checker_texture(tex_size, 16, seed * 1000 + i)])
        layouts[tid] = 1  # Tiled
    view_sizes = rng.choice(np.asarray(sizes), n_views)
    d = rng.normal(size=(n_views, 3))
    eyes = 3.0 * d / np.linalg.norm(d, axis=1, keepdims=True)
    base = 1 + n_scenes
    for v in range(n_views):
        w = int(view_sizes[v])
        st.texture(base + 2 * v, w, w, usage=S.USAGE_COLOR | S.USAGE_SAMPLER)
        st.tex_init(base + 2 * v)
        st.texture(base + 2 * v + 1, w, w, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH)
        st.tex_init(base + 2 * v + 1)
        st.fbo(1 + v, color=base + 2 * v, depth=base + 2 * v + 1)
    st.frame()
    proj = projection(math.radians(60.0), 1.0, 0.1)
    model = np.eye(4)
    rs = _common_states(S.DEPTH_LESS, cull=True)
    for v in range(n_views):
        w = int(view_sizes[v])
        i = v % n_scenes
        eye = tuple(eyes[v])
        up = (0, 1, 0) if abs(eye[1]) < 0.99 * 3.0 else (0, 0, 1)
        view = look_at(eye, (0, 0, 0), up)
        st.begin(1 + v, True, True, (0.05, 0.05, 0.08, 1.0), 1.0)
        st.viewport(0, 0, w, w)
        st.draw(1 + i, S.SHADING_BLINN_PHONG, S.DEF_ALBEDO, rs, S.uniforms_model(model, proj @ view @ model),
                S.uniforms_scene((0.15, 0.15, 0.15), eye, (2.0, 3.0, 4.0), (1.0, 1.0, 1.0)),
                S.uniforms_material(1, 0, 0, 1.0, 0.5, (1, 1, 1, 1)), samplers=[(S.SLOT_ALBEDO, 1 + i)])
        st.end()
    if reads:
        for v in range(n_views):
            st.read(base + 2 * v, which=0)
            st.read(base + 2 * v + 1, which=0)
    pixels = int(sum(int(s) ** 2 for s in view_sizes))
    return Scene("c4", st, int(max(sizes)), int(max(sizes)), False,
                 dict(layouts=layouts, scene_triangles=n_tris, view_sizes=[int(s) for s in view_sizes],
                      view_scene=[v % n_scenes for v in range(n_views)], pixels=pixels,
                      view_triangles=int(sum(n_tris[v % n_scenes] for v in range(n_views)))))


# ----------------------------------------------------------------------------- IBL inputs

def _sh_basis(d):
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    return np.stack([np.full_like(x, 0.282095), 0.488603 * y, 0.488603 * z, 0.488603 * x, 1.092548 * x * y,
                     1.092548 * y * z, 0.315392 * (3 * z * z - 1), 1.092548 * x * z, 0.546274 * (x * x - y * y)], -1)


def _cube_dirs(size):
    """World directions of texel centres of the six faces, in the face order and
    orientation the reference's cube sampler uses (SamplerSoft.h:311-373)."""
    c = (np.arange(size) + 0.5) / size * 2 - 1
    vc, uc = np.meshgrid(c, c, indexing="ij")
    vc = -vc  # the sampler flips v
    one = np.ones_like(uc)
    faces = [np.stack([one, vc, -uc], -1), np.stack([-one, vc, uc], -1), np.stack([uc, one, -vc], -1),
             np.stack([uc, -one, vc], -1), np.stack([uc, vc, one], -1), np.stack([-uc, vc, -one], -1)]
    return [f / np.linalg.norm(f, axis=-1, keepdims=True) for f in faces]


def ibl_coefficients(seed=7):
    """Seeded smooth low-frequency radiance: order-2 spherical harmonics per RGB channel."""
    rng = np.random.default_rng(seed)
    coef = rng.normal(0, 0.25, size=(9, 3))
    coef[0] = rng.uniform(1.2, 1.8, size=3)
    return coef


def ibl_maps(seed=7, irr_size=32, pref_size=128, pref_levels=5):
    """RGBA8 irradiance cube and prefiltered specular mip chain (the reference's IBL map
    formats, Environment.h:20-22) for an SH radiance field. Irradiance uses the exact
    cosine-lobe convolution of the SH bands; level k of the prefiltered chain
    attenuates band l by exp(-r^2 l(l+1)) with r = k / (levels-1)."""
    coef = ibl_coefficients(seed)
    band = np.array([0, 1, 1, 1, 2, 2, 2, 2, 2])
    irr_w = np.array([np.pi, 2 * np.pi / 3, np.pi / 4])[band] / np.pi
    irr = [_to_u8(np.clip(_sh_basis(d) @ (coef * irr_w[:, None]), 0, 1) * 0.8) for d in _cube_dirs(irr_size)]
    irr = [np.concatenate([f, np.full(f.shape[:2] + (1,), 255, np.uint8)], -1) for f in irr]
    levels = []
    nlev = int(np.floor(np.log2(pref_size))) + 1
    for k in range(nlev):
        s = max(1, pref_size >> k)
        r = min(k, pref_levels - 1) / (pref_levels - 1)
        w = np.exp(-(r * r) * band * (band + 1))
        faces = [_to_u8(np.clip(_sh_basis(d) @ (coef * w[:, None]), 0, 1) * 0.8) for d in _cube_dirs(s)]
        levels.append([np.concatenate([f, np.full(f.shape[:2] + (1,), 255, np.uint8)], -1) for f in faces])
    return irr, levels


def env_cube_faces(size, seed=7):
    """The six RGBA8 faces of a seeded smooth low-frequency radiance cube (SH radiance, quantised to bytes:
    the reference's cube maps are RGBA8, Texture.h:63-66), face order and orientation as the cube sampler's."""
    coef = ibl_coefficients(seed)
    out = []
    for d in _cube_dirs(size):
        f = _to_u8(np.clip(_sh_basis(d) @ coef, 0, 1) * 0.8)
        out.append(np.concatenate([f, np.full(f.shape[:2] + (1,), 255, np.uint8)], -1))
    return out


def cube_mesh():
    """A cube of half-size 1 around the origin, 12 triangles of unshared vertices (the IBL and skybox mesh,
    ModelLoader.cpp:30-46), in the 64-byte Vertex layout."""
    q = [((1, 1, 1), (1, -1, 1), (1, -1, -1), (1, 1, -1)), ((-1, 1, -1), (-1, -1, -1), (-1, -1, 1), (-1, 1, 1)),
         ((-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)), ((-1, -1, 1), (-1, -1, -1), (1, -1, -1), (1, -1, 1)),
         ((-1, 1, 1), (-1, -1, 1), (1, -1, 1), (1, 1, 1)), ((1, 1, -1), (1, -1, -1), (-1, -1, -1), (-1, 1, -1))]
    pos = []
    for a, b, c, d in q:
        pos += [a, b, c, a, c, d]
    pos = np.array(pos, np.float32)
    z2 = np.zeros((len(pos), 2), np.float32)
    z3 = np.zeros((len(pos), 3), np.float32)
    return _pack_vertices(pos, z2, z3, z3), np.arange(len(pos), dtype=np.int32).reshape(-1, 3)


# IBLGenerator::drawCubeFaces (Environment.cpp:152-195): look-at targets and up vectors of the six faces
_CUBE_VIEWS = [((1, 0, 0), (0, -1, 0)), ((-1, 0, 0), (0, -1, 0)), ((0, 1, 0), (0, 0, 1)),
               ((0, -1, 0), (0, 0, -1)), ((0, 0, 1), (0, -1, 0)), ((0, 0, -1), (0, -1, 0))]


def ibl_precompute(st, vao, src_tex, irr_tex, pref_tex, src_size, irr_size, pref_size, pref_levels=5, first_fbo=100):
    """Appends the reference's IBL precompute (Environment.cpp:48-103, 152-195) to a stream: the irradiance
    convolution of `src_tex` into the faces of `irr_tex`, then GGX prefiltering into `pref_size`-px faces of
    `pref_levels` mip levels of `pref_tex`, each face one render pass of the cube mesh with a 90-degree
    camera at the origin (rotation only). Returns the next free framebuffer id."""
    proj = projection(math.radians(90.0), 1.0, 0.1)
    fid = first_fbo
    ident = np.eye(4)
    jobs = [(S.SHADING_IBL_IRRADIANCE, irr_tex, 0, irr_size, bytes(48))]
    for lev in range(pref_levels):
        rough = lev / (pref_levels - 1)
        mat = struct.pack("<2f", float(src_size), rough) + bytes(40)
        jobs.append((S.SHADING_IBL_PREFILTER, pref_tex, lev, max(1, pref_size >> lev), mat))
    for shading, tex, level, size, mat in jobs:
        for face, (center, up) in enumerate(_CUBE_VIEWS):
            view = look_at((0, 0, 0), center, up)
            view3 = np.eye(4)
            view3[:3, :3] = view[:3, :3]
            st.fbo(fid, color=tex, color_layer=face, color_level=level)
            st.begin(fid, True, False, (0, 0, 0, 0), 1.0)
            st.viewport(0, 0, size, size)
            st.draw(vao, shading, 0, S.RenderStates(), S.uniforms_model(ident, proj @ view3 @ ident),
                    bytes(64), mat, samplers=[(S.SLOT_CUBE, src_tex)])
            st.end()
            fid += 1
    return fid


def scene_ibl(src_size=32, irr_size=8, pref_size=16, pref_levels=5, seed=7, width=96, height=64, equirect=False,
              fxaa=False, flat=None):
    """IBL precompute (irradiance and prefiltered maps drawn from a mipmapped source cube), then a skybox pass
    over a shaded sphere (the cube map, or an equirectangular image with `equirect`), optionally followed by
    an FXAA pass over the result; every map and the final image are read back. `flat` (RGB bytes) makes
    every source texel that one colour."""
    st = S.Stream()
    cv, ci = cube_mesh()
    st.vao(1, cv, ci)
    trilinear = dict(filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR, wrap=S.WRAP_CLAMP_TO_EDGE)
    st.texture(10, src_size, src_size, type=S.TEX_CUBE, mipmaps=True, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD,
               **trilinear)
    faces = env_cube_faces(src_size, seed)
    if flat is not None:
        for f in faces:
            f[..., :3] = np.asarray(flat, np.uint8)
    st.tex_data(10, faces)
    st.texture(11, irr_size, irr_size, type=S.TEX_CUBE, usage=S.USAGE_COLOR | S.USAGE_SAMPLER,
               filter_min=S.FILTER_LINEAR, filter_mag=S.FILTER_LINEAR, wrap=S.WRAP_CLAMP_TO_EDGE)
    st.tex_init(11)
    st.texture(12, pref_size, pref_size, type=S.TEX_CUBE, mipmaps=True, usage=S.USAGE_COLOR | S.USAGE_SAMPLER,
               **trilinear)
    st.tex_init(12)
    if equirect:
        rng = np.random.default_rng(seed)
        eq = np.zeros((src_size, 2 * src_size, 4), np.uint8)
        eq[..., :3] = _to_u8(0.5 + 0.45 * np.sin(np.linspace(0, 6 * np.pi, 2 * src_size))[None, :, None]
                             * np.cos(np.linspace(0, 3 * np.pi, src_size))[:, None, None]
                             * rng.uniform(0.6, 1.0, (1, 1, 3)))
        eq[..., 3] = 255
        if flat is not None:
            eq[..., :3] = np.asarray(flat, np.uint8)
        st.texture(13, 2 * src_size, src_size, mipmaps=True, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD,
                   filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR, wrap=S.WRAP_REPEAT)
        st.tex_data(13, [eq])
    st.texture(2, width, height, usage=S.USAGE_COLOR | S.USAGE_SAMPLER, filter_min=S.FILTER_LINEAR,
               filter_mag=S.FILTER_LINEAR, wrap=S.WRAP_CLAMP_TO_EDGE)
    st.tex_init(2)
    st.texture(3, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH)
    st.tex_init(3)
    st.fbo(1, color=2, depth=3)
    sv, si = uv_sphere(24, 12, 0.6)
    st.vao(2, sv, si)
    if fxaa:
        st.texture(4, width, height, usage=S.USAGE_COLOR)
        st.tex_init(4)
        st.fbo(2, color=4)
        qv = np.zeros((4, 16), np.float32)
        qv[:, 0:2] = [[1, -1], [-1, -1], [1, 1], [-1, 1]]
        qv[:, 4:6] = [[1, 0], [0, 0], [1, 1], [0, 1]]
        st.vao(3, qv, np.array([[0, 1, 2], [1, 2, 3]], np.int32))
    ibl_precompute(st, 1, 10, 11, 12, src_size, irr_size, pref_size, pref_levels)
    st.frame()
    eye = (0.0, 0.4, 2.2)
    view = look_at(eye, (0, 0, 0), (0, 1, 0))
    proj = projection(math.radians(60.0), width / height, 0.1)
    st.begin(1, True, True, (0, 0, 0, 0), 1.0)
    st.viewport(0, 0, width, height)
    st.draw(2, S.SHADING_PBR, 0, _common_states(S.DEPTH_LESS), S.uniforms_model(np.eye(4), proj @ view),
            S.uniforms_scene((0.1, 0.1, 0.1), eye, (2.0, 2.0, 2.0), (1.0, 1.0, 1.0)),
            S.uniforms_material(0, 1, 0, 1.0, 0.5, (0.8, 0.6, 0.4, 1.0)),
            samplers=[(S.SLOT_IRRADIANCE, 11), (S.SLOT_PREFILTER, 12)])
    sky = S.RenderStates(depth_test=True, depth_func=S.DEPTH_LEQUAL)
    sky.depth_mask = False
    view3 = np.eye(4)
    view3[:3, :3] = view[:3, :3]
    st.draw(1, S.SHADING_SKYBOX, S.DEF_EQUIRECT if equirect else 0, sky, S.uniforms_model(np.eye(4), proj @ view3),
            bytes(64), bytes(48), samplers=[(S.SLOT_EQUIRECT, 13)] if equirect else [(S.SLOT_CUBE, 10)])
    st.end()
    if fxaa:
        st.begin(2, True, False, (0, 0, 0, 0), 1.0)
        st.viewport(0, 0, width, height)
        st.draw(3, S.SHADING_FXAA, 0, S.RenderStates(), S.uniforms_model(np.eye(4), np.eye(4)), bytes(64),
                struct.pack("<2f", float(width), float(height)) + bytes(40), samplers=[(S.SLOT_SCREEN, 2)])
        st.end()
    for face in range(6):
        st.read(11, layer=face)
    for lev in range(pref_levels):
        for face in range(6):
            st.read(12, layer=face, level=lev)
    st.read(2)
    st.read(3)
    if fxaa:
        st.read(4)
    return Scene("ibl", st, width, height, False, {})


def pbr_textures(size, seed):
    albedo = noise_texture(size, seed, lambda n: np.concatenate(
        [_to_u8(0.15 + 0.8 * n), np.full(n.shape[:2] + (1,), 255, np.uint8)], -1))

    def nmap(n):
        t = (n[..., :2] - 0.5) * 0.8
        v = np.concatenate([t, np.ones(t.shape[:2] + (1,))], -1)
        v /= np.linalg.norm(v, axis=-1, keepdims=True)
        return np.concatenate([_to_u8(v * 0.5 + 0.5), np.full(n.shape[:2] + (1,), 255, np.uint8)], -1)

    normal = noise_texture(size, seed + 1, nmap)
    mr = noise_texture(size, seed + 2, lambda n: np.stack(
        [np.zeros(n.shape[:2], np.uint8), _to_u8(0.2 + 0.7 * n[..., 0]), _to_u8((n[..., 1] > 0.55) * 0.9 + 0.05),
         np.full(n.shape[:2], 255, np.uint8)], -1))
    return albedo, normal, mr


def _translate(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m


def scene_c2(width=1920, height=1080, freq=224, tex_size=2048, shadow_size=2048, seed=3, msaa=True,
             layout="morton", read_shadow=False, read_samples=False, translucent=True, shadow=True):
    """C2: 1920x1080 MSAA-4x, reversed Z; displaced icosphere (freq 224 -> 1,003,520 triangles, 3-octave value
    noise, amplitude 0.1); PBR metallic-roughness with 2048^2 albedo/normal/metal-rough maps (full mips, Morton
    layout); IBL irradiance (32^2) and prefiltered (128^2, 5 used levels) RGBA8 cubes from a seeded SH radiance;
    a translucent sphere (alpha 0.5) blended over it; a 2048^2 shadow-map depth pass first."""
    st = S.Stream()
    verts, idx = icosphere(freq, 1.0, displace=0.1, seed=seed)
    tverts, tidx = icosphere(32, 0.35)
    albedo, normal, mr = pbr_textures(tex_size, seed)
    irr, pref = ibl_maps(seed + 4)
    lay = {"linear": 0, "tiled": 1, "morton": 2}[layout]
    mip = dict(mipmaps=True, filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR,
               wrap=S.WRAP_REPEAT, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    for tid, img in ((1, albedo), (2, normal), (3, mr)):
        st.texture(tid, tex_size, tex_size, **mip)
        st.tex_data(tid, [img])
    st.texture(4, 32, 32, type=S.TEX_CUBE, filter_min=S.FILTER_LINEAR, filter_mag=S.FILTER_LINEAR,
               usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    st.tex_data(4, irr)
    st.texture(5, 128, 128, type=S.TEX_CUBE, mipmaps=True, filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR,
               filter_mag=S.FILTER_LINEAR, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    st.tex_data(5, pref[0])
    for lev in range(1, len(pref)):
        for face in range(6):
            st.tex_level(5, face, lev, pref[lev][face])
    st.vao(1, verts, idx)
    st.vao(2, tverts, tidx)
    st.texture(6, width, height, usage=S.USAGE_COLOR | S.USAGE_SAMPLER, msaa=msaa, filter_min=S.FILTER_LINEAR,
               filter_mag=S.FILTER_LINEAR)
    st.tex_init(6)
    st.texture(7, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH | S.USAGE_SAMPLER, msaa=msaa)
    st.tex_init(7)
    st.fbo(1, color=6, depth=7)
    if shadow:
        st.texture(8, shadow_size, shadow_size, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH | S.USAGE_SAMPLER,
                   wrap=S.WRAP_CLAMP_TO_BORDER)
        st.tex_init(8)
        st.fbo(2, depth=8)
    st.frame()
    light = (1.6, 2.0, 2.4)
    eye = (0.0, 0.0, 2.6)
    model = np.eye(4)
    if shadow:
        lview = look_at(light, (0, 0, 0), (0, 1, 0))
        lproj = projection(math.radians(60.0), 1.0, 0.1, reverse_z=True)
        rs = _common_states(S.DEPTH_GREATER, cull=True)
        st.begin(2, False, True, (0, 0, 0, 0), 0.0)
        st.viewport(0, 0, shadow_size, shadow_size)
        st.draw(1, S.SHADING_BASIC, 0, rs, S.uniforms_model(model, lproj @ lview @ model, reverse_z=True),
                bytes(S.SCENE_BYTES), S.uniforms_material(0, 0, 0), samplers=[])
        st.end()
    view = look_at(eye, (0, 0, 0), (0, 1, 0))
    proj = projection(math.radians(60.0), width / height, 0.1, reverse_z=True)
    scene = S.uniforms_scene((0.1, 0.1, 0.1), eye, light, (2.0, 2.0, 2.0))
    st.begin(1, True, True, (0.02, 0.02, 0.05, 1.0), 0.0)
    st.viewport(0, 0, width, height)
    st.draw(1, S.SHADING_PBR, S.DEF_ALBEDO | S.DEF_NORMAL | S.DEF_METALROUGH,
            _common_states(S.DEPTH_GREATER, cull=True),
            S.uniforms_model(model, proj @ view @ model, reverse_z=True), scene,
            S.uniforms_material(1, 1, 0, 1.0, 0.5, (1, 1, 1, 1)),
            samplers=[(S.SLOT_ALBEDO, 1), (S.SLOT_NORMAL, 2), (S.SLOT_METALROUGH, 3), (S.SLOT_IRRADIANCE, 4),
                      (S.SLOT_PREFILTER, 5)])
    if translucent:
        tm = _translate((0.45, -0.3, 1.25))
        st.draw(2, S.SHADING_PBR, 0, _common_states(S.DEPTH_GREATER, cull=True, blend=True),
                S.uniforms_model(tm, proj @ view @ tm, reverse_z=True), scene,
                S.uniforms_material(1, 1, 0, 1.0, 0.5, (0.2, 0.6, 1.0, 0.5)),
                samplers=[(S.SLOT_IRRADIANCE, 4), (S.SLOT_PREFILTER, 5)])
    st.end()
    st.read(6, which=0)
    st.read(7, which=1 if msaa else 0)
    if read_samples and msaa:
        st.read(6, which=1)
    if read_shadow and shadow:
        st.read(8, which=0)
    return Scene("c2", st, width, height, msaa,
                 dict(triangles=int(idx.shape[0]), vertices=int(verts.shape[0]),
                      translucent_triangles=int(tidx.shape[0]), tex_size=tex_size, shadow_size=shadow_size,
                      layouts={1: lay, 2: lay, 3: lay}))


# ----------------------------------------------------------------------------- small state variants

def random_triangles(n, width, height, seed, area_lo=0.25, area_hi=16.0, z_lo=0.0, z_hi=1.0, big=0):
    """Screen-space random triangles in NDC (w = 1) with log-uniform pixel area, uniform
    centres and NDC depth, flat RGBA32F colours per vertex; plus `big` screen-covering
    triangles at random constant depths. Vertex layout {vec3 a_position; vec4 a_color}
    (8 floats). Returns vertices (3 per triangle, unshared) and indices."""
    rng = np.random.default_rng(seed)
    area = np.exp(rng.uniform(np.log(area_lo), np.log(area_hi), n))
    cx, cy = rng.uniform(0, width, n), rng.uniform(0, height, n)
    ang = rng.uniform(0, 2 * np.pi, (n, 1)) + np.array([0, 2.1, 4.2]) + rng.uniform(-0.4, 0.4, (n, 3))
    rad = np.sqrt(area / (0.5 * 3 * np.sin(2.1)))[:, None] * rng.uniform(0.8, 1.2, (n, 3))
    px = cx[:, None] + rad * np.cos(ang)
    py = cy[:, None] + rad * np.sin(ang)
    z = rng.uniform(z_lo, z_hi, n)
    pos = np.stack([px / width * 2 - 1, py / height * 2 - 1, np.repeat(z[:, None], 3, 1)], -1)
    col = rng.uniform(0, 1, (n, 3, 4)).astype(np.float32)
    col[..., 3] = 1.0
    if big:
        bz = rng.uniform(z_lo, z_hi, big)
        bp = np.stack([np.tile([[-1.0, -1.0], [3.0, -1.0], [-1.0, 3.0]], (big, 1, 1))[..., 0],
                       np.tile([[-1.0, -1.0], [3.0, -1.0], [-1.0, 3.0]], (big, 1, 1))[..., 1],
                       np.repeat(bz[:, None], 3, 1)], -1)
        bc = rng.uniform(0, 1, (big, 3, 4)).astype(np.float32)
        bc[..., 3] = 1.0
        pos = np.concatenate([pos, bp])
        col = np.concatenate([col, bc])
    m = pos.shape[0]
    v = np.zeros((m * 3, 8), np.float32)
    v[:, 0:3] = pos.reshape(-1, 3)
    v[:, 4:8] = col.reshape(-1, 4)
    return v, np.arange(m * 3, dtype=np.int32).reshape(-1, 3)


def scene_c3(width=3840, height=2160, n=2_000_000, big=2000, seed=5, msaa=True, read_samples=False, crop=None):
    """C3 binning/raster stress: `n` random triangles (screen area log-uniform in 0.25-16 px, centres uniform,
    NDC z uniform in [0, 1]) and `big` screen-covering triangles at random depths after them; per-vertex RGBA32F
    colours, depth LESS, no culling, 3840x2160 MSAA-4x.

    crop=(x0, y0, w, h) renders the same triangles into a w x h framebuffer through the viewport
    (-x0, -y0, width, height): the window of the full frame at (x0, y0), every triangle still drawn."""
    v, i = random_triangles(n, width, height, seed, 0.25, 16.0, 0.0, 1.0, big=big)
    fx, fy, fw, fh = crop if crop else (0, 0, width, height)
    st = S.Stream()
    st.texture(2, fw, fh, usage=S.USAGE_COLOR, msaa=msaa)
    st.tex_init(2)
    st.texture(3, fw, fh, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH, msaa=msaa)
    st.tex_init(3)
    st.fbo(1, color=2, depth=3)
    st.vao(1, v, i)
    st.frame()
    st.begin(1, True, True, (0, 0, 0, 0), 1.0)
    st.viewport(-fx, -fy, width, height)
    rs = S.RenderStates(depth_test=True, depth_func=S.DEPTH_LESS)
    st.draw(1, S.SHADING_VERTEX_COLOR, 0, rs, S.uniforms_model(np.eye(4), np.eye(4)), bytes(64), bytes(48))
    st.end()
    st.read(2)
    st.read(3, which=1)
    if read_samples:
        st.read(2, which=1)
    return Scene("c3", st, fw, fh, msaa, {"triangles": n + big})


VARIANTS = ["blend_alpha", "blend_add_one", "blend_minmax", "blend_sub", "depth_lequal", "depth_gequal_rz",
            "depth_always_nomask", "depth_equal", "depth_notequal", "depth_never", "cull_off", "near_clip",
            "side_clip", "filter_nearest", "filter_linear", "filter_nn", "filter_ln", "filter_nl", "wrap_mirror",
            "wrap_clamp_edge", "wrap_border_white", "shadow_blinn", "point_light", "vertex_color",
            "basic_nodepth", "viewport_offset", "msaa_blend", "msaa_clip", "msaa_near", "layout_tiled",
            "layout_morton", "depth_only_msaa", "ao_emissive", "normal_blinn", "normal_blinn_dir"]


def scene_variant(name, width=160, height=120, seed=11, read_between=False):
    """One small frame exercising one pipeline state (for fixtures and parity tests). With `read_between`, the
    colour target (and its resolved image at 4x) is also read between a pass's draws."""
    st = S.Stream()
    msaa = name.startswith("msaa") or name == "depth_only_msaa"
    layouts = {}
    verts, idx = uv_sphere(24, 12)
    tex = checker_texture(64, 8, seed)
    filt = {"filter_nearest": S.FILTER_NEAREST, "filter_linear": S.FILTER_LINEAR,
            "filter_nn": S.FILTER_NEAREST_MIPMAP_NEAREST, "filter_ln": S.FILTER_LINEAR_MIPMAP_NEAREST,
            "filter_nl": S.FILTER_NEAREST_MIPMAP_LINEAR}.get(name, S.FILTER_LINEAR_MIPMAP_LINEAR)
    wrap = {"wrap_mirror": S.WRAP_MIRRORED_REPEAT, "wrap_clamp_edge": S.WRAP_CLAMP_TO_EDGE,
            "wrap_border_white": S.WRAP_CLAMP_TO_BORDER}.get(name, S.WRAP_REPEAT)
    st.texture(1, 64, 64, mipmaps=filt > S.FILTER_LINEAR, filter_min=filt, filter_mag=S.FILTER_LINEAR, wrap=wrap,
               border=S.BORDER_WHITE if name == "wrap_border_white" else S.BORDER_BLACK,
               usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
    st.tex_data(1, [tex])
    if name.startswith("normal_"):
        # a tangent-space normal map (Blinn-Phong's normal-mapped path)
        _, nmap, _ = pbr_textures(64, seed + 1)
        st.texture(5, 64, 64, mipmaps=True, filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR,
                   wrap=S.WRAP_REPEAT, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD)
        st.tex_data(5, [nmap])
    if name == "layout_tiled":
        layouts[1] = 1
    if name == "layout_morton":
        layouts[1] = 2
    if name.startswith("wrap_"):
        # texture coordinates outside [0, 1]
        verts = verts.copy()
        verts[:, 4:6] = verts[:, 4:6] * 3.0 - 1.0
    st.vao(1, verts, idx)
    st.texture(2, width, height, usage=S.USAGE_COLOR, msaa=msaa)
    st.tex_init(2)
    st.texture(3, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH, msaa=msaa)
    st.tex_init(3)
    st.fbo(1, color=-1 if name == "depth_only_msaa" else 2, depth=3)
    eye = (0.0, 0.0, 3.0)
    rz = name == "depth_gequal_rz"
    fov = math.radians(60.0)
    center = (0.0, 0.0, 0.0)
    if name == "near_clip":
        eye = (0.0, 0.3, 1.02)
    if name == "msaa_near":
        eye = (0.0, 0.3, 1.048)  # fragments between n/2 and n get z < 0: dropped at 4x, clamped at 1x
    if name in ("side_clip", "msaa_clip"):
        eye, center = (0.0, 0.0, 2.2), (0.9, 0.5, 0.0)
    view = look_at(eye, center, (0, 1, 0))
    proj = projection(fov, width / height, 0.1, reverse_z=rz)
    model = np.eye(4)
    mvp = proj @ view @ model
    scene_u = S.uniforms_scene((0.2, 0.2, 0.2), eye, (1.0, 1.2, 1.5), (0.9, 0.9, 0.9))
    if name in ("shadow_blinn", "point_light", "normal_blinn"):
        scene_u = S.uniforms_scene((0.15, 0.15, 0.15), eye, (1.2, 1.6, 2.0), (1.0, 1.0, 1.0))
    shadow_mvp = None
    if name == "shadow_blinn":
        # a depth pass from the light first, sampled by the Blinn-Phong shadow lookup
        st.texture(4, 128, 128, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH | S.USAGE_SAMPLER,
                   wrap=S.WRAP_CLAMP_TO_BORDER)
        st.tex_init(4)
        st.fbo(2, depth=4)
        verts2, idx2 = uv_sphere(16, 8, 0.45)
        verts2[:, 0:3] += np.array([0.5, 0.9, 1.0], np.float32)
        st.vao(2, verts2, idx2)
    st.frame()
    clear_depth = 0.0 if rz else 1.0
    if name == "shadow_blinn":
        lview = look_at((1.2, 1.6, 2.0), (0, 0, 0), (0, 1, 0))
        lproj = projection(math.radians(70.0), 1.0, 0.1)
        bias = np.array([[0.5, 0, 0, 0.5], [0, 0.5, 0, 0.5], [0, 0, 1, 0], [0, 0, 0, 1]])
        shadow_mvp = bias @ lproj @ lview
        st.begin(2, False, True, (0, 0, 0, 0), 1.0)
        st.viewport(0, 0, 128, 128)
        rsd = _common_states(S.DEPTH_LESS, cull=True)
        for vid in (1, 2):
            st.draw(vid, S.SHADING_BASIC, 0, rsd, S.uniforms_model(model, lproj @ lview), bytes(S.SCENE_BYTES),
                    S.uniforms_material(0, 0, 0))
        st.end()
    st.begin(1, True, True, (0.1, 0.2, 0.3, 1.0), clear_depth)
    if name == "viewport_offset":
        st.viewport(20, 10, width - 40, height - 30)
    else:
        st.viewport(0, 0, width, height)
    rs = _common_states(S.DEPTH_LESS, cull=True)
    shading, defines = S.SHADING_BLINN_PHONG_DIR, S.DEF_ALBEDO
    samplers = [(S.SLOT_ALBEDO, 1)]
    mat = S.uniforms_material(1, 0, 0, 1.0, 0.5, (1, 1, 1, 1))
    second = None
    if name.startswith("blend") or name == "msaa_blend":
        rs = _common_states(S.DEPTH_LESS, cull=False, blend=True)
        if name == "blend_add_one":
            rs.set_blend_factor(S.BF_ONE, S.BF_ONE)
        if name == "blend_minmax":
            rs.blend_func_rgb, rs.blend_func_alpha = S.BLEND_MIN, S.BLEND_MAX
            rs.blend_src_rgb, rs.blend_dst_rgb = S.BF_SRC_COLOR, S.BF_ONE_MINUS_DST_COLOR
        if name == "blend_sub":
            rs.blend_func_rgb, rs.blend_func_alpha = S.BLEND_SUBTRACT, S.BLEND_REVERSE_SUBTRACT
            rs.blend_src_rgb, rs.blend_dst_rgb = S.BF_DST_ALPHA, S.BF_ONE_MINUS_DST_ALPHA
            rs.blend_src_alpha, rs.blend_dst_alpha = S.BF_SRC_ALPHA, S.BF_DST_COLOR
        mat = S.uniforms_material(1, 0, 0, 1.0, 0.5, (1, 1, 1, 0.6))
        defines = 0
        samplers = []
    elif name in ("depth_lequal", "depth_gequal_rz", "depth_always_nomask", "depth_equal", "depth_notequal",
                  "depth_never"):
        fn = {"depth_lequal": S.DEPTH_LEQUAL, "depth_gequal_rz": S.DEPTH_GEQUAL, "depth_always_nomask": S.DEPTH_ALWAYS,
              "depth_equal": S.DEPTH_EQUAL, "depth_notequal": S.DEPTH_NOTEQUAL, "depth_never": S.DEPTH_NEVER}[name]
        rs = _common_states(S.DEPTH_LESS, cull=False)
        # first draw with LESS (or GREATER for reverse Z), second with the function under test
        second = _common_states(fn, cull=False)
        if name == "depth_always_nomask":
            second.depth_mask = False
        if rz:
            rs.depth_func = S.DEPTH_GREATER
    elif name == "cull_off":
        rs = _common_states(S.DEPTH_LESS, cull=False)
    elif name in ("shadow_blinn", "point_light"):
        shading = S.SHADING_BLINN_PHONG
        mat = S.uniforms_material(1, 0, 1 if name == "shadow_blinn" else 0, 1.0, 0.8, (1, 1, 1, 1))
        if name == "shadow_blinn":
            samplers = [(S.SLOT_ALBEDO, 1), (S.SLOT_SHADOW, 4)]
    elif name in ("normal_blinn", "normal_blinn_dir"):
        shading = S.SHADING_BLINN_PHONG if name == "normal_blinn" else S.SHADING_BLINN_PHONG_DIR
        defines = S.DEF_ALBEDO | S.DEF_NORMAL
        samplers = [(S.SLOT_ALBEDO, 1), (S.SLOT_NORMAL, 5)]
        mat = S.uniforms_material(1, 0, 0, 1.0, 0.8, (1, 1, 1, 1))
    elif name == "basic_nodepth":
        shading, defines, samplers = S.SHADING_BASIC, 0, []
        rs = S.RenderStates()
        mat = S.uniforms_material(0, 0, 0, 1.0, 0.0, (0.9, 0.4, 0.2, 1.0))
    elif name == "ao_emissive":
        shading = S.SHADING_BLINN_PHONG
        defines = S.DEF_ALBEDO | S.DEF_AO | S.DEF_EMISSIVE
        samplers = [(S.SLOT_ALBEDO, 1), (S.SLOT_AO, 1), (S.SLOT_EMISSIVE, 1)]
    elif name == "depth_only_msaa":
        shading, defines, samplers = S.SHADING_BASIC, 0, []
    model_u = S.uniforms_model(model, mvp, shadow_mvp, reverse_z=rz)
    if name == "vertex_color":
        v, i = random_triangles(300, width, height, seed, 1.0, 200.0, 0.0, 1.0, big=2)
        st2 = st  # appended VAO must precede the frame: rebuild the stream with it
        st = S.Stream()
        st.texture(2, width, height, usage=S.USAGE_COLOR)
        st.tex_init(2)
        st.texture(3, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH)
        st.tex_init(3)
        st.fbo(1, color=2, depth=3)
        st.vao(3, v, i)
        st.frame()
        st.begin(1, True, True, (0, 0, 0, 0), 1.0)
        st.viewport(0, 0, width, height)
        st.draw(3, S.SHADING_VERTEX_COLOR, 0, _common_states(S.DEPTH_LESS, cull=False),
                S.uniforms_model(np.eye(4), np.eye(4)), bytes(S.SCENE_BYTES), S.uniforms_material(0, 0, 0))
        st.end()
        st.read(2)
        st.read(3)
        del st2
        return Scene(name, st, width, height, False, dict(layouts={}))
    st.draw(1, shading, defines, rs, model_u, scene_u, mat, samplers=samplers)
    if read_between and name != "depth_only_msaa":
        st.read(2, which=0)
        if msaa:
            st.read(2, which=1)
    if name.startswith("blend") or name == "msaa_blend":
        # a second, offset sphere blended over the first
        tm = _translate((0.35, 0.2, 0.4))
        st.draw(1, shading, defines, rs, S.uniforms_model(tm, proj @ view @ tm), scene_u,
                S.uniforms_material(1, 0, 0, 1.0, 0.5, (0.3, 0.9, 0.5, 0.5)), samplers=samplers)
    if second is not None:
        sm = np.diag([1.0, 1.0, 1.0, 1.0])
        sm[0, 3] = 0.15
        st.draw(1, S.SHADING_BASIC, 0, second, S.uniforms_model(sm, proj @ view @ sm, reverse_z=rz), scene_u,
                S.uniforms_material(0, 0, 0, 1.0, 0.0, (0.9, 0.2, 0.1, 1.0)))
    st.end()
    if name != "depth_only_msaa":
        st.read(2, which=0)
        if msaa:
            st.read(2, which=1)
    st.read(3, which=1 if msaa else 0)
    return Scene(name, st, width, height, msaa, dict(layouts=layouts))


def _colour_vertices(pos, col):
    """Vertices {vec3 a_position; vec4 a_color} (8 floats, the vertex colour shader's layout)."""
    v = np.zeros((len(pos), 8), np.float32)
    v[:, 0:3] = pos
    v[:, 4:8] = col
    return v


def scene_points_lines(width=160, height=120, seed=17, msaa=False, read_samples=False):
    """Points, lines and triangles in point and line mode (RendererSoft.cpp:575-718), in the ways the Viewer
    draws them (wireframe meshes, the world axis, the point light: Viewer.cpp:264-275, ModelLoader.cpp:60-85)
    and past them: over a shaded sphere with depth testing; lines and points crossing the near plane and the
    frustum sides (clipped, or dropped), wide lines and big points running off the framebuffer, points on
    repeated indices and on the shared vertices of a mesh in point mode (each rasterization starts where the
    last one of that vertex ended), a line within one pixel column, blended and vertex-coloured lines, a
    culled wireframe, and a depth-only pass that draws nothing."""
    rng = np.random.default_rng(seed)
    st = S.Stream()
    sv, si = uv_sphere(16, 8, 0.8)
    st.vao(1, sv, si)
    wv, wi = uv_sphere(12, 6, 0.95)
    st.vao(2, wv, wi)
    # world axis: three lines from the origin, and lines through the near plane and past the sides
    ax = np.array([[0, 0, 0], [1.5, 0, 0], [0, 0, 0], [0, 1.5, 0], [0, 0, 0], [0, 0, 1.5],
                   [-0.3, -0.2, 3.5], [0.4, 0.3, -1.0], [-4.0, 0.5, 0.2], [4.0, -0.6, 0.1],
                   [0.2, -3.0, 2.9], [0.1, 3.0, 2.95], [5.0, 5.0, 5.0], [6.0, 6.0, -5.0],
                   [0.71, -0.4, 0.2], [0.71, 0.6, 0.2]], np.float32)
    axc = rng.uniform(0, 1, (len(ax), 4)).astype(np.float32)
    axc[:, 3] = rng.uniform(0.3, 1.0, len(ax))
    st.vao(3, _colour_vertices(ax, axc), np.arange(len(ax), dtype=np.int32).reshape(-1, 2))
    # points: inside, at the framebuffer edge, behind the camera, outside the sides, and repeated indices
    pts = np.concatenate([rng.uniform(-1.2, 1.2, (24, 3)), [[0, 0, 4.0], [3.5, 0, 0], [-1.45, 1.05, 0.0],
                                                            [1.5, -1.1, 0.3]]]).astype(np.float32)
    ptc = rng.uniform(0, 1, (len(pts), 4)).astype(np.float32)
    pidx = np.concatenate([np.arange(len(pts)), [3, 3, 5, 3, 26, 26]]).astype(np.int32)
    st.vao(4, _colour_vertices(pts, ptc), pidx)
    st.texture(2, width, height, usage=S.USAGE_COLOR, msaa=msaa)
    st.tex_init(2)
    st.texture(3, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH, msaa=msaa)
    st.tex_init(3)
    st.fbo(1, color=2, depth=3)
    st.texture(5, width, height, format=S.FMT_FLOAT32, usage=S.USAGE_DEPTH)
    st.tex_init(5)
    st.fbo(2, depth=5)
    st.frame()
    eye = (0.6, 0.8, 3.0)
    view = look_at(eye, (0, 0, 0), (0, 1, 0))
    proj = projection(math.radians(60.0), width / height, 0.1)
    mvp = proj @ view
    scene_u = S.uniforms_scene((0.2, 0.2, 0.2), eye, (1.0, 1.2, 1.5), (0.9, 0.9, 0.9))
    st.begin(1, True, True, (0.05, 0.1, 0.15, 1.0), 1.0)
    st.viewport(0, 0, width, height)
    st.draw(1, S.SHADING_BLINN_PHONG, 0, _common_states(S.DEPTH_LESS, cull=True), S.uniforms_model(np.eye(4), mvp),
            scene_u, S.uniforms_material(1, 0, 0, 1.0, 0.5, (0.8, 0.7, 0.6, 1.0)))

    def basic(vao, prim, mode, colour, point_size=1.0, width_=1.0, depth_func=S.DEPTH_LEQUAL, cull=False,
              blend=False, depth_test=True, model=np.eye(4)):
        rs = _common_states(depth_func, cull=cull, blend=blend)
        rs.depth_test = depth_test
        rs.primitive_type, rs.polygon_mode, rs.line_width = prim, mode, width_
        st.draw(vao, S.SHADING_BASIC, 0, rs, S.uniforms_model(model, mvp @ model), bytes(S.SCENE_BYTES),
                S.uniforms_material(0, 0, 0, point_size, 0.5, colour))

    # wireframe of a larger sphere: back faces culled, then all edges on top without depth testing, thick
    basic(2, S.PRIM_TRIANGLE, S.POLY_LINE, (0.9, 0.9, 0.2, 1.0), cull=True)
    basic(2, S.PRIM_TRIANGLE, S.POLY_LINE, (0.2, 0.9, 0.9, 0.5), width_=2.0, blend=True, depth_test=False,
          model=_translate((0.9, -0.5, -0.6)))
    # the axis and the clipped lines, width 1 then 3
    basic(3, S.PRIM_LINE, S.POLY_FILL, (1.0, 0.3, 0.3, 1.0))
    basic(3, S.PRIM_LINE, S.POLY_FILL, (0.3, 1.0, 0.3, 1.0), width_=3.0, depth_func=S.DEPTH_LESS,
          model=_translate((0.12, 0.2, 0.0)))
    # vertex-coloured lines, blended
    rs = _common_states(S.DEPTH_LEQUAL, cull=False, blend=True)
    rs.primitive_type = S.PRIM_LINE
    tl = _translate((-0.25, -0.3, 0.2))
    st.draw(3, S.SHADING_VERTEX_COLOR, 0, rs, S.uniforms_model(tl, mvp @ tl), bytes(64), bytes(48))
    # points of size 10 (the point light), 1 and 4.5, with repeated indices
    basic(4, S.PRIM_POINT, S.POLY_FILL, (1.0, 1.0, 1.0, 1.0), point_size=10.0, depth_func=S.DEPTH_LESS)
    basic(4, S.PRIM_POINT, S.POLY_FILL, (0.0, 0.0, 1.0, 1.0), point_size=1.0, depth_test=False)
    basic(4, S.PRIM_POINT, S.POLY_FILL, (1.0, 0.5, 0.0, 0.6), point_size=4.5, blend=True, depth_test=False,
          model=_translate((0.15, 0.05, 0.0)))
    # a mesh in point mode: shared vertices are rasterized several times
    basic(1, S.PRIM_TRIANGLE, S.POLY_POINT, (0.6, 0.1, 0.9, 1.0), point_size=3.0, depth_test=False, cull=True,
          model=_translate((-0.7, 0.4, 0.0)))
    # vertex-coloured points
    rs = _common_states(S.DEPTH_ALWAYS, cull=False)
    rs.primitive_type = S.PRIM_POINT
    tp = _translate((0.0, -0.12, 0.05))
    st.draw(4, S.SHADING_VERTEX_COLOR, 0, rs, S.uniforms_model(tp, mvp @ tp), bytes(64), bytes(48))
    st.end()
    # a depth-only pass: points and lines draw nothing without a colour attachment
    st.begin(2, False, True, (0, 0, 0, 0), 1.0)
    st.viewport(0, 0, width, height)
    basic(3, S.PRIM_LINE, S.POLY_FILL, (1, 1, 1, 1), depth_func=S.DEPTH_ALWAYS)
    basic(2, S.PRIM_TRIANGLE, S.POLY_LINE, (1, 1, 1, 1), depth_func=S.DEPTH_ALWAYS)
    st.end()
    st.read(2)
    st.read(3, which=1 if msaa else 0)  # a multisampled depth buffer has no single-sample image
    if read_samples and msaa:
        st.read(2, which=1)
    st.read(5)
    return Scene("points_lines", st, width, height, msaa, {})


def scene_ibl_precompute(src_size=512, irr_size=32, pref_size=128, pref_levels=5, seed=7):
    """The C2 environment's IBL precompute alone: a seeded smooth radiance cube of `src_size`-px RGBA8 faces
    with a full mip chain (RendererSoft stores colour textures as RGBA8: Texture.h:63-66), convolved into a
    `irr_size`-px irradiance cube and GGX-prefiltered into a `pref_size`-px cube of `pref_levels` levels
    (Environment.cpp:48-103, 152-195). The frame is the 36 render passes; the maps are read back."""
    st = S.Stream()
    cv, ci = cube_mesh()
    st.vao(1, cv, ci)
    trilinear = dict(filter_min=S.FILTER_LINEAR_MIPMAP_LINEAR, filter_mag=S.FILTER_LINEAR, wrap=S.WRAP_CLAMP_TO_EDGE)
    st.texture(10, src_size, src_size, type=S.TEX_CUBE, mipmaps=True, usage=S.USAGE_SAMPLER | S.USAGE_UPLOAD,
               **trilinear)
    st.tex_data(10, env_cube_faces(src_size, seed))
    st.texture(11, irr_size, irr_size, type=S.TEX_CUBE, usage=S.USAGE_COLOR | S.USAGE_SAMPLER,
               filter_min=S.FILTER_LINEAR, filter_mag=S.FILTER_LINEAR, wrap=S.WRAP_CLAMP_TO_EDGE)
    st.tex_init(11)
    st.texture(12, pref_size, pref_size, type=S.TEX_CUBE, mipmaps=True, usage=S.USAGE_COLOR | S.USAGE_SAMPLER,
               **trilinear)
    st.tex_init(12)
    st.frame()
    ibl_precompute(st, 1, 10, 11, 12, src_size, irr_size, pref_size, pref_levels)
    for face in range(6):
        st.read(11, layer=face)
    for lev in range(pref_levels):
        for face in range(6):
            st.read(12, layer=face, level=lev)
    return Scene("ibl_precompute", st, pref_size, pref_size, False,
                 dict(src_size=src_size, irr_size=irr_size, pref_size=pref_size, pref_levels=pref_levels))
