"""Seeded synthetic scenes, recorded as rhip_stream.h command streams.

Each builder returns a `Scene` holding the stream plus metadata. The call
sequence mirrors what SoftGLRender's Viewer issues for a frame
(src/Viewer/Viewer.cpp:100-168 drawFrame/drawShadowMap, :388-396 pipelineDraw,
:597-626 setupPipelineStates, :684-716 uniform updates); the camera matrices
follow src/Viewer/Camera.cpp:26-69 (infinite-far projection, optional reverse Z).
"""
from __future__ import annotations

import math
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
