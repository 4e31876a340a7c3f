"""CPU restatement of the learned PBR shader's network -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's baseline leg use this; the
renderer never does. The learned shader has no counterpart in SoftGLRender, so
nothing of the reference is restated here: this module states the network the
GPU kernels compute (softglrender_amd/csrc/rhip_neural.h) in numpy, to check
their forward pass and gradients.

Network: 24 inputs (+ a constant 1 that carries layer 1's bias) -> 64 ReLU ->
64 ReLU -> 3 linear. Parameter vector layout (5955 floats): W1 (64x24), b1
(64), W2 (64x64), b2 (64), W3 (3x64), b3 (3). Loss of a batch: sum over
records and channels of (y - target)^2.

With `bf16=True` the operands are rounded to bfloat16 where the kernels round
them (inputs, weights, hidden activations, output gradients, hidden
gradients) and everything else is float64; with `bf16=False` it is plain
float64 math, which the finite-difference test pins.

Pinning (tests/test_nn_oracle.py): the float64 gradient equals central finite
differences of the float64 loss; the forward pass of a hand-set network
reduces to closed forms (a single active path); features of axis-aligned
vectors are checked against their definitions.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 24, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_OUT * N_HID
N_PARAMS = OFF_B3 + N_OUT

# capture record layout (32 floats): the network's inputs (N, V, albedo, roughness, metallic), AO, material,
# FP64 target, plain GPU colour, pixel, sample mask, then the plain shader's exact inputs (N, albedo, roughness,
# metallic; V is the same) that the FP64 target is computed from
REC_FLOATS = 32
REC_N, REC_V, REC_ALBEDO, REC_ROUGH, REC_METAL, REC_AO, REC_MATERIAL, REC_TARGET, REC_PLAIN = 0, 3, 6, 10, 11, 12, 13, \
    14, 17
REC_PIXEL, REC_MASK = 20, 21
REC_XN, REC_XALBEDO, REC_XROUGH, REC_XMETAL = 22, 25, 28, 29


def exact_from_inputs(rec):
    """Copies the network-input fields of records into their exact-input fields (records whose inputs are
    exact, as synthetic ones are)."""
    rec[:, REC_XN:REC_XN + 3] = rec[:, REC_N:REC_N + 3]
    rec[:, REC_XALBEDO:REC_XALBEDO + 3] = rec[:, REC_ALBEDO:REC_ALBEDO + 3]
    rec[:, REC_XROUGH], rec[:, REC_XMETAL] = rec[:, REC_ROUGH], rec[:, REC_METAL]
    return rec


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    w1 = p[OFF_W1:OFF_B1].reshape(N_HID, N_IN)
    b1 = p[OFF_B1:OFF_W2]
    w2 = p[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3].reshape(N_OUT, N_HID)
    b3 = p[OFF_B3:]
    return w1, b1, w2, b2, w3, b3


def pack(w1, b1, w2, b2, w3, b3):
    return np.concatenate([np.ravel(w1), b1, np.ravel(w2), b2, np.ravel(w3), b3]).astype(np.float64)


def _dot3_f32(a, b):
    """fma(z, z', fma(x, x', y * y')) in float32 (the shaders' dot product)."""
    a = a.astype(np.float32)
    b = b.astype(np.float32)
    yy = (a[..., 1] * b[..., 1]).astype(np.float32)
    t = (a[..., 0].astype(np.float64) * b[..., 0] + yy).astype(np.float32)
    return (a[..., 2].astype(np.float64) * b[..., 2] + t).astype(np.float32)


def features(rec):
    """The 24 network inputs of capture records (float32 like the kernels): N, V, max(N.V, 0), roughness,
    metallic, albedo (as sampled), R = 2 (N.V) N - V, sin(pi R), sin(2 pi R), sin(4 pi R) (the last two
    by double-angle identities from sin(pi R) and cos(pi R), as the kernels compute them)."""
    rec = np.asarray(rec, dtype=np.float32)
    N = rec[:, REC_N:REC_N + 3]
    V = rec[:, REC_V:REC_V + 3]
    nv = _dot3_f32(N, V)
    d2 = (np.float32(2.0) * nv).astype(np.float32)
    R = (d2[:, None] * N - V).astype(np.float32)
    f = np.zeros((rec.shape[0], N_IN), np.float32)
    f[:, 0:3] = N
    f[:, 3:6] = V
    f[:, 6] = np.maximum(nv, 0)
    f[:, 7] = rec[:, REC_ROUGH]
    f[:, 8] = rec[:, REC_METAL]
    f[:, 9:12] = rec[:, REC_ALBEDO:REC_ALBEDO + 3]
    f[:, 12:15] = R
    # sin(pi R), then sin(2 pi R) = 2 s c and sin(4 pi R) = 2 sin(2 pi R) cos(2 pi R), cos(2 pi R) = 1 - 2 s^2
    a = (np.float32(np.pi) * R).astype(np.float32)
    s1, c1 = np.sin(a).astype(np.float32), np.cos(a).astype(np.float32)
    s2 = (np.float32(2) * s1 * c1).astype(np.float32)
    c2 = (np.float32(1) - np.float32(2) * s1 * s1).astype(np.float32)
    f[:, 15:18], f[:, 18:21], f[:, 21:24] = s1, s2, (np.float32(2) * s2 * c2).astype(np.float32)
    return f


def forward(p, x, bf=True):
    """Returns (y, cache) for inputs x (n x 24)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    w1b = np.concatenate([w1, b1[:, None]], 1)  # bias as column 24
    xb = np.concatenate([np.asarray(x, np.float64), np.ones((x.shape[0], 1))], 1)
    X, W1, W2, W3 = r(xb), r(w1b), r(w2), r(w3)
    a1 = X @ W1.T
    h1 = r(np.maximum(a1, 0))
    f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf else (lambda a: np.asarray(a, np.float64))
    a2 = h1 @ W2.T + f32(b2)
    h2 = r(np.maximum(a2, 0))
    y = h2 @ W3.T + f32(b3)
    return y, (X, W1, W2, W3, h1, h2)


def gradient(p, x, target, bf=True):
    """Gradient of sum((y - target)^2) with respect to the parameter vector, and the loss."""
    y, (X, W1, W2, W3, h1, h2) = forward(p, x, bf)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    e = y - target
    loss = float(np.sum(e * e))
    dy = r(2 * e)
    dW3 = dy.T @ h2
    db3 = dy.sum(0)
    dh2 = r(np.where(h2 > 0, dy @ W3, 0))
    dW2 = dh2.T @ h1
    db2 = dh2.sum(0)
    dh1 = r(np.where(h1 > 0, dh2 @ W2, 0))
    dW1b = dh1.T @ X
    return pack(dW1b[:, :N_IN], dW1b[:, N_IN], dW2, db2, dW3, db3), loss


# ----------------------------------------------------------------------------- FP64 plain shader

def _cube_uv(d):
    """convertXYZ2UV (SamplerSoft.h:311-373) in float64: face (later faces win ties), u, v."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    ax, ay, az = np.abs(x), np.abs(y), np.abs(z)
    face = np.zeros(len(d), np.int64)
    ma, uc, vc = np.zeros(len(d)), np.zeros(len(d)), np.zeros(len(d))
    cases = [(x > 0) & (ax >= ay) & (ax >= az), ~(x > 0) & (ax >= ay) & (ax >= az),
             (y > 0) & (ay >= ax) & (ay >= az), ~(y > 0) & (ay >= ax) & (ay >= az),
             (z > 0) & (az >= ax) & (az >= ay), ~(z > 0) & (az >= ax) & (az >= ay)]
    vals = [(ax, -z, y), (ax, z, y), (ay, x, -z), (ay, x, z), (az, x, y), (az, -x, y)]
    for i, (c, (m_, u_, v_)) in enumerate(zip(cases, vals)):
        face = np.where(c, i, face)
        ma, uc, vc = np.where(c, m_, ma), np.where(c, u_, uc), np.where(c, v_, vc)
    return face, 0.5 * (uc / ma + 1.0), 0.5 * (-vc / ma + 1.0)


def _bilinear_clamp(levels_faces, level, face, u, v):
    """Bilinear RGB (float64, exact weights) of `face` at `level`, clamp-to-edge."""
    out = np.zeros((len(u), 3))
    for f in range(6):
        sel = face == f
        if not np.any(sel):
            continue
        img = levels_faces[level][f][..., :3].astype(np.float64) / 255.0
        h, w = img.shape[:2]
        fx, fy = u[sel] * w - 0.5, v[sel] * h - 0.5
        x0, y0 = np.floor(fx), np.floor(fy)
        ax, ay = (fx - x0)[:, None], (fy - y0)[:, None]
        xi, yi = x0.astype(np.int64), y0.astype(np.int64)
        cx = lambda t: np.clip(t, 0, w - 1)
        cy = lambda t: np.clip(t, 0, h - 1)
        t0 = img[cy(yi), cx(xi)] * (1 - ax) + img[cy(yi), cx(xi + 1)] * ax
        t1 = img[cy(yi + 1), cx(xi)] * (1 - ax) + img[cy(yi + 1), cx(xi + 1)] * ax
        out[sel] = t0 * (1 - ay) + t1 * ay
    return out


def pbr_ibl_f64(rec, irradiance_faces, prefilter_levels):
    """The image-based PBR shader (PbrSoft.h:225-322 with the point light off) in float64 for capture
    records, from the RGBA8 irradiance cube (6 faces) and prefiltered mip chain (list of 6-face levels):
    display colour before clamping. Cube sampling: bilinear in a face with clamp-to-edge, linear between
    mip levels floor(lod) and floor(lod) + 1 (lod = 4 roughness), no 8-bit rounding in between."""
    rec = np.asarray(rec, np.float64)
    N, V = rec[:, REC_XN:REC_XN + 3], rec[:, REC_V:REC_V + 3]
    rough, metal, ao = rec[:, REC_XROUGH], rec[:, REC_XMETAL], rec[:, REC_AO]
    alb = rec[:, REC_XALBEDO:REC_XALBEDO + 3] ** 2.2
    F0 = 0.04 * (1 - metal[:, None]) + alb * metal[:, None]
    ndv = np.sum(N * V, 1)
    nv = np.maximum(ndv, 0)
    p5 = np.clip(1 - nv, 0, 1) ** 5
    R = 2 * ndv[:, None] * N - V
    f, u, v = _cube_uv(N)
    irr = _bilinear_clamp([irradiance_faces], 0, f, u, v)
    f, u, v = _cube_uv(R)
    lod = rough * 4.0
    max_level = len(prefilter_levels) - 1
    hi = np.clip(np.floor(lod).astype(np.int64), 0, max_level)
    lo = np.minimum(hi + 1, max_level)
    frac = (lod - np.floor(lod))[:, None]
    pf = np.zeros((len(rec), 3))
    for lev in np.unique(hi):
        s = hi == lev
        a = _bilinear_clamp(prefilter_levels, lev, f[s], u[s], v[s])
        b = _bilinear_clamp(prefilter_levels, min(lev + 1, max_level), f[s], u[s], v[s])
        pf[s] = np.where((lo[s] == hi[s])[:, None], a, a * (1 - frac[s]) + b * frac[s])
    F = F0 + (np.maximum(1 - rough[:, None], F0) - F0) * p5[:, None]
    cx, cy = 1 - rough, 0.0425 - 0.0275 * rough
    cz, cw = 1.04 - 0.572 * rough, -0.04 + 0.022 * rough
    a004 = np.minimum(cx * cx, np.exp2(-9.28 * nv)) * cx + cy
    bx = -1.04 * a004 + cz
    by = (1.04 * a004 + cw) * np.clip(50.0 * F[:, 1], 0, 1)
    kD = (1 - F) * (1 - metal[:, None])
    amb = (kD * irr * alb + pf * (F * bx[:, None] + by[:, None])) * ao[:, None]
    return amb ** (1.0 / 2.2)
