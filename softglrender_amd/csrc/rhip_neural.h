/*
 * rhip_neural.h -- the learned PBR shader (no counterpart in SoftGLRender).
 *
 * A 24-64-64-3 MLP (ReLU, linear RGB output) stands in for the image-based
 * PBR shader (PbrSoft.h:225-322 with the point light off): from the normal,
 * view direction, N.V, roughness, metallic, albedo and a 3-band encoding of
 * the reflection vector it predicts the display colour, replacing the two
 * cube-map lookups, the split-sum terms and the gamma curves. Weights are
 * BF16 on MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, for
 * inference and for training; an FP32 master copy is updated with Adam. The
 * plain shader evaluated in FP64 is the reference both for the training
 * targets and for the error the network is judged by.
 *
 * Matrix convention: every MFMA operand is loaded as 8 consecutive bf16 of a
 * row-major [row][k] matrix in LDS (lane l: row l%16, k (l/16)*8..+7). A
 * layer computes H^T[unit][frag] = W[unit][k] . X[frag][k], so weights are
 * the A operand and activations the B operand; the result tile gives lane l
 * units (l/16)*4+v of fragment l%16.
 */
#pragma once
#include "rhip_kernels.h"

namespace rhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kNnIn = 24, kNnHid = 64, kNnOut = 3;
constexpr int kNnK1 = 32; /* layer-1 K: 24 features, a constant 1 (bias column), zeros */
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kNnHid * kNnIn, kOffW2 = kOffB1 + kNnHid,
              kOffB2 = kOffW2 + kNnHid * kNnHid, kOffW3 = kOffB2 + kNnHid, kOffB3 = kOffW3 + kNnOut * kNnHid;
constexpr int kNnParams = kOffB3 + kNnOut; /* 5955 */
constexpr int kNnMaxParts = 1024; /* most k_nn_grad blocks in one launch */
constexpr int kNnUpdateBlocks = (kNnParams + 15) / 16; /* k_nn_update: 16 parameters per block */

/* bf16 operands packed for MFMA (device): W1 with b1 in column 24, W2, W3
 * padded to 16 rows, and the transposes backward needs; biases of layers 2, 3.
 * w2r and w3r are W2 and W3 with their columns in the order nnRegUnit gives,
 * for the inference kernel that keeps activations in registers. */
struct NnPacked {
  __bf16 w1[kNnHid * kNnK1];
  __bf16 w2[kNnHid * kNnHid];
  __bf16 w3[16 * kNnHid];
  __bf16 w2t[kNnHid * kNnHid];
  __bf16 w3t[kNnHid * 32];
  __bf16 w2r[kNnHid * kNnHid];
  __bf16 w3r[16 * kNnHid];
  float b2[kNnHid];
  float b3[16];
};

/* The hidden unit behind K position `slot` (0..63) of a layer whose B operand
 * is built in registers from the previous layer's result tiles. Lane (g, c)
 * holds units 16m + 4g + v (v < 4) of result tile m; the K block kb of the next
 * layer takes tiles 2kb and 2kb+1, so its K positions 8g .. 8g+7 are those two
 * tiles' four units each. The weights' columns are stored in the same order. */
__host__ __device__ constexpr int nnRegUnit(int slot) {
  return 32 * (slot >> 5) + 4 * ((slot & 31) >> 3) + ((slot & 7) < 4 ? (slot & 3) : 16 + (slot & 3));
}
/* its inverse: the K position of hidden unit u */
__host__ __device__ constexpr int nnRegSlot(int u) {
  return 32 * (u >> 5) + 8 * ((u & 15) >> 2) + (u & 3) + ((u & 16) ? 4 : 0);
}
constexpr bool nnRegOrderOk() {
  for (int s = 0; s < 64; s++)
    if (nnRegSlot(nnRegUnit(s)) != s) return false;
  return true;
}
static_assert(nnRegOrderOk(), "nnRegSlot must invert nnRegUnit");

/* A captured fragment: what the network reads (the learned shader's inputs,
 * nnFastInputs), the plain shader's exact inputs, the plain GPU shader's colour
 * and the FP64 plain-shader colour (from the exact inputs). */
constexpr int kRecFloats = 32;
enum {
  kRecN = 0,        /* 3: shading normal (network input) */
  kRecV = 3,        /* 3: view direction (network input and exact) */
  kRecAlbedo = 6,   /* 4: albedo texel (sRGB-encoded), alpha (network input) */
  kRecRough = 10,   /* network input */
  kRecMetal = 11,   /* network input */
  kRecAo = 12,
  kRecMaterial = 13,
  kRecTarget = 14,  /* 3: FP64 plain shader on the exact inputs, clamped to [0, 1] */
  kRecPlain = 17,   /* 3: the GPU plain shader, clamped to [0, 1] */
  kRecPixel = 20,   /* pixel index y * width + x */
  kRecMask = 21,    /* samples the fragment covers (bit mask) */
  kRecXN = 22,      /* 3: the plain shader's shading normal */
  kRecXAlbedo = 25, /* 3: the plain shader's albedo texel */
  kRecXRough = 28,
  kRecXMetal = 29,
};

__device__ __forceinline__ f32x4 mma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* operand of tile (row0, k0) of a row-major bf16 matrix with `rs` elements per row */
__device__ __forceinline__ bf16x8 ldOp(const __bf16 *m, int rs, int row0, int k0) {
  const int l = threadIdx.x & 63;
  return *(const bf16x8 *)(m + (row0 + (l & 15)) * rs + k0 + (l >> 4) * 8);
}

/* The 24 inputs (+ the bias 1) of one fragment. */
__device__ __forceinline__ void nnFeatures(f3 N, f3 V, float rough, float metal, float ar, float ag, float ab,
                                           float *f) {
  const float nv = gmax(dot3(N, V), 0.f);
  const float d2 = 2.f * dot3(N, V);
  const f3 R{d2 * N.x - V.x, d2 * N.y - V.y, d2 * N.z - V.z};
  f[0] = N.x, f[1] = N.y, f[2] = N.z;
  f[3] = V.x, f[4] = V.y, f[5] = V.z;
  f[6] = nv, f[7] = rough, f[8] = metal;
  f[9] = ar, f[10] = ag, f[11] = ab;
  f[12] = R.x, f[13] = R.y, f[14] = R.z;
  /* sin(pi R), then sin(2 pi R) and sin(4 pi R) by the double-angle identities */
  const float pi = 3.14159265358979f;
  const float r3[3] = {R.x, R.y, R.z};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float s1 = __sinf(pi * r3[i]), c1 = __cosf(pi * r3[i]);
    const float s2 = 2.f * s1 * c1, c2 = 1.f - 2.f * s1 * s1;
    f[15 + i] = s1;
    f[18 + i] = s2;
    f[21 + i] = 2.f * s2 * c2;
  }
  f[24] = 1.f;
#pragma unroll
  for (int k = 25; k < kNnK1; k++) f[k] = 0.f;
}

__device__ __forceinline__ void recFeatures(const float *r, float *f) {
  nnFeatures({r[kRecN], r[kRecN + 1], r[kRecN + 2]}, {r[kRecV], r[kRecV + 1], r[kRecV + 2]}, r[kRecRough],
             r[kRecMetal], r[kRecAlbedo], r[kRecAlbedo + 1], r[kRecAlbedo + 2], f);
}

/* Inference LDS. Rows are padded by 8 bf16 so that the 16 rows one MFMA operand
 * load touches start in different banks. */
constexpr int kXs = kNnK1 + 8, kHs = kNnHid + 8;
struct NnWaveLds {
  __bf16 x[64 * kXs];
  float y[64 * 4];
};
/* the block's copy of the packed operands (read by all its wavefronts) */
struct NnBlockW {
  __bf16 w1[kNnHid * kXs];
  __bf16 w2r[kNnHid * kHs];
  __bf16 w3r[16 * kHs];
  float b2[kNnHid];
  float b3[16];
};

/* Copies the packed operands into the block's LDS (all threads of the block). */
__device__ __forceinline__ void nnLoadBlockW(const NnPacked &W, NnBlockW &B) {
  for (int i = threadIdx.x; i < kNnHid * kNnK1 / 8; i += blockDim.x) {
    const int r = i / (kNnK1 / 8), k = (i % (kNnK1 / 8)) * 8;
    *(bf16x8 *)&B.w1[r * kXs + k] = *(const bf16x8 *)&W.w1[r * kNnK1 + k];
  }
  for (int i = threadIdx.x; i < kNnHid * kNnHid / 8; i += blockDim.x) {
    const int r = i / (kNnHid / 8), k = (i % (kNnHid / 8)) * 8;
    *(bf16x8 *)&B.w2r[r * kHs + k] = *(const bf16x8 *)&W.w2r[r * kNnHid + k];
  }
  for (int i = threadIdx.x; i < 16 * kNnHid / 8; i += blockDim.x) {
    const int r = i / (kNnHid / 8), k = (i % (kNnHid / 8)) * 8;
    *(bf16x8 *)&B.w3r[r * kHs + k] = *(const bf16x8 *)&W.w3r[r * kNnHid + k];
  }
  for (int i = threadIdx.x; i < kNnHid; i += blockDim.x) B.b2[i] = W.b2[i];
  for (int i = threadIdx.x; i < 16; i += blockDim.x) B.b3[i] = W.b3[i];
}

/* Two result tiles (units 16m + 4g + v of one fragment group), ReLU'd, as the
 * next layer's B operand in nnRegUnit order. */
__device__ __forceinline__ bf16x8 reluPack(f32x4 lo, f32x4 hi) {
  bf16x8 o;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    o[v] = (__bf16)fmaxf(lo[v], 0.f);
    o[4 + v] = (__bf16)fmaxf(hi[v], 0.f);
  }
  return o;
}

/* The network on the 64 fragments whose input rows are in `X` (row stride kXs):
 * layer-1 operands from the rows, then the hidden activations in registers
 * (each layer's result tiles are the next layer's B operand, with the weights'
 * columns in nnRegUnit order). Returns, in lanes 0..15, channels 0..3 of
 * fragments 16q + c. `w1`, `w2r`, `w3r` have row strides s1, s2, s2; the same
 * arithmetic in the same order wherever they live. Every lane must call it. */
__device__ __forceinline__ void nnLayers(const __bf16 *w1, const __bf16 *w2r, const __bf16 *w3r, int s1, int s2,
                                         const float *b2, const float *b3, const __bf16 *X, f32x4 *yq) {
  const int g = (threadIdx.x & 63) >> 4;
  bf16x8 xb[4];
#pragma unroll
  for (int q = 0; q < 4; q++) xb[q] = ldOp(X, kXs, 16 * q, 0);
  /* layer 1: H1^T = W1 . X^T (bias in column 24), packed as layer 2's operands */
  bf16x8 h1[4][2];
#pragma unroll
  for (int kb = 0; kb < 2; kb++) {
    const bf16x8 a0 = ldOp(w1, s1, 32 * kb, 0), a1 = ldOp(w1, s1, 32 * kb + 16, 0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const f32x4 z = {0.f, 0.f, 0.f, 0.f};
      h1[q][kb] = reluPack(mma16(a0, xb[q], z), mma16(a1, xb[q], z));
    }
  }
  /* layer 2: H2^T = W2 . H1^T + b2, packed as layer 3's operands */
  bf16x8 h2[4][2];
#pragma unroll
  for (int kb = 0; kb < 2; kb++) {
    f32x4 d[2][4];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int m = 2 * kb + t;
      const bf16x8 a0 = ldOp(w2r, s2, 16 * m, 0), a1 = ldOp(w2r, s2, 16 * m, 32);
      const f32x4 bias = *(const f32x4 *)&b2[16 * m + 4 * g];
#pragma unroll
      for (int q = 0; q < 4; q++) d[t][q] = mma16(a1, h1[q][1], mma16(a0, h1[q][0], bias));
    }
#pragma unroll
    for (int q = 0; q < 4; q++) h2[q][kb] = reluPack(d[0][q], d[1][q]);
  }
  /* layer 3: Y^T = W3 . H2^T + b3 */
  const bf16x8 a0 = ldOp(w3r, s2, 0, 0), a1 = ldOp(w3r, s2, 0, 32);
  const f32x4 bias = *(const f32x4 *)&b3[4 * g];
#pragma unroll
  for (int q = 0; q < 4; q++) yq[q] = mma16(a1, h2[q][1], mma16(a0, h2[q][0], bias));
}

/* Forward pass of the 64 fragments of a wavefront: lane l's features `f`
 * (kNnK1 floats) in, lane l's RGB out. Every lane of the wavefront must call it. */
__device__ __forceinline__ void nnForward64(const NnBlockW &W, NnWaveLds &L, const float *f, float *out) {
  const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
#pragma unroll
  for (int k0 = 0; k0 < kNnK1; k0 += 8) {
    bf16x8 v;
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = (__bf16)f[k0 + k];
    *(bf16x8 *)&L.x[l * kXs + k0] = v;
  }
  waveSync();
  f32x4 yq[4];
  nnLayers(W.w1, W.w2r, W.w3r, kXs, kHs, W.b2, W.b3, L.x, yq);
  if (g == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) *(f32x4 *)&L.y[(16 * q + c) * 4] = yq[q];
  }
  waveSync();
  const f32x4 y = *(const f32x4 *)&L.y[l * 4];
  out[0] = y[0];
  out[1] = y[1];
  out[2] = y[2];
  waveSync();
}

/* --------------------------------------------------------------------------- FP64 plain shader */

__device__ __forceinline__ double texelF64(const TexDesc &d, const TexLevel &l, int x, int y, int ch) {
  return (double)chan(fetchWrap<uint32_t>(d, l, x, y, d.wrap), ch) / 255.0;
}

__device__ __forceinline__ void bilinearF64(const TexDesc &d, int layer, int level, double u, double v, double *o) {
  const TexLevel &l = d.lv[layer][level];
  const double fx0 = u * l.w - 0.5, fy0 = v * l.h - 0.5;
  const double flx = floor(fx0), fly = floor(fy0);
  const int x = (int)flx, y = (int)fly;
  const double ax = fx0 - flx, ay = fy0 - fly;
#pragma unroll
  for (int ch = 0; ch < 3; ch++) {
    const double t0 = texelF64(d, l, x, y, ch) * (1 - ax) + texelF64(d, l, x + 1, y, ch) * ax;
    const double t1 = texelF64(d, l, x, y + 1, ch) * (1 - ax) + texelF64(d, l, x + 1, y + 1, ch) * ax;
    o[ch] = t0 * (1 - ay) + t1 * ay;
  }
}

/* textureCubeLodImpl (SamplerSoft.h:289-373) with exact interpolation weights and no
 * intermediate 8-bit rounding: bilinear in a face, linear between mip levels. */
__device__ void cubeF64(const TexDesc &d, double x, double y, double z, double lod, double *o) {
  const double ax = fabs(x), ay = fabs(y), az = fabs(z);
  double ma = 0, uc = 0, vc = 0;
  int face = 0;
  if (x > 0 && ax >= ay && ax >= az) { ma = ax; uc = -z; vc = y; face = 0; }
  if (!(x > 0) && ax >= ay && ax >= az) { ma = ax; uc = z; vc = y; face = 1; }
  if (y > 0 && ay >= ax && ay >= az) { ma = ay; uc = x; vc = -z; face = 2; }
  if (!(y > 0) && ay >= ax && ay >= az) { ma = ay; uc = x; vc = z; face = 3; }
  if (z > 0 && az >= ax && az >= ay) { ma = az; uc = x; vc = y; face = 4; }
  if (!(z > 0) && az >= ax && az >= ay) { ma = az; uc = -x; vc = y; face = 5; }
  vc = -vc;
  const double u = 0.5 * (uc / ma + 1.0), v = 0.5 * (vc / ma + 1.0);
  if (d.filter_min <= 1 || d.levels == 1) {
    bilinearF64(d, face, 0, u, v, o);
    return;
  }
  const int maxLevel = d.levels - 1;
  const int hi = min(max((int)floor(lod), 0), maxLevel);
  const int lo = min(hi + 1, maxLevel);
  bilinearF64(d, face, hi, u, v, o);
  if (hi == lo) return;
  double t[3];
  bilinearF64(d, face, lo, u, v, t);
  const double f = lod - floor(lod);
#pragma unroll
  for (int ch = 0; ch < 3; ch++) o[ch] = o[ch] * (1 - f) + t[ch] * f;
}

/* The image-based PBR shader (PbrSoft.h:225-322, u_enableLight off) in FP64
 * from the fragment's exact shading inputs; returns the display colour (not clamped). */
__device__ void pbrIblF64(const DrawParams &P, const float *r, double *col) {
  const double N[3] = {r[kRecXN], r[kRecXN + 1], r[kRecXN + 2]};
  const double V[3] = {r[kRecV], r[kRecV + 1], r[kRecV + 2]};
  const double rough = r[kRecXRough], metal = r[kRecXMetal], ao = r[kRecAo];
  double alb[3], F0[3];
#pragma unroll
  for (int ch = 0; ch < 3; ch++) {
    alb[ch] = pow((double)r[kRecXAlbedo + ch], 2.2);
    F0[ch] = 0.04 * (1 - metal) + alb[ch] * metal;
  }
  const double ndv = N[0] * V[0] + N[1] * V[1] + N[2] * V[2];
  const double nv = ndv > 0 ? ndv : 0;
  const double p5 = pow(fmin(fmax(1 - nv, 0.0), 1.0), 5.0);
  const double R[3] = {2 * ndv * N[0] - V[0], 2 * ndv * N[1] - V[1], 2 * ndv * N[2] - V[2]};
  double irr[3], pf[3];
  cubeF64(*P.tex[RHS_SLOT_IRRADIANCE], N[0], N[1], N[2], 0.0, irr);
  cubeF64(*P.tex[RHS_SLOT_PREFILTER], R[0], R[1], R[2], rough * 4.0, pf);
  /* Fresnel with roughness, then EnvBRDFApprox (PbrSoft.h:209-223) on it */
  double F[3];
#pragma unroll
  for (int ch = 0; ch < 3; ch++) F[ch] = F0[ch] + (fmax(1 - rough, F0[ch]) - F0[ch]) * p5;
  const double cx = 1 - rough, cy = 0.0425 - 0.0275 * rough, cz = 1.04 - 0.572 * rough, cw = -0.04 + 0.022 * rough;
  const double a004 = fmin(cx * cx, exp2(-9.28 * nv)) * cx + cy;
  const double bx = -1.04 * a004 + cz, by = (1.04 * a004 + cw) * fmax(0.0, fmin(1.0, 50.0 * F[1]));
#pragma unroll
  for (int ch = 0; ch < 3; ch++) {
    const double kD = (1 - F[ch]) * (1 - metal);
    const double amb = (kD * irr[ch] * alb[ch] + pf[ch] * (F[ch] * bx + by)) * ao;
    col[ch] = pow(amb, 1.0 / 2.2);
  }
}

/* --------------------------------------------------------------------------- inputs */

/* The learned shader's inputs: the plain shader's (pbrInputs) read at lower
 * cost. The albedo and normal maps are sampled bilinearly at the one mip level
 * nearest the LOD (the plain shader blends two levels), the LOD taken from the
 * triangle's barycentric gradients at the fragment (the plain shader differences
 * the exact shading samples of its quad's pixels), and the metal-rough map at
 * the nearest texel of level 0 (the plain shader: bilinear). The capture
 * records these for training, and the network is judged against the FP64
 * plain shader evaluated on the exact inputs, which it records beside them. */
struct NnIn {
  f3 N, V;
  float rough, metal;
  RGBAf albedo;
};
/* Texture-coordinate change over one pixel in x (dx) and in y (dy) at a
 * fragment: the perspective-correct coordinate at the fragment's raw
 * barycentric `b` and at b moved by the barycentric gradients of one pixel. */
struct NnUvGrad {
  f2 dx, dy;
};
__device__ __forceinline__ NnUvGrad nnUvGrad(const DrawParams &P, const PrimSetup &ps, const Tri &T, float4 b) {
  if (!(P.vary_groups & 1u)) return {{0.f, 0.f}, {0.f, 0.f}}; /* no texture coordinates stored */
  /* texture coordinates: slots 0-1, the first stored group */
  const float2 t0 = *(const float2 *)(P.vary + (size_t)ps.v[0] * P.vary_stride);
  const float2 t1 = *(const float2 *)(P.vary + (size_t)ps.v[1] * P.vary_stride);
  const float2 t2 = *(const float2 *)(P.vary + (size_t)ps.v[2] * P.vary_stride);
  /* bary(): ux = (a1 b2 - a2 b1) / uz, uy = (a2 b0 - a0 b2) / uz with a2 = v0x - x, b2 = v0y - y */
  const float iuz = 1.f / T.uz;
  const float uxx = T.b1 * iuz, uyx = -T.b0 * iuz, uxy = -T.a1 * iuz, uyy = T.a0 * iuz;
  const float W0 = T.w[0] * b.x, W1 = T.w[1] * b.y, W2 = T.w[2] * b.z;
  const float D = W0 + W1 + W2;
  const float Nu = W0 * t0.x + W1 * t1.x + W2 * t2.x, Nv = W0 * t0.y + W1 * t1.y + W2 * t2.y;
  /* weights of the moved barycentric (1 - ux - uy, uy, ux) minus the fragment's */
  auto step = [&](float dux, float duy) {
    const float e0 = -(dux + duy) * T.w[0], e1 = duy * T.w[1], e2 = dux * T.w[2];
    const float iD = 1.f / (D + (e0 + e1 + e2)), iD0 = 1.f / D;
    return f2{(Nu + e0 * t0.x + e1 * t1.x + e2 * t2.x) * iD - Nu * iD0,
              (Nv + e0 * t0.y + e1 * t1.y + e2 * t2.y) * iD - Nv * iD0};
  };
  return {step(uxx, uyx), step(uxy, uyy)};
}

/* the mip level nearest the LOD of `g` for texture t (hardware log2: the level only needs the rounding) */
__device__ __forceinline__ int nnLevel(const TexDesc &t, const NnUvGrad &g) {
  if (t.filter_min < 2) return 0;
  const float w = (float)t.width, h = (float)t.height;
  const float dxx = g.dx.x * w, dyx = g.dy.x * w;
  const float dxy = g.dx.y * h, dyy = g.dy.y * h;
  const float d = fmaxf(fmaf(dxy, dxy, dxx * dxx), fmaf(dyy, dyy, dyx * dyx));
  const float lod = 0.5f * __builtin_amdgcn_logf(d);
  /* clamped as a float first: a zero footprint gives -inf, a NaN takes level 0 */
  return (int)fminf(fmaxf(floorf(lod + 0.5f), 0.f), (float)(t.levels - 1));
}
__device__ __forceinline__ NnIn nnFastInputs(const DrawParams &P, const Frag &f, const NnUvGrad &g) {
  const float *v = f.v;
  const Uniforms &u = P.u;
  NnIn in;
  const f2 uv{v[0], v[1]};
  const TexDesc *ta = P.tex[RHS_SLOT_ALBEDO], *tn = P.tex[RHS_SLOT_NORMAL];
  const bool am = (P.defines & RHS_DEF_ALBEDO_MAP) && ta, nm = (P.defines & RHS_DEF_NORMAL_MAP) && tn;
  RGBAf nt{0.5f, 0.5f, 1.f, 1.f};
  if (am && nm && sameShape(*ta, *tn)) {
    float4 ba, bn;
    sampleLevelPair(*ta, *tn, 0, nnLevel(*ta, g), uv, true, ba, bn);
    in.albedo = unpack01(ba);
    nt = unpack01(bn);
  } else {
    in.albedo = am ? unpack01(sampleLevel4(*ta, 0, nnLevel(*ta, g), uv, true))
                   : RGBAf{u.base_color[0], u.base_color[1], u.base_color[2], u.base_color[3]};
    if (nm) nt = unpack01(sampleLevel4(*tn, 0, nnLevel(*tn, g), uv, true));
  }
  in.N = nm ? normalFromTexel<true>(f, 20, 24, nt) : nrm3b({v[4], v[5], v[6]});
  in.V = nrm3b({v[12], v[13], v[14]});
  in.metal = 0.f;
  in.rough = 1.f;
  const TexDesc *tm = P.tex[RHS_SLOT_METAL_ROUGHNESS];
  if ((P.defines & RHS_DEF_METALROUGHNESS_MAP) && tm) {
    const RGBAf mr = unpack01(sampleLevel4(*tm, 0, 0, uv, false));
    in.metal = mr.b;
    in.rough = mr.g;
  }
  return in;
}

/* --------------------------------------------------------------------------- capture */

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

/* Records 1 in 2^shift of the fragments the two-phase raster queued for a
 * PBR draw: the shading inputs, the plain GPU colour and the FP64 colour. */
template <int S>
__global__ __launch_bounds__(256) void k_capture(DrawParams P, int need_deriv, float *rec, long long cap,
                                                 unsigned long long *count, int shift, uint32_t salt, int material) {
  const int n = P.dcount[2];
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    if (hash32((uint32_t)j ^ salt) & ((1u << shift) - 1u)) continue;
    const uint2 jb = P.jobs[j];
    const uint32_t pix = jb.x & 0x0fffffffu, w = jb.y;
    const int px = (int)(pix % (uint32_t)P.fb_w), py = (int)(pix / (uint32_t)P.fb_w);
    const int id = (int)(w >> 2), inst = (int)(w & 3u);
    const PrimSetup ps = P.prims[id];
    const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
    const float fx = (float)px, fy = (float)py;
    float4 bs;
    if (S == 4) {
      if (!bary(T, fx + 0.5f, fy + 0.5f, bs)) {
#pragma unroll
        for (int t = 3; t >= 0; t--) {
          float4 b;
          if (bary(T, sampleX(t) + fx, sampleY(t) + fy, b)) bs = b;
        }
      }
    } else {
      bary(T, fx + 0.5f, fy + 0.5f, bs);
    }
    const uint32_t smask = jb.x >> 28;
    const Frag fr = makeFrag<S, RHS_SHADING_PBR>(P, ps, T, px, py, bs, inst, need_deriv);
    const PbrInputs in = pbrInputs(P, fr);
    const NnIn fi = nnFastInputs(P, fr, nnUvGrad(P, ps, T, bs));
    const RGBAf plain = fsPbr(P, fr);
    const unsigned long long k = atomicAdd(count, 1ull);
    if ((long long)k >= cap) continue;
    float r[kRecFloats];
    r[kRecN] = fi.N.x, r[kRecN + 1] = fi.N.y, r[kRecN + 2] = fi.N.z;
    r[kRecV] = fi.V.x, r[kRecV + 1] = fi.V.y, r[kRecV + 2] = fi.V.z;
    r[kRecAlbedo] = fi.albedo.r, r[kRecAlbedo + 1] = fi.albedo.g;
    r[kRecAlbedo + 2] = fi.albedo.b, r[kRecAlbedo + 3] = fi.albedo.a;
    r[kRecRough] = fi.rough, r[kRecMetal] = fi.metal, r[kRecAo] = in.ao;
    r[kRecMaterial] = (float)material;
    r[kRecXN] = in.N.x, r[kRecXN + 1] = in.N.y, r[kRecXN + 2] = in.N.z;
    r[kRecXAlbedo] = in.albedo_rgba.r, r[kRecXAlbedo + 1] = in.albedo_rgba.g, r[kRecXAlbedo + 2] = in.albedo_rgba.b;
    r[kRecXRough] = in.rough, r[kRecXMetal] = in.metallic;
    r[30] = r[31] = 0.f;
    double col[3];
    pbrIblF64(P, r, col);
    r[kRecTarget] = (float)fmin(fmax(col[0], 0.0), 1.0);
    r[kRecTarget + 1] = (float)fmin(fmax(col[1], 0.0), 1.0);
    r[kRecTarget + 2] = (float)fmin(fmax(col[2], 0.0), 1.0);
    r[kRecPlain] = gclamp(plain.r, 0.f, 1.f);
    r[kRecPlain + 1] = gclamp(plain.g, 0.f, 1.f);
    r[kRecPlain + 2] = gclamp(plain.b, 0.f, 1.f);
    r[kRecPixel] = (float)pix;
    r[kRecMask] = (float)smask;
    float4 *o = (float4 *)(rec + (size_t)k * kRecFloats);
#pragma unroll
    for (int q = 0; q < kRecFloats / 4; q++) o[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
  }
}

/* --------------------------------------------------------------------------- inference in the pipeline */

/* The network's input row for job j of the two-phase raster (albedo, normal and
 * metal-rough maps sampled as the plain shader samples them): 24 features, the
 * bias 1, zeros, and the albedo alpha in the last column (which the weights ignore). */
template <int S>
__device__ __forceinline__ void nnJobRow(const DrawParams &P, int j, int need_deriv, float *f) {
  const uint2 jb = P.jobs[j];
  const uint32_t pix = jb.x & 0x0fffffffu, w = jb.y;
  const int px = (int)(pix % (uint32_t)P.fb_w), py = (int)(pix / (uint32_t)P.fb_w);
  const int id = (int)(w >> 2), inst = (int)(w & 3u);
  const PrimSetup ps = P.prims[id];
  const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
  const float fx = (float)px, fy = (float)py;
  float4 bs;
  if (S == 4) {
    if (!bary(T, fx + 0.5f, fy + 0.5f, bs)) {
#pragma unroll
      for (int t = 3; t >= 0; t--) {
        float4 b;
        if (bary(T, sampleX(t) + fx, sampleY(t) + fy, b)) bs = b;
      }
    }
  } else {
    bary(T, fx + 0.5f, fy + 0.5f, bs);
  }
  /* no quad partners: the LOD comes from the triangle's gradients */
  const Frag fr = makeFrag<S, RHS_SHADING_NEURAL_PBR>(P, ps, T, px, py, bs, inst, 0);
  const NnIn in = nnFastInputs(P, fr, nnUvGrad(P, ps, T, bs));
  nnFeatures(in.N, in.V, in.rough, in.metal, in.albedo.r, in.albedo.g, in.albedo.b, f);
  f[kNnK1 - 1] = in.albedo.a;
}

/* Learned shading of the fragments the two-phase raster queued, one kernel:
 * each wavefront takes 64 jobs, every lane writes its job's input row to the
 * wavefront's LDS tile and takes its layer-1 operands from there; the hidden
 * activations then stay in registers (each layer's result tiles are the next
 * layer's B operand, with the weights' columns in nnRegUnit order), and the
 * outputs go back to their fragments' lanes through the tile for coalesced
 * colour writes to every sample each fragment won. Weights come from L1/L2. */
template <int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nn_shade(DrawParams P, const NnPacked *W, int need_deriv) {
  __shared__ __bf16 xbuf[4][64 * kXs]; /* per wavefront: input rows, then outputs */
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  __bf16 *X = xbuf[wave];
  const int n = P.dcount[2];
  const int nchunks = (n + 63) / 64;
  unsigned long long nfs = 0, nwr = 0;
  const XcdShare xs = xcdShare(nchunks);
  for (int ch = xs.lo + xs.idx * 4 + wave; ch < xs.hi; ch += xs.cnt * 4) {
    const int jme = ch * 64 + l;
    const bool mine = jme < n;
    uint32_t jx = 0;
    float alpha = 0.f;
    {
      float f[kNnK1];
#pragma unroll
      for (int k = 0; k < kNnK1; k++) f[k] = 0.f;
      if (mine) {
        jx = P.jobs[jme].x;
        nnJobRow<S>(P, jme, need_deriv, f);
        alpha = (float)(__bf16)f[kNnK1 - 1];
      }
#pragma unroll
      for (int q = 0; q < kNnK1 / 8; q++) {
        bf16x8 v;
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = (__bf16)f[8 * q + k];
        *(bf16x8 *)&X[l * kXs + 8 * q] = v;
      }
    }
    waveSync();
    f32x4 yq[4];
    nnLayers(W->w1, W->w2r, W->w3r, kNnK1, kNnHid, W->b2, W->b3, X, yq);
    waveSync(); /* every lane has read its operands before the tile is overwritten */
    /* outputs to their fragments' lanes through the tile */
    float *Y = (float *)X;
    if (g == 0) {
#pragma unroll
      for (int q = 0; q < 4; q++) *(f32x4 *)&Y[(16 * q + c) * 4] = yq[q];
    }
    waveSync();
    if (mine) {
      const f32x4 y = *(const f32x4 *)&Y[l * 4];
      const uint32_t pix = jx & 0x0fffffffu, m = jx >> 28;
      const uint32_t packed = pack4(u8t(gclamp(y[0], 0.f, 1.f) * 255.f), u8t(gclamp(y[1], 0.f, 1.f) * 255.f),
                                    u8t(gclamp(y[2], 0.f, 1.f) * 255.f), u8t(gclamp(alpha, 0.f, 1.f) * 255.f));
      nfs++;
#pragma unroll
      for (int t = 0; t < S; t++)
        if ((m >> t) & 1u) {
          P.color[(size_t)pix * S + t] = packed;
          nwr++;
        }
    }
    waveSync();
  }
  addShadeCounts(P, nfs, nwr);
}

/* --------------------------------------------------------------------------- evaluation */

constexpr int kNnHistBins = 1024; /* bins of 1/16 of 1/255, the last one open */

__device__ __forceinline__ int histBin(float err255) {
  const int b = (int)(err255 * 16.f);
  return b < 0 ? 0 : (b >= kNnHistBins ? kNnHistBins - 1 : b);
}

/* Network and plain-shader errors against the FP64 colours over records
 * [first, first + count) (of material `material`, or all if < 0): per-channel
 * |error| * 255 into histograms, plus sums and maxima. Also writes the
 * network outputs when `out` is given. One wavefront per 64 records. */
__global__ __launch_bounds__(256) void k_nn_eval(const NnPacked *W, const float *rec, long long first, int count,
                                                 int material, unsigned int *hist_nn, unsigned int *hist_plain,
                                                 double *sums, float *out) {
  __shared__ NnWaveLds lds[4];
  __shared__ NnBlockW bw;
  __shared__ unsigned int h_nn[kNnHistBins], h_pl[kNnHistBins];
  __shared__ double s_sum[6];
  for (int i = threadIdx.x; i < kNnHistBins; i += blockDim.x) h_nn[i] = h_pl[i] = 0;
  if (threadIdx.x < 6) s_sum[threadIdx.x] = 0;
  nnLoadBlockW(*W, bw);
  __syncthreads();
  NnWaveLds &L = lds[threadIdx.x >> 6];
  double e_nn = 0, e_pl = 0, m_nn = 0, m_pl = 0, n_ok = 0;
  for (int b0 = blockIdx.x * blockDim.x; b0 < count; b0 += gridDim.x * blockDim.x) {
    const int i = b0 + threadIdx.x;
    const float *r = rec + (size_t)(first + (i < count ? i : 0)) * kRecFloats;
    const bool act = i < count && (material < 0 || (int)r[kRecMaterial] == material);
    float f[kNnK1];
    recFeatures(r, f);
    float y[3];
    nnForward64(bw, L, f, y);
    if (out && i < count) {
      out[(size_t)i * 3 + 0] = y[0];
      out[(size_t)i * 3 + 1] = y[1];
      out[(size_t)i * 3 + 2] = y[2];
    }
    if (!act) continue;
    n_ok += 1;
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
      const float en = fabsf(gclamp(y[ch], 0.f, 1.f) - r[kRecTarget + ch]) * 255.f;
      const float ep = fabsf(r[kRecPlain + ch] - r[kRecTarget + ch]) * 255.f;
      atomicAdd(&h_nn[histBin(en)], 1u);
      atomicAdd(&h_pl[histBin(ep)], 1u);
      e_nn += en;
      e_pl += ep;
      m_nn = fmax(m_nn, (double)en);
      m_pl = fmax(m_pl, (double)ep);
    }
  }
  atomicAdd(&s_sum[0], e_nn);
  atomicAdd(&s_sum[1], e_pl);
  atomicAdd(&s_sum[4], n_ok);
  __syncthreads();
  for (int i = threadIdx.x; i < kNnHistBins; i += blockDim.x) {
    if (h_nn[i]) atomicAdd(&hist_nn[i], h_nn[i]);
    if (h_pl[i]) atomicAdd(&hist_plain[i], h_pl[i]);
  }
  if (threadIdx.x == 0) {
    atomicAdd(&sums[0], s_sum[0]);
    atomicAdd(&sums[1], s_sum[1]);
    atomicAdd(&sums[4], s_sum[4]);
  }
  /* maxima: through the integer bits of non-negative doubles */
  atomicMax((unsigned long long *)&sums[2], (unsigned long long)__double_as_longlong(m_nn));
  atomicMax((unsigned long long *)&sums[3], (unsigned long long)__double_as_longlong(m_pl));
}

/* --------------------------------------------------------------------------- training */

/* Per-wavefront LDS of one training chunk (64 fragments). */
struct NnTrainLds {
  __bf16 x[64 * kNnK1];    /* X[frag][feat] */
  __bf16 xt[kNnK1 * 64];   /* X^T[feat][frag] */
  __bf16 h1[64 * kNnHid];  /* H1[frag][unit] */
  __bf16 h1t[kNnHid * 64]; /* H1^T, later dH1^T */
  __bf16 h2[64 * kNnHid];  /* H2, later dH2[frag][unit] */
  __bf16 h2t[kNnHid * 64]; /* H2^T, later dH2^T */
  __bf16 dy[64 * 32];      /* dY[frag][channel] (3 used) */
  __bf16 dyt[16 * 64];     /* dY^T[channel][frag] */
  float yf[64 * 4];        /* outputs of the chunk */
  float g[kNnParams];      /* gradient sums of this block */
};

/* sum over the 16 rows of a product tile: MFMA with an all-ones operand */
__device__ __forceinline__ bf16x8 onesOp() {
  bf16x8 o;
#pragma unroll
  for (int k = 0; k < 8; k++) o[k] = (__bf16)1.f;
  return o;
}

/* One training step's gradient of mean squared error over `batch` records
 * drawn from [first, first + count) (or the `batch` records starting at
 * `first` when `count` is 0): forward and backward on MFMA, gradients summed
 * per block in LDS. A block takes chunks of 64 records with four wavefronts:
 * wavefront w computes the output tiles of hidden units 16w..16w+15 of every
 * layer (and fragment tile w where a layer has one row tile), so each gradient
 * sum has one writer and the same order as with one wavefront; wavefront 0
 * stages the records and takes their errors. Each block writes its sums (FP32,
 * unscaled) to its own row of `partial` and its squared error to
 * lossp[block]: no atomics, so k_nn_update's fixed-order sum makes a step
 * reproducible bit for bit. */
constexpr int kNnGradThreads = 256;
__global__ __launch_bounds__(kNnGradThreads) void k_nn_grad(const NnPacked *W, const float *rec, long long first,
                                                            long long count, int batch, uint32_t seed,
                                                            float *partial, double *lossp) {
  __shared__ NnTrainLds L;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  const int m = w; /* this wavefront's row tile of the hidden layers */
  for (int i = threadIdx.x; i < kNnParams; i += kNnGradThreads) L.g[i] = 0.f;
  double lsum = 0;
  const bf16x8 ones = onesOp();
  for (int c0 = blockIdx.x * 64; c0 < batch; c0 += gridDim.x * 64) {
    const int bi = c0 + l;
    const bool act = bi < batch;
    long long ri = first + bi;
    if (count > 0) ri = first + (long long)(hash32((uint32_t)bi * 0x9e3779b9u ^ seed) % (uint32_t)count);
    const float *r = rec + (size_t)(act ? ri : first) * kRecFloats;
    if (w == 0) {
      float f[kNnK1];
      recFeatures(r, f);
      if (!act) {
#pragma unroll
        for (int k = 0; k < kNnK1; k++) f[k] = 0.f;
      }
#pragma unroll
      for (int k = 0; k < kNnK1; k++) {
        L.x[l * kNnK1 + k] = (__bf16)f[k];
        L.xt[k * 64 + l] = (__bf16)f[k];
      }
    }
    __syncthreads();
    /* forward: H1, H2 (and transposes), Y */
    {
      const bf16x8 a = ldOp(W->w1, kNnK1, 16 * m, 0);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mma16(a, ldOp(L.x, kNnK1, 16 * n, 0), d);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const __bf16 h = (__bf16)gmax(d[v], 0.f);
          const int u = 16 * m + 4 * g + v, fr = 16 * n + c;
          L.h1[fr * kNnHid + u] = h;
          L.h1t[u * 64 + fr] = h;
        }
      }
    }
    __syncthreads();
    {
      const bf16x8 a0 = ldOp(W->w2, kNnHid, 16 * m, 0), a1 = ldOp(W->w2, kNnHid, 16 * m, 32);
      f32x4 bias;
#pragma unroll
      for (int v = 0; v < 4; v++) bias[v] = W->b2[16 * m + 4 * g + v];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const f32x4 d = mma16(a1, ldOp(L.h1, kNnHid, 16 * n, 32), mma16(a0, ldOp(L.h1, kNnHid, 16 * n, 0), bias));
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const __bf16 h = (__bf16)gmax(d[v], 0.f);
          const int u = 16 * m + 4 * g + v, fr = 16 * n + c;
          L.h2[fr * kNnHid + u] = h;
          L.h2t[u * 64 + fr] = h;
        }
      }
    }
    __syncthreads();
    {
      const bf16x8 a0 = ldOp(W->w3, kNnHid, 0, 0), a1 = ldOp(W->w3, kNnHid, 0, 32);
      f32x4 bias;
#pragma unroll
      for (int v = 0; v < 4; v++) bias[v] = W->b3[4 * g + v];
      const int n = w, fr = 16 * n + c;
      const f32x4 d = mma16(a1, ldOp(L.h2, kNnHid, 16 * n, 32), mma16(a0, ldOp(L.h2, kNnHid, 16 * n, 0), bias));
      if (g == 0) {
#pragma unroll
        for (int v = 0; v < 4; v++) L.yf[fr * 4 + v] = d[v];
      }
    }
    __syncthreads();
    /* record l's error and dY (wavefront 0) */
    if (w == 0) {
      float dyv[3];
#pragma unroll
      for (int ch = 0; ch < 3; ch++) {
        const float y = L.yf[l * 4 + ch];
        const float e = act ? y - r[kRecTarget + ch] : 0.f;
        lsum += (double)e * e;
        dyv[ch] = 2.f * e;
      }
#pragma unroll
      for (int ch = 0; ch < 32; ch++) L.dy[l * 32 + ch] = (__bf16)(ch < 3 ? dyv[ch] : 0.f);
#pragma unroll
      for (int ch = 0; ch < 16; ch++) L.dyt[ch * 64 + l] = (__bf16)(ch < 3 ? dyv[ch] : 0.f);
    }
    __syncthreads();
    /* dW3 = dY^T . H2 (unit tile w), db3 = row sums of dY^T (wavefront 0) */
    {
      const bf16x8 a0 = ldOp(L.dyt, 64, 0, 0), a1 = ldOp(L.dyt, 64, 0, 32);
      const int n = w;
      f32x4 d = {0.f, 0.f, 0.f, 0.f};
      d = mma16(a1, ldOp(L.h2t, 64, 16 * n, 32), mma16(a0, ldOp(L.h2t, 64, 16 * n, 0), d));
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int ch = 4 * g + v;
        if (ch < kNnOut) L.g[kOffW3 + ch * kNnHid + 16 * n + c] += d[v];
      }
      if (w == 0) {
        f32x4 s = {0.f, 0.f, 0.f, 0.f};
        s = mma16(a1, ones, mma16(a0, ones, s));
        if (c == 0) {
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (4 * g + v < kNnOut) L.g[kOffB3 + 4 * g + v] += s[v];
        }
      }
    }
    __syncthreads(); /* H2^T is read above before it is overwritten below */
    /* dH2^T = W3^T . dY^T, masked by H2 > 0, written over H2 and H2^T */
    {
      bf16x8 bdy[4];
#pragma unroll
      for (int n = 0; n < 4; n++) bdy[n] = ldOp(L.dy, 32, 16 * n, 0);
      const bf16x8 a = ldOp(W->w3t, 32, 16 * m, 0);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mma16(a, bdy[n], d);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int u = 16 * m + 4 * g + v, fr = 16 * n + c;
          const bool on = (float)L.h2t[u * 64 + fr] > 0.f;
          const __bf16 dh = (__bf16)(on ? d[v] : 0.f);
          L.h2t[u * 64 + fr] = dh;
          L.h2[fr * kNnHid + u] = dh;
        }
      }
    }
    __syncthreads();
    /* dW2 = dH2^T . H1, db2 = row sums of dH2^T */
    {
      const bf16x8 a0 = ldOp(L.h2t, 64, 16 * m, 0), a1 = ldOp(L.h2t, 64, 16 * m, 32);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mma16(a1, ldOp(L.h1t, 64, 16 * n, 32), mma16(a0, ldOp(L.h1t, 64, 16 * n, 0), d));
#pragma unroll
        for (int v = 0; v < 4; v++) L.g[kOffW2 + (16 * m + 4 * g + v) * kNnHid + 16 * n + c] += d[v];
      }
      f32x4 s = {0.f, 0.f, 0.f, 0.f};
      s = mma16(a1, ones, mma16(a0, ones, s));
      if (c == 0) {
#pragma unroll
        for (int v = 0; v < 4; v++) L.g[kOffB2 + 16 * m + 4 * g + v] += s[v];
      }
    }
    __syncthreads(); /* H1^T is read above before it is overwritten below */
    /* dH1^T = W2^T . dH2^T, masked by H1 > 0, written over H1^T */
    {
      bf16x8 bd[4][2];
#pragma unroll
      for (int n = 0; n < 4; n++)
#pragma unroll
        for (int kb = 0; kb < 2; kb++) bd[n][kb] = ldOp(L.h2, kNnHid, 16 * n, 32 * kb);
      const bf16x8 a0 = ldOp(W->w2t, kNnHid, 16 * m, 0), a1 = ldOp(W->w2t, kNnHid, 16 * m, 32);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mma16(a1, bd[n][1], mma16(a0, bd[n][0], d));
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int u = 16 * m + 4 * g + v, fr = 16 * n + c;
          const bool on = (float)L.h1t[u * 64 + fr] > 0.f;
          L.h1t[u * 64 + fr] = (__bf16)(on ? d[v] : 0.f);
        }
      }
    }
    __syncthreads();
    /* dW1 (with db1 in column 24) = dH1^T . X */
    {
      const bf16x8 a0 = ldOp(L.h1t, 64, 16 * m, 0), a1 = ldOp(L.h1t, 64, 16 * m, 32);
#pragma unroll
      for (int n = 0; n < 2; n++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mma16(a1, ldOp(L.xt, 64, 16 * n, 32), mma16(a0, ldOp(L.xt, 64, 16 * n, 0), d));
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int u = 16 * m + 4 * g + v, k = 16 * n + c;
          if (k < kNnIn)
            L.g[kOffW1 + u * kNnIn + k] += d[v];
          else if (k == kNnIn)
            L.g[kOffB1 + u] += d[v];
        }
      }
    }
    __syncthreads(); /* the chunk's tiles are read before the next chunk overwrites them */
  }
  float *dst = partial + (size_t)blockIdx.x * kNnParams;
  for (int i = threadIdx.x; i < kNnParams; i += kNnGradThreads) dst[i] = L.g[i];
  if (w == 0) {
    double t = lsum;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (l == 0) lossp[blockIdx.x] = t;
  }
}

/* Parameter i's entries in the packed bf16 operands (the padding entries are
 * zero from allocation and never written). */
__device__ __forceinline__ void nnPackParam(NnPacked *W, int i, float x) {
  const __bf16 h = (__bf16)x;
  if (i < kOffB1) {
    W->w1[(i / kNnIn) * kNnK1 + i % kNnIn] = h;
  } else if (i < kOffW2) {
    W->w1[(i - kOffB1) * kNnK1 + kNnIn] = h;
  } else if (i < kOffB2) {
    const int j = i - kOffW2, u = j / kNnHid, k = j % kNnHid;
    W->w2[j] = h;
    W->w2t[k * kNnHid + u] = h;
    W->w2r[u * kNnHid + nnRegSlot(k)] = h;
  } else if (i < kOffW3) {
    W->b2[i - kOffB2] = x;
  } else if (i < kOffB3) {
    const int j = i - kOffW3, ch = j / kNnHid, k = j % kNnHid;
    W->w3[j] = h;
    W->w3t[k * 32 + ch] = h;
    W->w3r[ch * kNnHid + nnRegSlot(k)] = h;
  } else {
    W->b3[i - kOffB3] = x;
  }
}

/* The gradient of a step: the `nparts` block sums of k_nn_grad added in a
 * fixed order (16 interleaved partial sums per parameter, each over batches of
 * 8 independent loads, then those 16 in order), and the loss by a fixed tree.
 * With `gout` the gradient is only written there. Otherwise Adam (step > 0)
 * updates the FP32 parameters with the gradient scaled by `scale`, and every
 * parameter is repacked into the bf16 operands (step 0: repack only).
 * kNnUpdateParams parameters per block of 256 threads. */
constexpr int kNnUpdateSlices = 16, kNnUpdateParams = 256 / kNnUpdateSlices;
static_assert(kNnUpdateParams == 16 && kNnUpdateBlocks * kNnUpdateParams >= kNnParams, "k_nn_update grid");
__global__ __launch_bounds__(256) void k_nn_update(float *p, const float *partial, int nparts, const double *lossp,
                                                   double *loss, float *gout, float *m1, float *m2, NnPacked *W,
                                                   float lr, float b1, float b2, float eps, float scale, int step) {
  __shared__ float part[kNnUpdateSlices][kNnUpdateParams];
  __shared__ double lpart[256];
  const int t = threadIdx.x % kNnUpdateParams, s = threadIdx.x / kNnUpdateParams;
  const int i = blockIdx.x * kNnUpdateParams + t;
  float acc = 0.f;
  if (i < kNnParams) {
    for (int b0 = s; b0 < nparts; b0 += 8 * kNnUpdateSlices) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int b = b0 + j * kNnUpdateSlices;
        v[j] = b < nparts ? partial[(size_t)b * kNnParams + i] : 0.f;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) acc += v[j];
    }
  }
  part[s][t] = acc;
  __syncthreads();
  if (s == 0 && i < kNnParams) {
    float gs = 0.f;
#pragma unroll
    for (int k = 0; k < kNnUpdateSlices; k++) gs += part[k][t];
    if (gout) {
      gout[i] = gs;
    } else {
      if (step > 0) {
        const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
        const float gi = gs * scale;
        const float a = b1 * m1[i] + (1.f - b1) * gi, b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
      }
      nnPackParam(W, i, p[i]);
    }
  }
  if (loss && blockIdx.x == 0) {
    double v = 0;
    for (int b = threadIdx.x; b < nparts; b += 256) v += lossp[b];
    lpart[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) lpart[threadIdx.x] += lpart[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) *loss = lpart[0];
  }
}

}  // namespace rhip
