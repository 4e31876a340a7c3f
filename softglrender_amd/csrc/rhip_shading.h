/*
 * rhip_shading.h -- device-side texture sampling, vertex shaders and fragment
 * shaders of librhip (gfx950).
 *
 * The arithmetic follows SoftGLRender's software pipeline operation by
 * operation, including the fused multiply-adds its x86 build (-O3 -mavx2 -mfma)
 * emits, so that texel values, depths and coverage agree bit for bit where the
 * reference is bit-deterministic. Every kernel TU is compiled with
 * -ffp-contract=off: each fusion below is an explicit fmaf.
 */
#pragma once
#include "rhip_internal.h"

namespace rhip {

struct f2 {
  float x, y;
};
struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 scl3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
/* glm::dot(vec3) as compiled by the reference build: fma(z, z', fma(x, x', y*y')) */
__device__ __forceinline__ float dot3(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.x, b.x, a.y * b.y)); }
__device__ __forceinline__ f3 nrm3(f3 v) { return scl3(v, 1.0f / sqrtf(dot3(v, v))); }
/* glm::dot / glm::normalize as the reference build compiles them in the fragment shaders and the
 * normal-map set-up of the vertex shaders: fma(z, z', fma(y, y', x x')) */
__device__ __forceinline__ float dot3b(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ f3 nrm3b(f3 v) { return scl3(v, 1.0f / sqrtf(dot3b(v, v))); }
/* glm::dot where the reference build rounds each product on its own (products shared by two dot
 * products): (x x' + y y') + z z' */
__device__ __forceinline__ float dot3s(f3 a, f3 b) {
  const float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
  return (px + py) + pz;
}
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}
/* Correctly rounded powf (double precision, rounded once): the learned shader's
 * feature encoding and inputs outside the host-libm tables (rhip_libm.h). */
__device__ __forceinline__ float rpowf(float x, float y) { return (float)pow((double)x, (double)y); }

__device__ __forceinline__ float gmax(float a, float b) { return a < b ? b : a; } /* glm::max */
__device__ __forceinline__ float gmin(float a, float b) { return b < a ? b : a; } /* glm::min */
/* glm::clamp = min(max(x, lo), hi) with glm's comparisons (keeps -0 and NaN like the reference) */
__device__ __forceinline__ float gclamp(float v, float lo, float hi) { return gmin(gmax(v, lo), hi); }

/* mat4 * vec4(p, 1): (fma(c0, x, c1*y)) + (fma(c2, z, c3))  -- glm type_mat4x4.inl:536-575 */
__device__ __forceinline__ float4 m4p(const float *m, float x, float y, float z) {
  float r[4];
#pragma unroll
  for (int i = 0; i < 4; i++) r[i] = fmaf(m[i], x, m[4 + i] * y) + fmaf(m[8 + i], z, m[12 + i]);
  return make_float4(r[0], r[1], r[2], r[3]);
}
/* mat3 * vec3: fma(c2, z, fma(c0, x, c1*y)); columns `cs` floats apart */
__device__ __forceinline__ f3 m3v(const float *m, int cs, f3 v) {
  return {fmaf(m[2 * cs + 0], v.z, fmaf(m[0], v.x, m[cs + 0] * v.y)),
          fmaf(m[2 * cs + 1], v.z, fmaf(m[1], v.x, m[cs + 1] * v.y)),
          fmaf(m[2 * cs + 2], v.z, fmaf(m[2], v.x, m[cs + 2] * v.y))};
}

/* --------------------------------------------------------------------------- texel addressing
 * Buffer.h:31-33 (linear), :151-158 (4x4 tiles), :185-202 (32x32 tiles, 16-bit Morton inside). */
__device__ __forceinline__ uint32_t morton16(uint32_t x, uint32_t y) {
  uint32_t r = x | (y << 16);
  r = (r | (r << 4)) & 0x0f0f0f0fu;
  r = (r | (r << 2)) & 0x33333333u;
  r = (r | (r << 1)) & 0x55555555u;
  return (r | (r >> 15)) & 0xffffu;
}

__device__ __forceinline__ size_t texelIndex(const TexLevel &l, int layout, int x, int y) {
  if (layout == LAYOUT_TILED) {
    uint32_t tx = (uint32_t)x >> 2, ty = (uint32_t)y >> 2;
    return ((size_t)(ty * (uint32_t)l.tiles_x + tx) << 4) + ((uint32_t)(y & 3) << 2) + (uint32_t)(x & 3);
  }
  if (layout == LAYOUT_MORTON) {
    uint32_t tx = (uint32_t)x >> 5, ty = (uint32_t)y >> 5;
    return ((size_t)(ty * (uint32_t)l.tiles_x + tx) << 10) + morton16(x & 31, y & 31);
  }
  return (size_t)y * (uint32_t)l.w + (uint32_t)x;
}

__device__ __forceinline__ int chan(uint32_t c, int i) { return (int)((c >> (8 * i)) & 255u); }
__device__ __forceinline__ uint32_t pack4(int r, int g, int b, int a) {
  return (uint32_t)(r & 255) | ((uint32_t)(g & 255) << 8) | ((uint32_t)(b & 255) << 16) | ((uint32_t)(a & 255) << 24);
}
/* float -> u8 the way the reference's conversion compiles: truncate to int32, keep the low byte. */
__device__ __forceinline__ int u8t(float v) { return (int)(uint8_t)(int32_t)v; }

/* glm::mix on u8vec4 as compiled: u8(fma(y, a, x * (1 - a))) per channel (SamplerSoft.h:164, :265) */
__device__ __forceinline__ uint32_t mixRGBA(uint32_t x, uint32_t y, float a) {
  float na = 1.0f - a;
  return pack4(u8t(fmaf((float)chan(y, 0), a, (float)chan(x, 0) * na)),
               u8t(fmaf((float)chan(y, 1), a, (float)chan(x, 1) * na)),
               u8t(fmaf((float)chan(y, 2), a, (float)chan(x, 2) * na)),
               u8t(fmaf((float)chan(y, 3), a, (float)chan(x, 3) * na)));
}
__device__ __forceinline__ float mixF(float x, float y, float a) { return fmaf(y, a, x * (1.0f - a)); }

template <typename T>
struct Texel;
template <>
struct Texel<uint32_t> {
  __device__ static uint32_t mix(uint32_t a, uint32_t b, float t) { return mixRGBA(a, b, t); }
  __device__ static uint32_t border(const TexDesc &d) { return d.border == 1 ? 0xffffffffu : 0u; }
};
template <>
struct Texel<float> {
  __device__ static float mix(float a, float b, float t) { return mixF(a, b, t); }
  __device__ static float border(const TexDesc &d) { return d.border == 1 ? 1.f : 0.f; }
};

/* pixelWithWrapMode (SamplerSoft.h:170-213). REPEAT masks with (n-1): the
 * reference's CoordMod macro parses as i & (2n-1) & (n-1). Where the read at
 * (x, y) lands: 0 the texel at `idx`, 1 the border colour, 2 zero. */
__device__ __forceinline__ int wrapTexel(const TexDesc &d, const TexLevel &l, int x, int y, int wrap, size_t &idx) {
  int w = l.w, h = l.h;
  if (wrap == 0) {
    x &= w - 1;
    y &= h - 1;
  } else if (wrap == 1) {
    x &= 2 * w - 1;
    y &= 2 * h - 1;
    x -= w;
    y -= h;
    x = x >= 0 ? x : (-1 - x);
    y = y >= 0 ? y : (-1 - y);
    x = w - 1 - x;
    y = h - 1 - y;
  } else if (wrap == 2) {
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
  } else {
    if (x < 0 || x >= w || y < 0 || y >= h) return 1;
  }
  if ((unsigned)x >= (unsigned)w || (unsigned)y >= (unsigned)h) return 2;
  idx = texelIndex(l, d.layout, x, y);
  return 0;
}

template <typename T>
__device__ __forceinline__ T fetchWrap(const TexDesc &d, const TexLevel &l, int x, int y, int wrap) {
  int w = l.w, h = l.h;
  if (wrap == 0) {
    x &= w - 1;
    y &= h - 1;
  } else if (wrap == 1) {
    x &= 2 * w - 1;
    y &= 2 * h - 1;
    x -= w;
    y -= h;
    x = x >= 0 ? x : (-1 - x);
    y = y >= 0 ? y : (-1 - y);
    x = w - 1 - x;
    y = h - 1 - y;
  } else if (wrap == 2) {
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
  } else {
    if (x < 0 || x >= w || y < 0 || y >= h) return Texel<T>::border(d);
  }
  if ((unsigned)x >= (unsigned)w || (unsigned)y >= (unsigned)h) return T(0);
  const T *p = (const T *)(d.base + l.offset);
  return p[texelIndex(l, d.layout, x, y)];
}

/* samplePixelBilinear (SamplerSoft.h:254-266) at texel-space (ux, uy) */
template <typename T>
__device__ __forceinline__ T bilinear(const TexDesc &d, const TexLevel &l, float ux, float uy, int wrap) {
  float fx0 = ux - 0.5f, fy0 = uy - 0.5f;
  float flx = floorf(fx0), fly = floorf(fy0);
  int x = (int)flx, y = (int)fly;
  T s1 = fetchWrap<T>(d, l, x, y, wrap);
  T s2 = fetchWrap<T>(d, l, x + 1, y, wrap);
  T s3 = fetchWrap<T>(d, l, x, y + 1, wrap);
  T s4 = fetchWrap<T>(d, l, x + 1, y + 1, wrap);
  float fx = fx0 - flx, fy = fy0 - fly;
  return Texel<T>::mix(Texel<T>::mix(s1, s2, fx), Texel<T>::mix(s3, s4, fx), fy);
}

template <typename T>
__device__ __forceinline__ T sampleLevel(const TexDesc &d, int layer, int level, f2 uv, bool linear) {
  const TexLevel &l = d.lv[layer][level];
  float tx = uv.x * (float)l.w, ty = uv.y * (float)l.h;
  if (linear) return bilinear<T>(d, l, tx, ty, d.wrap);
  return fetchWrap<T>(d, l, (int)floorf(tx), (int)floorf(ty), d.wrap);
}

/* BaseSampler::textureImpl (SamplerSoft.h:117-168) */
template <typename T>
__device__ T textureImpl(const TexDesc &d, int layer, f2 uv, float lod) {
  int fm = d.filter_min;
  if (fm == 0) return sampleLevel<T>(d, layer, 0, uv, false);
  if (fm == 1) return sampleLevel<T>(d, layer, 0, uv, true);
  int maxLevel = d.levels - 1;
  if (fm == 2 || fm == 3) {
    int level = min(max((int)ceilf(lod + 0.5f) - 1, 0), maxLevel);
    return sampleLevel<T>(d, layer, level, uv, fm == 3);
  }
  int hi = min(max((int)floorf(lod), 0), maxLevel);
  int lo = min(max(hi + 1, 0), maxLevel);
  T th = sampleLevel<T>(d, layer, hi, uv, fm == 5);
  if (hi == lo) return th;
  T tl = sampleLevel<T>(d, layer, lo, uv, fm == 5);
  return Texel<T>::mix(th, tl, lod - floorf(lod));
}

/* ------------------------------------------------------------------ RGBA8 sampling on byte-valued floats
 *
 * The same arithmetic as textureImpl<uint32_t>, with each channel kept as a
 * float holding its byte instead of being repacked after every mix: the
 * reference's u8 conversion of a mix result in [0, 256) is its truncation, and
 * a NaN weight (which makes every mix of that bilinear NaN) converts to 0 at
 * the end of the bilinear, as the byte chain gives. */
__device__ __forceinline__ float4 bytesF(uint32_t p) {
  return make_float4((float)(p & 255u), (float)((p >> 8) & 255u), (float)((p >> 16) & 255u), (float)(p >> 24));
}
__device__ __forceinline__ float4 mixB(float4 x, float4 y, float a) {
  const float na = 1.0f - a;
  return make_float4(truncf(fmaf(y.x, a, x.x * na)), truncf(fmaf(y.y, a, x.y * na)), truncf(fmaf(y.z, a, x.z * na)),
                     truncf(fmaf(y.w, a, x.w * na)));
}
__device__ __forceinline__ float4 nanToZero(float4 v) {
  return make_float4(v.x == v.x ? v.x : 0.f, v.y == v.y ? v.y : 0.f, v.z == v.z ? v.z : 0.f, v.w == v.w ? v.w : 0.f);
}

/* A texel index (texelIndex) is the sum of a column part and a row part: linear
 * x + y w; 4x4 tiles ((x >> 2) << 4) + (x & 3) and ((y >> 2) tiles_x << 4) +
 * ((y & 3) << 2); 32x32 tiles with the Morton order inside, whose two halves
 * interleave the low five bits of x and of y. So the four texels of a bilinear
 * footprint need the wrap and the parts of two columns and two rows only. */
__device__ __forceinline__ uint32_t spread5(uint32_t v) {
  v &= 31u;
  v = (v | (v << 4)) & 0x0f0fu;
  v = (v | (v << 2)) & 0x3333u;
  return (v | (v << 1)) & 0x5555u;
}
__device__ __forceinline__ uint32_t colPart(const TexLevel &l, int layout, uint32_t x) {
  if (layout == LAYOUT_TILED) return ((x >> 2) << 4) + (x & 3u);
  if (layout == LAYOUT_MORTON) return ((x >> 5) << 10) + spread5(x);
  return x;
}
__device__ __forceinline__ uint32_t rowPart(const TexLevel &l, int layout, uint32_t y) {
  if (layout == LAYOUT_TILED) return (((y >> 2) * (uint32_t)l.tiles_x) << 4) + ((y & 3u) << 2);
  if (layout == LAYOUT_MORTON) return (((y >> 5) * (uint32_t)l.tiles_x) << 10) + (spread5(y) << 1);
  return y * (uint32_t)l.w;
}

/* One axis of pixelWithWrapMode (SamplerSoft.h:170-213): the wrapped
 * coordinate; `out` when CLAMP_TO_BORDER finds it outside (border colour),
 * `bad` when the wrapped coordinate is still outside (zero). */
struct AxisTexel {
  uint32_t v;
  bool out, bad;
};
__device__ __forceinline__ AxisTexel wrapAxis(int v, int n, int wrap) {
  AxisTexel a{0u, false, false};
  if (wrap == 0) {
    v &= n - 1;
  } else if (wrap == 1) {
    v &= 2 * n - 1;
    v -= n;
    v = v >= 0 ? v : (-1 - v);
    v = n - 1 - v;
  } else if (wrap == 2) {
    v = v < 0 ? 0 : (v >= n ? n - 1 : v);
  } else {
    a.out = v < 0 || v >= n;
  }
  a.bad = (unsigned)v >= (unsigned)n;
  a.v = (uint32_t)v;
  return a;
}

/* The texel of column c and row r: border colour, zero or the image's texel (fetchWrap's order of tests). */
__device__ __forceinline__ float4 readTexel(const TexDesc &d, const uint32_t *img, int wrap, const AxisTexel &c,
                                            uint32_t cp, const AxisTexel &r, uint32_t rp) {
  if (wrap == 3 && (c.out || r.out)) return bytesF(Texel<uint32_t>::border(d));
  if (c.bad || r.bad) return make_float4(0.f, 0.f, 0.f, 0.f);
  return bytesF(img[cp + rp]);
}

__device__ __forceinline__ float4 bilinear4(const TexDesc &d, const TexLevel &l, float ux, float uy, int wrap) {
  const float fx0 = ux - 0.5f, fy0 = uy - 0.5f;
  const float flx = floorf(fx0), fly = floorf(fy0);
  const int x = (int)flx, y = (int)fly;
  const AxisTexel c0 = wrapAxis(x, l.w, wrap), c1 = wrapAxis(x + 1, l.w, wrap);
  const AxisTexel r0 = wrapAxis(y, l.h, wrap), r1 = wrapAxis(y + 1, l.h, wrap);
  const uint32_t cp0 = colPart(l, d.layout, c0.v), cp1 = colPart(l, d.layout, c1.v);
  const uint32_t rp0 = rowPart(l, d.layout, r0.v), rp1 = rowPart(l, d.layout, r1.v);
  const uint32_t *img = (const uint32_t *)(d.base + l.offset);
  const float4 s1 = readTexel(d, img, wrap, c0, cp0, r0, rp0);
  const float4 s2 = readTexel(d, img, wrap, c1, cp1, r0, rp0);
  const float4 s3 = readTexel(d, img, wrap, c0, cp0, r1, rp1);
  const float4 s4 = readTexel(d, img, wrap, c1, cp1, r1, rp1);
  const float fx = fx0 - flx, fy = fy0 - fly;
  return nanToZero(mixB(mixB(s1, s2, fx), mixB(s3, s4, fx), fy));
}

__device__ __forceinline__ float4 sampleLevel4(const TexDesc &d, int layer, int level, f2 uv, bool linear) {
  const TexLevel &l = d.lv[layer][level];
  const float tx = uv.x * (float)l.w, ty = uv.y * (float)l.h;
  if (linear) return bilinear4(d, l, tx, ty, d.wrap);
  return bytesF(fetchWrap<uint32_t>(d, l, (int)floorf(tx), (int)floorf(ty), d.wrap));
}

/* textureImpl<uint32_t> with byte-valued float channels */
__device__ __forceinline__ float4 textureImpl4(const TexDesc &d, int layer, f2 uv, float lod) {
  const int fm = d.filter_min;
  if (fm == 0) return sampleLevel4(d, layer, 0, uv, false);
  if (fm == 1) return sampleLevel4(d, layer, 0, uv, true);
  const int maxLevel = d.levels - 1;
  if (fm == 2 || fm == 3) {
    const int level = min(max((int)ceilf(lod + 0.5f) - 1, 0), maxLevel);
    return sampleLevel4(d, layer, level, uv, fm == 3);
  }
  const int hi = min(max((int)floorf(lod), 0), maxLevel);
  const int lo = min(max(hi + 1, 0), maxLevel);
  const float4 th = sampleLevel4(d, layer, hi, uv, fm == 5);
  if (hi == lo) return th;
  const float4 tl = sampleLevel4(d, layer, lo, uv, fm == 5);
  return nanToZero(mixB(th, tl, lod - floorf(lod)));
}

/* Two textures of one shape (size, levels, layout, wrap and filter) read at the
 * same coordinate and LOD: the levels, weights and texel indices are the same,
 * so they are computed once and both images are read with them. */
__device__ __forceinline__ bool sameShape(const TexDesc &a, const TexDesc &b) {
  return a.width == b.width && a.height == b.height && a.levels == b.levels && a.layout == b.layout &&
         a.wrap == b.wrap && a.filter_min == b.filter_min && a.layers == b.layers && a.type == b.type &&
         a.format == b.format;
}

/* One texel read of a shape: where the wrap mode sends it (wrapTexel) and the index there. */
struct TexelAt {
  int kind; /* 0 the texel at idx, 1 the border colour, 2 zero */
  uint32_t idx;
};
__device__ __forceinline__ TexelAt texelAt(const TexDesc &a, const TexLevel &l, int x, int y) {
  size_t idx = 0;
  const int k = wrapTexel(a, l, x, y, a.wrap, idx);
  return {k, (uint32_t)idx};
}
__device__ __forceinline__ float4 readAt(const TexDesc &d, const TexLevel &l, TexelAt t) {
  if (t.kind == 0) return bytesF(((const uint32_t *)(d.base + l.offset))[t.idx]);
  return bytesF(t.kind == 1 ? Texel<uint32_t>::border(d) : 0u);
}

__device__ __forceinline__ void sampleLevelPair(const TexDesc &a, const TexDesc &b, int layer, int level, f2 uv,
                                                bool linear, float4 &ra, float4 &rb) {
  const TexLevel &la = a.lv[layer][level], &lb = b.lv[layer][level];
  const float tx = uv.x * (float)la.w, ty = uv.y * (float)la.h;
  if (!linear) {
    const TexelAt t = texelAt(a, la, (int)floorf(tx), (int)floorf(ty));
    ra = readAt(a, la, t);
    rb = readAt(b, lb, t);
    return;
  }
  const float fx0 = tx - 0.5f, fy0 = ty - 0.5f;
  const float flx = floorf(fx0), fly = floorf(fy0);
  const int x = (int)flx, y = (int)fly;
  const int wrap = a.wrap;
  const AxisTexel c0 = wrapAxis(x, la.w, wrap), c1 = wrapAxis(x + 1, la.w, wrap);
  const AxisTexel r0 = wrapAxis(y, la.h, wrap), r1 = wrapAxis(y + 1, la.h, wrap);
  const uint32_t cp0 = colPart(la, a.layout, c0.v), cp1 = colPart(la, a.layout, c1.v);
  const uint32_t rp0 = rowPart(la, a.layout, r0.v), rp1 = rowPart(la, a.layout, r1.v);
  const float fx = fx0 - flx, fy = fy0 - fly;
  /* one image at a time: the second's texels are read after the first is mixed */
  const uint32_t *ia = (const uint32_t *)(a.base + la.offset), *ib = (const uint32_t *)(b.base + lb.offset);
  ra = nanToZero(mixB(mixB(readTexel(a, ia, wrap, c0, cp0, r0, rp0), readTexel(a, ia, wrap, c1, cp1, r0, rp0), fx),
                      mixB(readTexel(a, ia, wrap, c0, cp0, r1, rp1), readTexel(a, ia, wrap, c1, cp1, r1, rp1), fx), fy));
  rb = nanToZero(mixB(mixB(readTexel(b, ib, wrap, c0, cp0, r0, rp0), readTexel(b, ib, wrap, c1, cp1, r0, rp0), fx),
                      mixB(readTexel(b, ib, wrap, c0, cp0, r1, rp1), readTexel(b, ib, wrap, c1, cp1, r1, rp1), fx), fy));
}

/* textureImpl of two textures of one shape at one coordinate and LOD */
__device__ __forceinline__ void textureImplPair(const TexDesc &a, const TexDesc &b, int layer, f2 uv, float lod,
                                                float4 &ra, float4 &rb) {
  const int fm = a.filter_min;
  if (fm <= 1) {
    sampleLevelPair(a, b, layer, 0, uv, fm == 1, ra, rb);
    return;
  }
  const int maxLevel = a.levels - 1;
  if (fm == 2 || fm == 3) {
    const int level = min(max((int)ceilf(lod + 0.5f) - 1, 0), maxLevel);
    sampleLevelPair(a, b, layer, level, uv, fm == 3, ra, rb);
    return;
  }
  const int hi = min(max((int)floorf(lod), 0), maxLevel);
  const int lo = min(max(hi + 1, 0), maxLevel);
  sampleLevelPair(a, b, layer, hi, uv, fm == 5, ra, rb);
  if (hi == lo) return;
  float4 la, lb;
  sampleLevelPair(a, b, layer, lo, uv, fm == 5, la, lb);
  const float t = lod - floorf(lod);
  ra = nanToZero(mixB(ra, la, t));
  rb = nanToZero(mixB(rb, lb, t));
}

/* convertXYZ2UV (SamplerSoft.h:311-373); later faces win ties */
__device__ __forceinline__ void cubeFace(f3 c, int *face, f2 *uv) {
  float ax = fabsf(c.x), ay = fabsf(c.y), az = fabsf(c.z);
  bool px = c.x > 0, py = c.y > 0, pz = c.z > 0;
  float ma = 0.f, uc = 0.f, vc = 0.f;
  int idx = 0;
  if (px && ax >= ay && ax >= az) { ma = ax; uc = -c.z; vc = c.y; idx = 0; }
  if (!px && ax >= ay && ax >= az) { ma = ax; uc = c.z; vc = c.y; idx = 1; }
  if (py && ay >= ax && ay >= az) { ma = ay; uc = c.x; vc = -c.z; idx = 2; }
  if (!py && ay >= ax && ay >= az) { ma = ay; uc = c.x; vc = c.z; idx = 3; }
  if (pz && az >= ax && az >= ay) { ma = az; uc = c.x; vc = c.y; idx = 4; }
  if (!pz && az >= ax && az >= ay) { ma = az; uc = -c.x; vc = c.y; idx = 5; }
  vc = -vc;
  *face = idx;
  uv->x = 0.5f * (uc / ma + 1.0f);
  uv->y = 0.5f * (vc / ma + 1.0f);
}

struct RGBAf {
  float r, g, b, a;
};
/* k / 255.f for a byte k, correctly rounded like the division (ShaderSoft.h:80-83 divides the texel by 255):
 * q = k c with c = 1/255 rounded, then one Newton step with fused multiply-adds; equal to the division for
 * every byte (all 256 checked exactly). */
__device__ __forceinline__ float div255(float k) {
  const float c = 1.0f / 255.0f;
  const float q = k * c;
  return fmaf(fmaf(-q, 255.0f, k), c, q);
}
__device__ __forceinline__ RGBAf unpack01(uint32_t p) {
  return {div255((float)chan(p, 0)), div255((float)chan(p, 1)), div255((float)chan(p, 2)), div255((float)chan(p, 3))};
}
__device__ __forceinline__ RGBAf unpack01(float4 b) { return {div255(b.x), div255(b.y), div255(b.z), div255(b.w)}; }

/* --------------------------------------------------------------------------- vertex shaders */

__device__ __forceinline__ int varyingCount(int shading) {
  return (shading == RHS_SHADING_BLINN_PHONG || shading == RHS_SHADING_BLINN_PHONG_DIR) ? 32
         : (shading == RHS_SHADING_PBR || shading == RHS_SHADING_NEURAL_PBR)            ? 28
         : shading == RHS_SHADING_VERTEX_COLOR                                          ? 4
                                                                                        : 0;
}

/* BasicSoft.h:66-69, BlinnPhongSoft.h:103-121, PbrSoft.h:110-127, plus the
 * directional-light Blinn-Phong (constant light vector) and per-vertex colour. */
__device__ __forceinline__ float4 runVS(const DrawParams &P, const float *a, float *v) {
  const Uniforms &u = P.u;
  float4 clip = m4p(u.mvp, a[0], a[1], a[2]);
  int sh = P.shading;
  if (sh == RHS_SHADING_BLINN_PHONG || sh == RHS_SHADING_BLINN_PHONG_DIR || sh == RHS_SHADING_PBR ||
      sh == RHS_SHADING_NEURAL_PBR) {
    bool bp = sh == RHS_SHADING_BLINN_PHONG || sh == RHS_SHADING_BLINN_PHONG_DIR;
    v[0] = a[4];
    v[1] = a[5];
    float4 wp4 = m4p(u.model, a[0], a[1], a[2]);
    f3 wp{wp4.x, wp4.y, wp4.z};
    f3 nrm{a[8], a[9], a[10]};
    f3 nv = m3v(u.model, 4, nrm);
    v[4] = nv.x, v[5] = nv.y, v[6] = nv.z;
    v[8] = wp.x, v[9] = wp.y, v[10] = wp.z;
    f3 cam{u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]};
    f3 lp{u.light_pos[0], u.light_pos[1], u.light_pos[2]};
    f3 cd = sub3(cam, wp);
    v[12] = cd.x, v[13] = cd.y, v[14] = cd.z;
    f3 ld = sh == RHS_SHADING_BLINN_PHONG_DIR ? lp : sub3(lp, wp);
    v[16] = ld.x, v[17] = ld.y, v[18] = ld.z;
    if (bp) {
      float4 s = m4p(u.shadow_mvp, a[0], a[1], a[2]);
      v[20] = s.x, v[21] = s.y, v[22] = s.z, v[23] = s.w;
    }
    if (P.defines & RHS_DEF_NORMAL_MAP) {
      f3 tg{a[12], a[13], a[14]};
      /* as compiled: x-first normalisations, T - dot(T, N) N fused */
      f3 N = nrm3b(m3v(u.inv_t, 4, nrm));
      f3 T = nrm3b(m3v(u.inv_t, 4, tg));
      const float d = dot3b(T, N);
      f3 Tn = nrm3b({fmaf(-d, N.x, T.x), fmaf(-d, N.y, T.y), fmaf(-d, N.z, T.z)});
      int o = bp ? 24 : 20;
      v[o] = N.x, v[o + 1] = N.y, v[o + 2] = N.z;
      v[o + 4] = Tn.x, v[o + 5] = Tn.y, v[o + 6] = Tn.z;
    }
  } else if (sh == RHS_SHADING_VERTEX_COLOR) {
    v[0] = a[4], v[1] = a[5], v[2] = a[6], v[3] = a[7];
  } else if (sh == RHS_SHADING_SKYBOX) {
    /* SkyboxSoft.h:66-75: depth at the far plane (w/w = 1), or 0 with reversed Z */
    clip = make_float4(clip.x, clip.y, clip.w, clip.w);
    if (u.reverse_z) clip.z = 0.f;
    v[0] = a[0], v[1] = a[1], v[2] = a[2];
  } else if (sh == RHS_SHADING_IBL_IRRADIANCE || sh == RHS_SHADING_IBL_PREFILTER) {
    /* IBLIrradianceSoft.h:62-67, IBLPrefilterSoft.h:66-71 */
    clip.z = clip.w;
    v[0] = a[0], v[1] = a[1], v[2] = a[2];
  } else if (sh == RHS_SHADING_FXAA) {
    /* FxaaSoft.h:72-75: the quad is already in clip space */
    clip = make_float4(a[0], a[1], a[2], 1.0f);
    v[0] = a[4], v[1] = a[5];
  }
  return clip;
}

/* countFrustumClipMask, RendererSoft.cpp:994-1003 */
__device__ __forceinline__ int clipMask(float4 c) {
  int m = 0;
  if (c.w < c.x) m |= 1;
  if (c.w < -c.x) m |= 2;
  if (c.w < c.y) m |= 4;
  if (c.w < -c.y) m |= 8;
  if (c.w < c.z) m |= 16;
  if (c.w < -c.z) m |= 32;
  return m;
}

/* perspectiveDivideImpl + viewportTransformImpl, RendererSoft.cpp:981-992:
 * multiply and add are separate roundings there. */
__device__ __forceinline__ float4 toScreen(const DrawParams &P, float4 c) {
  float iw = 1.f / c.w;
  float x = c.x * iw, y = c.y * iw, z = c.z * iw;
  return make_float4(x * P.p_x + P.o_x, y * P.p_y + P.o_y, z * 1.f + 0.f, iw * 1.f + 0.f);
}

/* --------------------------------------------------------------------------- fragment shaders */

/* The quad context a fragment shader needs: its own varyings and the texture
 * coordinates of quad pixels 0, 1, 2 for the coarse derivative. */
struct Frag {
  float v[kVaryStride]; /* own interpolated varyings (indexed like the VS output) */
  f2 tc[3];             /* v_texCoord of quad pixels p0, p1, p2 */
};

/* ShaderSoft::getSampler2DLod (ShaderSoft.h:117-133), used only when the sampler filters with mipmaps */
__device__ __forceinline__ float lodFor(const DrawParams &P, const TexDesc *t, const Frag &f) {
  if (!t || t->filter_min <= 1) return 0.f;
  float w = (float)t->width, h = (float)t->height;
  float dxx = (f.tc[1].x - f.tc[0].x) * w, dyx = (f.tc[2].x - f.tc[0].x) * w;
  float dxy = (f.tc[1].y - f.tc[0].y) * h, dyy = (f.tc[2].y - f.tc[0].y) * h;
  float ddx = fmaf(dxy, dxy, dxx * dxx);
  float ddy = fmaf(dyy, dyy, dyx * dyx);
  float d = ddy < ddx ? ddx : ddy;
  float lod = 0.5f * hlog2f(P.libm, d);
  return 0.f < lod ? lod : 0.f;
}

__device__ __forceinline__ RGBAf tex2D(const TexDesc *t, f2 uv, float lod) {
  if (!t) return {0, 0, 0, 0};
  return unpack01(textureImpl4(*t, 0, uv, lod));
}

__device__ __forceinline__ RGBAf texCube(const TexDesc *t, f3 dir, float lod) {
  if (!t) return {0, 0, 0, 0};
  int face;
  f2 uv;
  cubeFace(dir, &face, &uv);
  return unpack01(textureImpl4(*t, face, uv, lod));
}

/* The normal from the normal map's texel t4 and the interpolated TBN frame (BlinnPhongSoft.h:150-163,
 * PbrSoft.h:153-166) as the reference build compiles each: x-first normalisations and T - dot(T, N) N
 * fused in both; TBN * tn as fma(N, tn.z, fma(T, tn.x, B tn.y)) per component in the PBR shader, while
 * the Blinn-Phong shader computes x and y in a vector register as fma(N, tn.z, fma(B, tn.y, T tn.x)). */
template <bool PBR>
__device__ __forceinline__ f3 normalFromTexel(const Frag &f, int offN, int offT, RGBAf t4) {
  const float *v = f.v;
  f3 N = nrm3b({v[offN], v[offN + 1], v[offN + 2]});
  f3 T = nrm3b({v[offT], v[offT + 1], v[offT + 2]});
  const float d = dot3b(T, N);
  T = nrm3b({fmaf(-d, N.x, T.x), fmaf(-d, N.y, T.y), fmaf(-d, N.z, T.z)});
  f3 B = cross3(T, N);
  f3 tn{t4.r * 2.0f - 1.0f, t4.g * 2.0f - 1.0f, t4.b * 2.0f - 1.0f};
  f3 r;
  if (PBR)
    r = {fmaf(N.x, tn.z, fmaf(T.x, tn.x, B.x * tn.y)), fmaf(N.y, tn.z, fmaf(T.y, tn.x, B.y * tn.y)),
         fmaf(N.z, tn.z, fmaf(T.z, tn.x, B.z * tn.y))};
  else
    r = {fmaf(N.x, tn.z, fmaf(B.x, tn.y, T.x * tn.x)), fmaf(N.y, tn.z, fmaf(B.y, tn.y, T.y * tn.x)),
         fmaf(N.z, tn.z, fmaf(T.z, tn.x, B.z * tn.y))};
  return nrm3b(r);
}

template <bool PBR>
__device__ __forceinline__ f3 normalFromMap(const DrawParams &P, const Frag &f, int offN, int offT) {
  const float *v = f.v;
  if (P.defines & RHS_DEF_NORMAL_MAP) {
    const TexDesc *nt = P.tex[RHS_SLOT_NORMAL];
    return normalFromTexel<PBR>(f, offN, offT, tex2D(nt, {v[0], v[1]}, lodFor(P, nt, f)));
  }
  return nrm3b({v[4], v[5], v[6]});
}

/* BlinnPhongSoft.h:165-189 */
__device__ __forceinline__ float shadowCalc(const DrawParams &P, float4 sp, f3 n, f3 lightDir) {
  f3 pc{sp.x / sp.w, sp.y / sp.w, sp.z / sp.w};
  float cur = pc.z;
  if (cur < 0.f || cur > 1.f) return 0.0f;
  float bias = gmax(0.00025f * (1.0f - dot3b(n, nrm3b(lightDir))), 0.00005f); /* x-first, as compiled */
  const TexDesc *t = P.tex[RHS_SLOT_SHADOW];
  if (!t) return 0.f;
  float ox = 1.0f / (float)t->width, oy = 1.0f / (float)t->height;
  float s = 0.f;
  for (int x = -1; x <= 1; ++x)
    for (int y = -1; y <= 1; ++y) {
      float d = textureImpl<float>(*t, 0, {pc.x + (float)x * ox, pc.y + (float)y * oy}, 0.f);
      if (P.u.reverse_z)
        s += cur + bias < d ? 1.0f : 0.0f;
      else
        s += cur - bias > d ? 1.0f : 0.0f;
    }
  return s / 9.0f;
}

/* BlinnPhongSoft.h:191-242; `directional` = constant light vector, no range attenuation */
__device__ __forceinline__ RGBAf fsBlinnPhong(const DrawParams &P, const Frag &f, bool directional) {
  const float *v = f.v;
  const Uniforms &u = P.u;
  f2 uv{v[0], v[1]};
  const TexDesc *ta = P.tex[RHS_SLOT_ALBEDO];
  RGBAf base = (P.defines & RHS_DEF_ALBEDO_MAP)
                   ? tex2D(ta, uv, lodFor(P, ta, f))
                   : RGBAf{u.base_color[0], u.base_color[1], u.base_color[2], u.base_color[3]};
  f3 N = normalFromMap<false>(P, f, 24, 28);
  const TexDesc *tao = P.tex[RHS_SLOT_AO];
  float ao = (P.defines & RHS_DEF_AO_MAP) ? tex2D(tao, uv, lodFor(P, tao, f)).r : 1.f;
  f3 b3{base.r, base.g, base.b};
  f3 amb = scl3(mul3(b3, {u.ambient[0], u.ambient[1], u.ambient[2]}), ao);
  f3 dif{0, 0, 0}, spe{0, 0, 0}, emi{0, 0, 0};
  if (u.enable_light) {
    f3 ldv{v[16], v[17], v[18]};
    /* the reference build sums every dot product x first, and computes L + normalize(C) with the
     * normalisation's multiply fused into the add */
    f3 L = nrm3b(ldv);
    float diffuse = gmax(dot3b(N, L), 0.0f);
    f3 lc{u.light_color[0], u.light_color[1], u.light_color[2]};
    if (directional) {
      dif = scl3(mul3(lc, b3), diffuse);
    } else {
      f3 l = scl3(ldv, 1.0f / 5.f);
      float att = gclamp(1.0f - dot3b(l, l), 0.0f, 1.0f);
      dif = scl3(scl3(mul3(lc, b3), diffuse), att);
    }
    const f3 cr{v[12], v[13], v[14]};
    const float ic = 1.0f / sqrtf(dot3b(cr, cr));
    f3 H = nrm3b({fmaf(cr.x, ic, L.x), fmaf(cr.y, ic, L.y), fmaf(cr.z, ic, L.z)});
    float sa = gmax(dot3b(N, H), 0.0f);
    float sp = u.k_specular * hpowi<128>(P.libm, sa);
    spe = {sp, sp, sp};
    if (u.enable_shadow) {
      float s = 1.0f - shadowCalc(P, make_float4(v[20], v[21], v[22], v[23]), N, ldv);
      dif = scl3(dif, s);
      spe = scl3(spe, s);
    }
  }
  if (P.defines & RHS_DEF_EMISSIVE_MAP) {
    const TexDesc *te = P.tex[RHS_SLOT_EMISSIVE];
    RGBAf e = tex2D(te, uv, lodFor(P, te, f));
    emi = {e.r, e.g, e.b};
  }
  f3 col = add3(add3(add3(amb, dif), spe), emi);
  return {col.x, col.y, col.z, base.a};
}

/* PbrSoft.h:168-223 */
constexpr float kPI = 3.14159265359f;
__device__ __forceinline__ float distGGX(f3 N, f3 H, float r) {
  float a = r * r, a2 = a * a;
  float nh = gmax(dot3b(N, H), 0.0f);
  float nh2 = nh * nh;
  float den = fmaf(a2 - 1.0f, nh2, 1.0f); /* the reference build fuses NdotH2 * (a2 - 1) + 1 */
  den = kPI * den * den;
  return a2 / den;
}
__device__ __forceinline__ float schlickGGX(float nv, float r) {
  float k = ((r + 1.0f) * (r + 1.0f)) / 8.0f;
  return nv / (nv * (1.0f - k) + k);
}
/* every multiply-add of EnvBRDFApprox is fused in the reference build */
__device__ __forceinline__ f3 envBRDFApprox(const DrawParams &P, f3 spec, float r, float nv) {
  float rx = fmaf(r, -1.f, 1.f), ry = fmaf(r, -0.0275f, 0.0425f), rz = fmaf(r, -0.572f, 1.04f),
        rw = fmaf(r, 0.022f, -0.04f);
  float a004 = fmaf(gmin(rx * rx, hexp2f(P.libm, -9.28f * nv)), rx, ry);
  float ax = fmaf(-1.04f, a004, rz), ay = fmaf(1.04f, a004, rw);
  ay *= gmax(0.f, gmin(1.f, 50.0f * spec.y));
  return {fmaf(spec.x, ax, ay), fmaf(spec.y, ax, ay), fmaf(spec.z, ax, ay)};
}

/* powf(x, 2.2) of a channel sampled from an RGBA8 texture: x is k / 255 for a
 * byte k, and the host libm's values of all 256 are in the gamma table. */
__device__ __forceinline__ float gammaIn(const DrawParams &P, float x) {
  return P.gamma_lut[(int)(x * 255.f + 0.5f)];
}

/* The byte the two-phase raster writes for a PBR colour channel c without an
 * emissive map: u8(clamp(powf(c, 1/2.2), 0, 1) * 255), which the host libm makes
 * monotonic in c (checked over every float of (0, 1] when the table is built).
 * The host found the smallest c giving each byte k (gammaTablesHost); a
 * hardware log2/exp2 estimate picks the starting byte and the thresholds
 * correct it. Returned as a float that truncates back to k. */
__device__ __forceinline__ float gammaOutByte(const DrawParams &P, float c) {
  /* powf(c, 1/2.2) off (0, inf): +0 at zero (either sign), NaN below zero and for NaN, inf at inf */
  if (!(c > 0.f)) return c == 0.f ? 0.f : __builtin_nanf("");
  if (!(c < INFINITY)) return INFINITY;
  const float *T = P.gamma_lut + 256;
  const float e = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(c) * (1.0f / 2.2f));
  int k = (int)(gmin(e, 1.f) * 255.f);
  k = k < 0 ? 0 : (k > 255 ? 255 : k);
  while (k < 255 && c >= T[k + 1]) k++;
  while (k > 0 && c < T[k]) k--;
  return ((float)k + 0.5f) * (1.f / 255.f);
}

/* Material inputs of the PBR shader, shared with the learned shader. */
struct PbrInputs {
  RGBAf albedo_rgba;
  f3 albedo;
  float metallic, rough, ao;
  f3 N, V, R;
  float dnv; /* dot(N, V) as the reference build computes it: -dot(N, -V) from the same products */
};

__device__ __forceinline__ PbrInputs pbrInputs(const DrawParams &P, const Frag &f) {
  const float *v = f.v;
  const Uniforms &u = P.u;
  PbrInputs in;
  f2 uv{v[0], v[1]};
  const TexDesc *ta = P.tex[RHS_SLOT_ALBEDO], *tn = P.tex[RHS_SLOT_NORMAL];
  /* albedo and normal maps of one shape share their LOD, levels, weights and texel indices */
  const bool pair = (P.defines & RHS_DEF_ALBEDO_MAP) && (P.defines & RHS_DEF_NORMAL_MAP) && ta && tn &&
                    sameShape(*ta, *tn);
  RGBAf ntexel{0.f, 0.f, 0.f, 0.f};
  if (pair) {
    float4 ba, bn;
    textureImplPair(*ta, *tn, 0, uv, lodFor(P, ta, f), ba, bn);
    in.albedo_rgba = unpack01(ba);
    ntexel = unpack01(bn);
    /* gammaIn's table index is the byte itself */
    in.albedo = {P.gamma_lut[(int)ba.x], P.gamma_lut[(int)ba.y], P.gamma_lut[(int)ba.z]};
  } else if (P.defines & RHS_DEF_ALBEDO_MAP) {
    in.albedo_rgba = tex2D(ta, uv, lodFor(P, ta, f));
    in.albedo = {gammaIn(P, in.albedo_rgba.r), gammaIn(P, in.albedo_rgba.g), gammaIn(P, in.albedo_rgba.b)};
  } else {
    /* powf(base colour, 2.2), by the host libm once per draw */
    in.albedo_rgba = RGBAf{u.base_color[0], u.base_color[1], u.base_color[2], u.base_color[3]};
    in.albedo = {P.base_lin[0], P.base_lin[1], P.base_lin[2]};
  }
  in.metallic = 0.0f;
  in.rough = 1.0f;
  if (P.defines & RHS_DEF_METALROUGHNESS_MAP) {
    /* this sampler never gets a LOD callback (PbrSoft.h:138-151): level-0 lookups */
    RGBAf mr = tex2D(P.tex[RHS_SLOT_METAL_ROUGHNESS], uv, 0.f);
    in.metallic = mr.b;
    in.rough = mr.g;
  }
  const TexDesc *tao = P.tex[RHS_SLOT_AO];
  in.ao = (P.defines & RHS_DEF_AO_MAP) ? tex2D(tao, uv, lodFor(P, tao, f)).r : 1.f;
  in.N = pair ? normalFromTexel<true>(f, 20, 24, ntexel) : normalFromMap<true>(P, f, 20, 24);
  in.V = nrm3b({v[12], v[13], v[14]});
  /* glm::reflect(-V, N); the reference build rounds the products of dot(N, I) on their own and reuses
   * them, negated, for dot(N, V) */
  f3 I{-in.V.x, -in.V.y, -in.V.z};
  float dNI = dot3s(in.N, I);
  in.dnv = -dNI;
  in.R = sub3(I, scl3(in.N, dNI * 2.0f));
  return in;
}

/* PbrSoft.h:225-322. U8OUT: the caller only truncates the colour to bytes with
 * nothing blended in between, so the gamma step may return byte representatives. */
template <bool U8OUT = false>
__device__ __forceinline__ RGBAf fsPbr(const DrawParams &P, const Frag &f) {
  const float *v = f.v;
  const Uniforms &u = P.u;
  PbrInputs in = pbrInputs(P, f);
  f3 albedo = in.albedo, N = in.N, V = in.V;
  float metallic = in.metallic, rough = in.rough;
  f3 F0{mixF(0.04f, albedo.x, metallic), mixF(0.04f, albedo.y, metallic), mixF(0.04f, albedo.z, metallic)};
  f3 Lo{0, 0, 0};
  if (u.enable_light) {
    f3 ldv{v[16], v[17], v[18]};
    f3 L = nrm3b(ldv);
    f3 H = nrm3b(add3(V, L));
    f3 ld = scl3(ldv, 1.0f / 5.f);
    float att = gclamp(1.0f - dot3b(ld, ld), 0.0f, 1.0f);
    f3 radiance = scl3({u.light_color[0], u.light_color[1], u.light_color[2]}, att);
    float NDF = distGGX(N, H, rough);
    float nv = gmax(in.dnv, 0.0f), nl = gmax(dot3b(N, L), 0.0f);
    float G = schlickGGX(nl, rough) * schlickGGX(nv, rough);
    float ct = gmax(dot3b(H, V), 0.0f);
    float p5 = hpowi<5>(P.libm, gclamp(1.0f - ct, 0.0f, 1.0f));
    f3 F{fmaf(1.0f - F0.x, p5, F0.x), fmaf(1.0f - F0.y, p5, F0.y), fmaf(1.0f - F0.z, p5, F0.z)};
    float ng = NDF * G;
    f3 num = scl3(F, ng);
    float den = fmaf(4.0f * nv, nl, 0.0001f);
    f3 spec{num.x / den, num.y / den, num.z / den};
    f3 kD{(1.0f - F.x) * (1.0f - metallic), (1.0f - F.y) * (1.0f - metallic), (1.0f - F.z) * (1.0f - metallic)};
    f3 t{kD.x * albedo.x / kPI + spec.x, kD.y * albedo.y / kPI + spec.y, kD.z * albedo.z / kPI + spec.z};
    Lo = scl3(mul3(t, radiance), nl);
  }
  f3 ambient;
  if (u.enable_ibl) {
    float nv = gmax(in.dnv, 0.0f);
    float p5 = hpowi<5>(P.libm, gclamp(1.0f - nv, 0.0f, 1.0f));
    float omr = 1.0f - rough;
    f3 F{fmaf(gmax(omr, F0.x) - F0.x, p5, F0.x), fmaf(gmax(omr, F0.y) - F0.y, p5, F0.y),
         fmaf(gmax(omr, F0.z) - F0.z, p5, F0.z)};
    f3 kD{(1.0f - F.x) * (1.0f - metallic), (1.0f - F.y) * (1.0f - metallic), (1.0f - F.z) * (1.0f - metallic)};
    RGBAf irr = texCube(P.tex[RHS_SLOT_IRRADIANCE], N, 0.f);
    f3 diffuse{irr.r * albedo.x, irr.g * albedo.y, irr.b * albedo.z};
    RGBAf pf = texCube(P.tex[RHS_SLOT_PREFILTER], in.R, rough * 4.0f);
    f3 brdf = envBRDFApprox(P, F, rough, nv);
    /* (kD * diffuse + prefiltered * brdf) with the second product fused into the add */
    f3 pre{fmaf(brdf.x, pf.r, kD.x * diffuse.x), fmaf(brdf.y, pf.g, kD.y * diffuse.y),
           fmaf(brdf.z, pf.b, kD.z * diffuse.z)};
    ambient = scl3(pre, in.ao);
  } else {
    ambient = scl3(mul3({u.ambient[0], u.ambient[1], u.ambient[2]}, albedo), in.ao);
  }
  f3 col = add3(ambient, Lo);
  if (U8OUT && !(P.defines & RHS_DEF_EMISSIVE_MAP))
    col = {gammaOutByte(P, col.x), gammaOutByte(P, col.y), gammaOutByte(P, col.z)};
  else
    col = {hpowGamma(P.libm, col.x), hpowGamma(P.libm, col.y), hpowGamma(P.libm, col.z)};
  if (P.defines & RHS_DEF_EMISSIVE_MAP) {
    const TexDesc *te = P.tex[RHS_SLOT_EMISSIVE];
    RGBAf e = tex2D(te, {v[0], v[1]}, lodFor(P, te, f));
    col = add3(col, {e.r, e.g, e.b});
  }
  return {col.x, col.y, col.z, in.albedo_rgba.a};
}

/* --------------------------------------------------------------------------- environment shaders */

/* textureLodOffset (ShaderSoft.h:105-110, SamplerSoft.h:216-238): the texel
 * offset is added after scaling to the level (bilinear: to the texel-space
 * coordinate; nearest: to the floored texel). The reference build fuses the
 * bilinear `uv * size + offset` into one multiply-add: a coordinate that
 * rounds to a texel centre unshifted can land just past it shifted. */
__device__ __forceinline__ uint32_t sampleLevelOff(const TexDesc &d, int level, f2 uv, bool linear, int ox, int oy) {
  const TexLevel &l = d.lv[0][level];
  if (linear)
    return bilinear<uint32_t>(d, l, fmaf(uv.x, (float)l.w, (float)ox), fmaf(uv.y, (float)l.h, (float)oy), d.wrap);
  const float tx = uv.x * (float)l.w, ty = uv.y * (float)l.h;
  return fetchWrap<uint32_t>(d, l, (int)floorf(tx) + ox, (int)floorf(ty) + oy, d.wrap);
}
__device__ RGBAf tex2DOffset(const TexDesc *t, f2 uv, float lod, int ox, int oy) {
  if (!t) return {0, 0, 0, 0};
  const TexDesc &d = *t;
  const int fm = d.filter_min;
  if (fm <= 1) return unpack01(sampleLevelOff(d, 0, uv, fm == 1, ox, oy));
  const int maxLevel = d.levels - 1;
  if (fm == 2 || fm == 3) {
    const int level = min(max((int)ceilf(lod + 0.5f) - 1, 0), maxLevel);
    return unpack01(sampleLevelOff(d, level, uv, fm == 3, ox, oy));
  }
  const int hi = min(max((int)floorf(lod), 0), maxLevel), lo = min(max(hi + 1, 0), maxLevel);
  const uint32_t th = sampleLevelOff(d, hi, uv, fm == 5, ox, oy);
  if (hi == lo) return unpack01(th);
  return unpack01(mixRGBA(th, sampleLevelOff(d, lo, uv, fm == 5, ox, oy), lod - floorf(lod)));
}

/* glm::dot and glm::normalize as the reference build compiles them in the
 * skybox, prefilter and FXAA shaders: fma(z, z', fma(y, y', x x')), x first (the
 * lighting shaders' order, dot3, starts from y). */

/* SkyboxSoft.h:78-96. The equirectangular lookup calls glibc's atan2f
 * (reproduced step for step, fdAtan2f) and the double-precision asin, rounded
 * to float; the reference build fuses `uv * scale + 0.5`. Only the Blinn-Phong
 * and PBR shaders give their samplers a derivative LOD (BlinnPhongSoft.h:135-148,
 * PbrSoft.h:138-151): the skybox's and FXAA's samplers read at LOD 0. */
__device__ __forceinline__ RGBAf fsSkybox(const DrawParams &P, const Frag &f) {
  const f3 wp{f.v[0], f.v[1], f.v[2]};
  if (P.defines & RHS_DEF_EQUIRECTANGULAR_MAP) {
    const f3 d = nrm3b(wp);
    const float ax = fdAtan2f(d.z, d.x), ay = (float)asin((double)-d.y);
    const f2 uv{fmaf(ax, 0.1591f, 0.5f), fmaf(ay, 0.3183f, 0.5f)};
    return tex2D(P.tex[RHS_SLOT_EQUIRECT], uv, 0.f);
  }
  return texCube(P.tex[RHS_SLOT_CUBE], wp, 0.f);
}

/* IBLIrradianceSoft.h:74-101: the cosine-weighted hemisphere sum in the
 * reference's loop order; sin and cos of the loop angles come from the host
 * libm (IblTables). The sum's terms (irrTerm) are independent; only their
 * addition is ordered, so a caller may compute them in parallel and add them
 * in order (k_ord_shade_ibl). */
struct IrrSetup {
  f3 N, right, up;
};
__device__ __forceinline__ IrrSetup irrSetup(const Frag &f) {
  IrrSetup s; /* x-first normalisations, as compiled */
  s.N = nrm3b({f.v[0], f.v[1], f.v[2]});
  f3 up{0.f, 1.f, 0.f};
  s.right = nrm3b(cross3(up, s.N));
  s.up = nrm3b(cross3(s.N, s.right));
  return s;
}
/* Term k of the sum (phi index k / ntheta, theta index k % ntheta): the sample's
 * colour times cos(theta), and sin(theta). The reference build fuses the last
 * product into the running sum (irrAdd), and the sample vector into two
 * multiply-adds. */
__device__ __forceinline__ float4 irrTerm(const DrawParams &P, const IrrSetup &s, int k) {
  const IblTables &T = *P.ibl;
  const int i = k / T.ntheta, j = k - i * T.ntheta;
  const float sp = T.sphi[i], cp = T.cphi[i], st = T.stheta[j], ct = T.ctheta[j];
  const f3 ts{st * cp, st * sp, ct};
  const f3 sv{fmaf(s.N.x, ts.z, fmaf(s.up.x, ts.y, s.right.x * ts.x)),
              fmaf(s.N.y, ts.z, fmaf(s.up.y, ts.y, s.right.y * ts.x)),
              fmaf(s.N.z, ts.z, fmaf(s.up.z, ts.y, s.right.z * ts.x))};
  const RGBAf c = texCube(P.tex[RHS_SLOT_CUBE], sv, 0.f);
  return make_float4(c.r * ct, c.g * ct, c.b * ct, st);
}
__device__ __forceinline__ f3 irrAdd(f3 irr, float4 t) {
  return {fmaf(t.x, t.w, irr.x), fmaf(t.y, t.w, irr.y), fmaf(t.z, t.w, irr.z)};
}
__device__ __forceinline__ RGBAf irrFinish(f3 irr, float nr) {
  irr = scl3(scl3(irr, kPI), 1.0f / nr);
  return {irr.x, irr.y, irr.z, 1.f};
}
__device__ RGBAf fsIrradiance(const DrawParams &P, const Frag &f) {
  const IrrSetup s = irrSetup(f);
  const int n = P.ibl->nphi * P.ibl->ntheta;
  f3 irr{0.f, 0.f, 0.f};
  float nr = 0.f;
  for (int k = 0; k < n; k++) {
    irr = irrAdd(irr, irrTerm(P, s, k));
    nr += 1.f;
  }
  return irrFinish(irr, nr);
}

/* IBLPrefilterSoft.h:88-168: GGX importance sampling over 1024 Hammersley
 * points, each sample read from the source mip level its pdf selects. The
 * reference build contracts (a^2 - 1) Xi.y + 1, 1 - cos^2, the tangent-space
 * sum, 2 (V.H) H - V, the GGX denominator, 1024 pdf + 1e-4 and the colour
 * sum into fused multiply-adds; so does this one, term for term. As for the
 * irradiance, the samples are independent and only their sum is ordered. */
struct PrefSetup {
  f3 N, V, tangent, bitangent;
  float rough, ra2, saTexel;
};
__device__ __forceinline__ PrefSetup prefSetup(const DrawParams &P, const Frag &f) {
  PrefSetup s;
  s.N = nrm3b({f.v[0], f.v[1], f.v[2]});
  s.V = s.N;
  s.rough = P.u.aux[1];
  const float res = P.u.aux[0];
  const float a = s.rough * s.rough;
  s.ra2 = a * a;
  const f3 upv = (double)fabsf(s.N.z) < 0.999 ? f3{0.f, 0.f, 1.f} : f3{1.f, 0.f, 0.f};
  s.tangent = nrm3b(cross3(upv, s.N));
  s.bitangent = cross3(s.N, s.tangent);
  s.saTexel = 4.0f * kPI / (6.0f * res * res);
  return s;
}
/* Sample i: the source texel's colour and N.L (the sample counts only when N.L > 0). */
__device__ __forceinline__ float4 prefSample(const DrawParams &P, const PrefSetup &s, uint32_t i) {
  const IblTables &T = *P.ibl;
  const f3 N = s.N, V = s.V, tangent = s.tangent, bitangent = s.bitangent;
  const float ra2 = s.ra2;
  const float xiy = (float)((double)__builtin_bitreverse32(i) * 2.3283064365386963e-10);
  const float cosT = sqrtf((1.0f - xiy) / fmaf(ra2 - 1.0f, xiy, 1.0f));
  const float sinT = sqrtf(fmaf(-cosT, cosT, 1.0f));
  const f3 Ht{T.hcos[i] * sinT, T.hsin[i] * sinT, cosT};
  const f3 H = nrm3b({fmaf(N.x, Ht.z, fmaf(bitangent.x, Ht.y, tangent.x * Ht.x)),
                      fmaf(N.y, Ht.z, fmaf(bitangent.y, Ht.y, tangent.y * Ht.x)),
                      fmaf(N.z, Ht.z, fmaf(bitangent.z, Ht.y, tangent.z * Ht.x))});
  const float vh2 = 2.0f * dot3b(V, H);
  const f3 L = nrm3b({fmaf(H.x, vh2, -V.x), fmaf(H.y, vh2, -V.y), fmaf(H.z, vh2, -V.z)});
  const float NdotL = gmax(dot3b(N, L), 0.0f);
  if (!(NdotL > 0.0f)) return make_float4(0.f, 0.f, 0.f, 0.f);
  const float NdotH = gmax(dot3b(N, H), 0.0f), HdotV = gmax(dot3b(H, V), 0.0f);
  const float den = fmaf(NdotH * NdotH, ra2 - 1.0f, 1.0f);
  const float D = ra2 / (kPI * den * den);
  const float pdf = D * NdotH / (4.0f * HdotV) + 0.0001f;
  const float saSample = 1.0f / fmaf(1024.0f, pdf, 0.0001f);
  const float mip = s.rough == 0.0f ? 0.0f : 0.5f * hlog2f(P.libm, saSample / s.saTexel);
  const RGBAf c = texCube(P.tex[RHS_SLOT_CUBE], L, mip);
  return make_float4(c.r, c.g, c.b, NdotL);
}
__device__ RGBAf fsPrefilter(const DrawParams &P, const Frag &f) {
  const PrefSetup s = prefSetup(P, f);
  f3 acc{0.f, 0.f, 0.f};
  float tw = 0.f;
  for (uint32_t i = 0; i < 1024u; i++) {
    const float4 t = prefSample(P, s, i);
    if (t.w > 0.0f) {
      acc = {fmaf(t.x, t.w, acc.x), fmaf(t.y, t.w, acc.y), fmaf(t.z, t.w, acc.z)};
      tw += t.w;
    }
  }
  return {acc.x / tw, acc.y / tw, acc.z / tw, 1.f};
}

/* FxaaSoft.h:81-266. The reference build contracts the luma dot product (red
 * first), -2 l + corners, |.| 2 + |.|, 2 (DU + LR) + LC, (average - centre)
 * and the sub-pixel smoothstep into fused multiply-adds; so does this one. */
__device__ __forceinline__ float lumaOf(RGBAf c) { return dot3b({c.r, c.g, c.b}, {0.299f, 0.587f, 0.114f}); }
__device__ __forceinline__ float fxaaQuality(float q) {
  return q < 5.f ? 1.0f : (q > 5.f ? (q < 10.f ? 2.0f : (q < 11.f ? 4.0f : 8.0f)) : 1.5f);
}
__device__ RGBAf fsFxaa(const DrawParams &P, const Frag &f) {
  const TexDesc *t = P.tex[RHS_SLOT_SCREEN];
  const f2 tc{f.v[0], f.v[1]};
  const float lod = 0.f;
  const f2 inv{1.0f / P.u.aux[0], 1.0f / P.u.aux[1]};
  const RGBAf cc = tex2D(t, tc, lod);
  const float lC = lumaOf(cc);
  const float lD = lumaOf(tex2DOffset(t, tc, 0.f, 0, -1)), lU = lumaOf(tex2DOffset(t, tc, 0.f, 0, 1));
  const float lL = lumaOf(tex2DOffset(t, tc, 0.f, -1, 0)), lR = lumaOf(tex2DOffset(t, tc, 0.f, 1, 0));
  const float lMin = gmin(lC, gmin(gmin(lD, lU), gmin(lL, lR)));
  const float lMax = gmax(lC, gmax(gmax(lD, lU), gmax(lL, lR)));
  const float range = lMax - lMin;
  if (range < gmax(0.0312f, lMax * 0.125f)) return {cc.r, cc.g, cc.b, 1.f};
  const float lDL = lumaOf(tex2DOffset(t, tc, 0.f, -1, -1)), lUR = lumaOf(tex2DOffset(t, tc, 0.f, 1, 1));
  const float lUL = lumaOf(tex2DOffset(t, tc, 0.f, -1, 1)), lDR = lumaOf(tex2DOffset(t, tc, 0.f, 1, -1));
  const float lDU = lD + lU, lLR = lL + lR;
  const float lLC = lDL + lUL, lDC = lDL + lDR, lRC = lDR + lUR, lUC = lUR + lUL;
  const float eH = fmaf(fabsf(fmaf(-2.0f, lC, lDU)), 2.0f, fabsf(fmaf(-2.0f, lL, lLC))) + fabsf(fmaf(-2.0f, lR, lRC));
  const float eV = fmaf(fabsf(fmaf(-2.0f, lC, lLR)), 2.0f, fabsf(fmaf(-2.0f, lU, lUC))) + fabsf(fmaf(-2.0f, lD, lDC));
  const bool hor = eH >= eV;
  float step = hor ? inv.y : inv.x;
  const float l1 = hor ? lD : lL, l2 = hor ? lU : lR;
  const float g1 = l1 - lC, g2 = l2 - lC;
  const bool steep1 = fabsf(g1) >= fabsf(g2);
  const float gs = 0.25f * gmax(fabsf(g1), fabsf(g2));
  float avg;
  if (steep1) {
    step = -step;
    avg = 0.5f * (l1 + lC);
  } else {
    avg = 0.5f * (l2 + lC);
  }
  f2 cur = tc;
  if (hor)
    cur.y += step * 0.5f;
  else
    cur.x += step * 0.5f;
  const f2 off = hor ? f2{inv.x, 0.0f} : f2{0.0f, inv.y};
  f2 uv1{cur.x - off.x * fxaaQuality(0.f), cur.y - off.y * fxaaQuality(0.f)};
  f2 uv2{cur.x + off.x * fxaaQuality(0.f), cur.y + off.y * fxaaQuality(0.f)};
  float e1 = 0.f, e2 = 0.f;
  bool r1 = false, r2 = false;
  for (int i = 1; i < 12; i++) {
    if (!r1) {
      e1 = lumaOf(tex2D(t, uv1, lod)) - avg;
      r1 = fabsf(e1) >= gs;
    }
    if (!r2) {
      e2 = lumaOf(tex2D(t, uv2, lod)) - avg;
      r2 = fabsf(e2) >= gs;
    }
    const bool both = r1 && r2;
    const float q = fxaaQuality((float)i);
    if (!r1) uv1 = {uv1.x - off.x * q, uv1.y - off.y * q};
    if (!r2) uv2 = {uv2.x + off.x * q, uv2.y + off.y * q};
    if (both) break;
  }
  const float d1 = hor ? (tc.x - uv1.x) : (tc.y - uv1.y);
  const float d2 = hor ? (uv2.x - tc.x) : (uv2.y - tc.y);
  const bool dir1 = d1 < d2;
  const float dmin = gmin(d1, d2);
  const bool smaller = lC < avg;
  const bool ok1 = (e1 < 0.0f) != smaller, ok2 = (e2 < 0.0f) != smaller;
  const bool ok = dir1 ? ok1 : ok2;
  const float edgeLen = d1 + d2;
  const float pxOff = -dmin / edgeLen + 0.5f;
  float fin = ok ? pxOff : 0.0f;
  const float lSum = fmaf(2.0f, lDU + lLR, lLC) + lRC;
  const float sp1 = gclamp(fabsf(fmaf(lSum, 1.0f / 12.0f, -lC)) / range, 0.0f, 1.0f);
  const float sp2 = fmaf(-2.0f, sp1, 3.0f) * sp1 * sp1;
  const float spF = sp2 * sp2 * 0.75f;
  fin = gmax(fin, spF);
  f2 fuv = tc;
  if (hor)
    fuv.y += fin * step;
  else
    fuv.x += fin * step;
  const RGBAf fc = tex2D(t, fuv, lod);
  return {fc.r, fc.g, fc.b, 1.f};
}

/* The environment shaders, one specialisation switching on the draw's model. */
__device__ __forceinline__ RGBAf fsEnv(const DrawParams &P, const Frag &f) {
  switch (P.shading) {
    case RHS_SHADING_SKYBOX: return fsSkybox(P, f);
    case RHS_SHADING_IBL_IRRADIANCE: return fsIrradiance(P, f);
    case RHS_SHADING_IBL_PREFILTER: return fsPrefilter(P, f);
    default: return fsFxaa(P, f);
  }
}

/* --------------------------------------------------------------------------- per-sample operations */

/* DepthSoft.h:13-25 */
__device__ __forceinline__ bool depthPass(float a, float b, int fn) {
  const float eps = 1.19209290e-07f;
  switch (fn) {
    case 0: return false;
    case 1: return a < b;
    case 2: return fabsf(a - b) <= eps;
    case 3: return a <= b;
    case 4: return a > b;
    case 5: return fabsf(a - b) > eps;
    case 6: return a >= b;
    default: return true;
  }
}

/* BlendSoft.h:13-56 */
__device__ __forceinline__ float blendFactor(int f, float s, float sa, float d, float da) {
  switch (f) {
    case 0: return 0.f;
    case 1: return 1.f;
    case 2: return s;
    case 3: return sa;
    case 4: return d;
    case 5: return da;
    case 6: return 1.f - s;
    case 7: return 1.f - sa;
    case 8: return 1.f - d;
    case 9: return 1.f - da;
  }
  return 0.f;
}
__device__ __forceinline__ float blendFunc(int fn, float s, float d) {
  switch (fn) {
    case 1: return s - d;
    case 2: return d - s;
    case 3: return gmin(s, d);
    case 4: return gmax(s, d);
    default: return s + d;
  }
}
__device__ __forceinline__ RGBAf blendColor(RGBAf s, RGBAf d, const RhsStates &st) {
  float sc[3] = {s.r, s.g, s.b}, dc[3] = {d.r, d.g, d.b}, r[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float fs = blendFactor(st.blendSrcRgb, sc[i], s.a, dc[i], d.a);
    float fd = blendFactor(st.blendDstRgb, sc[i], s.a, dc[i], d.a);
    r[i] = blendFunc(st.blendFuncRgb, sc[i] * fs, dc[i] * fd);
  }
  float fsa = blendFactor(st.blendSrcAlpha, s.a, s.a, d.a, d.a);
  float fda = blendFactor(st.blendDstAlpha, s.a, s.a, d.a, d.a);
  return {r[0], r[1], r[2], blendFunc(st.blendFuncAlpha, s.a * fsa, d.a * fda)};
}

}  // namespace rhip
