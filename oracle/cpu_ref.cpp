/*
 * cpu_ref.cpp -- TEST INFRASTRUCTURE ONLY: the CPU restatement of SoftGLRender's
 * RendererSoft triangle path, used as the checker for RendererHIP. The product
 * never links or loads it.
 *
 * It restates the reference's algorithm step by step, with the reference's
 * floating-point recipes where they decide coverage, depth or texel values
 * (FMA contraction as emitted by the reference's own build flags
 * -O3 -mavx2 -mfma, and the dpps/vector orders noted below). This file is
 * compiled with -ffp-contract=off so every fused operation is an explicit fmaf.
 *
 * Citations are path:line in SoftGLRender (src/...).
 */
#include "cpu_ref.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "rhip_stream.h"

namespace cref {

static int g_threads = 1;
static int g_trace_x = -1, g_trace_y = -1; /* debugging aid: CPU_REF_TRACE=x,y */

/* The reference calls the host libm's powf / log2f / exp2f (glibc: within ~0.8 ulp,
 * not always correctly rounded). With g_cr_math set, the checker evaluates them in
 * double precision and rounds once, i.e. correctly rounded in all but vanishingly
 * rare cases; that is the mode the GPU path must match bit for bit. */
static bool g_cr_math = false;
static inline float m_powf(float x, float y) { return g_cr_math ? (float)pow((double)x, (double)y) : powf(x, y); }
static inline float m_log2f(float x) { return g_cr_math ? (float)log2((double)x) : log2f(x); }
static inline float m_exp2f(float x) { return g_cr_math ? (float)exp2((double)x) : exp2f(x); }
static inline float m_sinf(float x) { return g_cr_math ? (float)sin((double)x) : sinf(x); }
static inline float m_cosf(float x) { return g_cr_math ? (float)cos((double)x) : cosf(x); }
static inline float m_atan2f(float y, float x) { return g_cr_math ? (float)atan2((double)y, (double)x) : atan2f(y, x); }
static long long g_stats[8];
/* unique-texel counting: while on (single-threaded raster), every texel a draw's shaders read is marked, and
 * the draw's count of distinct texels read, times 4 bytes, is appended to g_draw_texel_bytes */
static bool g_touch = false;
static std::vector<long long> g_draw_texel_bytes;

/* ------------------------------------------------------------------ small vector helpers */

struct V2 {
  float x, y;
};
struct V3 {
  float x, y, z;
};
struct V4 {
  float x, y, z, w;
};

static inline V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 add3(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 mul3(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 scale3(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
/* glm::dot(vec3) with the reference flags: fma(z, z', fma(x, x', y*y')). */
static inline float dot3(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.x, b.x, a.y * b.y)); }
static inline V3 normalize3(V3 v) {
  float inv = 1.0f / sqrtf(dot3(v, v)); /* glm::normalize = v * inversesqrt(dot(v, v)) */
  return scale3(v, inv);
}
/* glm::dot / glm::normalize as the reference build compiles them in the PBR, skybox, prefilter and FXAA
 * shaders: fma(z, z', fma(y, y', x x')) */
static inline float dot3b(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
static inline V3 normalize3b(V3 v) { return scale3(v, 1.0f / sqrtf(dot3b(v, v))); }
/* glm::dot where the reference build rounds each product on its own (it shares them between two dot
 * products): (x x' + y y') + z z' */
static inline float dot3s(V3 a, V3 b) {
  const float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
  return (px + py) + pz;
}
static inline V3 cross3(V3 a, V3 b) {
  return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}
static inline float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }
static inline float maxf(float a, float b) { return a < b ? b : a; } /* glm::max */
static inline float minf(float a, float b) { return b < a ? b : a; } /* glm::min */

/* glm::mat4 * vec4(p, 1) under the reference flags: (fma(c0, x, c1*y)) + (fma(c2, z, c3)),
 * glm/detail/type_mat4x4.inl:536-575 (Mul0+Mul1, Mul2+Mul3, Add0+Add1). */
static inline V4 m4p(const float *m, float x, float y, float z) {
  V4 r;
  float *o = &r.x;
  for (int i = 0; i < 4; i++) {
    float a0 = fmaf(m[i], x, m[4 + i] * y);
    float a1 = fmaf(m[8 + i], z, m[12 + i]);
    o[i] = a0 + a1;
  }
  return r;
}

/* glm::mat3 * vec3: fma(c2, z, fma(c0, x, c1*y)). `m` is 3 columns of stride `cs`. */
static inline V3 m3v(const float *m, int cs, V3 v) {
  V3 r;
  float *o = &r.x;
  for (int i = 0; i < 3; i++) o[i] = fmaf(m[2 * cs + i], v.z, fmaf(m[i], v.x, m[cs + i] * v.y));
  return r;
}

/* ------------------------------------------------------------------ uniforms (Material.h:63-95) */

struct Uniforms {
  int reverseZ;
  float model[16], mvp[16], invT[12], shadowMvp[16];
  V3 ambient, camPos, lightPos, lightColor;
  int enableLight, enableIBL, enableShadow;
  float pointSize, kSpecular;
  V4 baseColor;
  float aux[4]; /* UniformsIBLPrefilter {srcResolution, roughness} or UniformsQuadFilter {screenSize} */
};

static void parseUniforms(const RhsDraw &d, Uniforms &u) {
  memcpy(&u.reverseZ, d.model, 4);
  memcpy(u.model, d.model + 16, 64);
  memcpy(u.mvp, d.model + 80, 64);
  memcpy(u.invT, d.model + 144, 48);
  memcpy(u.shadowMvp, d.model + 192, 64);
  memcpy(&u.ambient, d.scene + 0, 12);
  memcpy(&u.camPos, d.scene + 16, 12);
  memcpy(&u.lightPos, d.scene + 32, 12);
  memcpy(&u.lightColor, d.scene + 48, 12);
  memcpy(&u.enableLight, d.material + 0, 4);
  memcpy(&u.enableIBL, d.material + 4, 4);
  memcpy(&u.enableShadow, d.material + 8, 4);
  memcpy(&u.pointSize, d.material + 12, 4);
  memcpy(&u.kSpecular, d.material + 16, 4);
  memcpy(&u.baseColor, d.material + 32, 16);
  memcpy(u.aux, d.material, 16);
}

/* ------------------------------------------------------------------ textures (TextureSoft.h, SamplerSoft.h) */

struct Level {
  int w = 0, h = 0;
  std::vector<uint32_t> rgba;  /* RGBA8 texels, x + y*w */
  std::vector<float> f;        /* FLOAT32 texels */
  std::vector<uint32_t> rgbaMs; /* 4 samples per pixel, contiguous (TextureSoft.h:33) */
  std::vector<float> fMs;
  bool hasSingle = false, hasMs = false;
  mutable std::vector<uint8_t> touched; /* texels read by the current draw (unique-texel counting) */
};

struct Texture {
  int type = 0, format = 0, w = 0, h = 0, usage = 0;
  bool mips = false, msaa = false;
  int filterMin = 0, filterMag = 0, wrapS = 2, border = 0;
  std::vector<std::vector<Level>> layers; /* [layer][level] */
};

static inline uint32_t packRGBA(int r, int g, int b, int a) {
  return (uint32_t)(r & 255) | ((uint32_t)(g & 255) << 8) | ((uint32_t)(b & 255) << 16) | ((uint32_t)(a & 255) << 24);
}
static inline int chan(uint32_t c, int i) { return (int)((c >> (8 * i)) & 255u); }

/* float -> u8 as the reference's glm conversion compiles: cvttss2si to int32, low byte. */
static inline int u8trunc(float v) { return (int)(uint8_t)(int32_t)v; }

/* glm::mix on u8vec4 (func_common.inl:104-113) as compiled: u8(fma(y, a, x*(1-a))) per channel. */
static inline uint32_t mixRGBA(uint32_t x, uint32_t y, float a) {
  float na = 1.0f - a;
  int c[4];
  for (int i = 0; i < 4; i++) c[i] = u8trunc(fmaf((float)chan(y, i), a, (float)chan(x, i) * na));
  return packRGBA(c[0], c[1], c[2], c[3]);
}
static inline float mixF(float x, float y, float a) { return fmaf(y, a, x * (1.0f - a)); }

/* SamplerSoft.h:14 CoordMod: `(i & (n-1) + n) & (n-1)` parses as i & (2n-1) & (n-1) = i & (n-1). */
static inline int coordMod(int i, int n) { return i & (n - 1); }

template <typename T>
struct TexelOps;
template <>
struct TexelOps<uint32_t> {
  static const uint32_t *data(const Level &l) { return l.rgba.data(); }
  static uint32_t mix(uint32_t a, uint32_t b, float t) { return mixRGBA(a, b, t); }
  static uint32_t zero() { return 0; }
};
template <>
struct TexelOps<float> {
  static const float *data(const Level &l) { return l.f.data(); }
  static float mix(float a, float b, float t) { return mixF(a, b, t); }
  static float zero() { return 0.f; }
};

/* pixelWithWrapMode, SamplerSoft.h:170-213 */
template <typename T>
static T pixelWrap(const Level &l, int x, int y, int wrap, T border) {
  int w = l.w, h = l.h;
  switch (wrap) {
    case 0: /* REPEAT */
      x = coordMod(x, w);
      y = coordMod(y, h);
      break;
    case 1: /* MIRRORED_REPEAT */
      x = coordMod(x, 2 * w);
      y = coordMod(y, 2 * h);
      x -= w;
      y -= h;
      x = x >= 0 ? x : (-1 - x);
      y = y >= 0 ? y : (-1 - y);
      x = w - 1 - x;
      y = h - 1 - y;
      break;
    case 2: /* CLAMP_TO_EDGE */
      if (x < 0) x = 0;
      if (y < 0) y = 0;
      if (x >= w) x = w - 1;
      if (y >= h) y = h - 1;
      break;
    case 3: /* CLAMP_TO_BORDER */
      if (x < 0 || x >= w) return border;
      if (y < 0 || y >= h) return border;
      break;
  }
  if ((size_t)x >= (size_t)w || (size_t)y >= (size_t)h) return TexelOps<T>::zero(); /* Buffer::get -> nullptr */
  g_stats[6]++;
  if (g_touch) {
    if (l.touched.size() != (size_t)w * h) l.touched.assign((size_t)w * h, 0);
    l.touched[(size_t)y * w + x] = 1;
  }
  return TexelOps<T>::data(l)[(size_t)y * w + x];
}

/* samplePixelBilinear, SamplerSoft.h:254-266 */
template <typename T>
static T bilinearAt(const Level &l, float ux, float uy, int wrap, T border) {
  float fx0 = ux - 0.5f, fy0 = uy - 0.5f;
  float flx = floorf(fx0), fly = floorf(fy0);
  int x = (int)flx, y = (int)fly;
  T s1 = pixelWrap<T>(l, x, y, wrap, border);
  T s2 = pixelWrap<T>(l, x + 1, y, wrap, border);
  T s3 = pixelWrap<T>(l, x, y + 1, wrap, border);
  T s4 = pixelWrap<T>(l, x + 1, y + 1, wrap, border);
  float fx = fx0 - flx, fy = fy0 - fly;
  return TexelOps<T>::mix(TexelOps<T>::mix(s1, s2, fx), TexelOps<T>::mix(s3, s4, fx), fy);
}

/* sampleNearest / sampleBilinear, SamplerSoft.h:215-238 */
template <typename T>
static T sampleNearest(const Level &l, V2 uv, int wrap, T border) {
  float tx = uv.x * (float)l.w, ty = uv.y * (float)l.h;
  return pixelWrap<T>(l, (int)floorf(tx), (int)floorf(ty), wrap, border);
}
template <typename T>
static T sampleBilinear(const Level &l, V2 uv, int wrap, T border) {
  return bilinearAt<T>(l, uv.x * (float)l.w, uv.y * (float)l.h, wrap, border);
}

/* BaseSampler::textureImpl, SamplerSoft.h:117-168 */
template <typename T>
static T textureImpl(const Texture &t, int layer, V2 uv, float lod, T border) {
  const std::vector<Level> &lv = t.layers[layer];
  if (lv.empty()) return TexelOps<T>::zero();
  int fm = t.filterMin, wrap = t.wrapS;
  if (fm == 0) return sampleNearest<T>(lv[0], uv, wrap, border);
  if (fm == 1) return sampleBilinear<T>(lv[0], uv, wrap, border);
  int maxLevel = (int)lv.size() - 1;
  if (fm == 2 || fm == 3) {
    int level = std::min(std::max((int)ceilf(lod + 0.5f) - 1, 0), maxLevel);
    return fm == 2 ? sampleNearest<T>(lv[level], uv, wrap, border) : sampleBilinear<T>(lv[level], uv, wrap, border);
  }
  int hi = std::min(std::max((int)floorf(lod), 0), maxLevel);
  int lo = std::min(std::max(hi + 1, 0), maxLevel);
  T th = fm == 4 ? sampleNearest<T>(lv[hi], uv, wrap, border) : sampleBilinear<T>(lv[hi], uv, wrap, border);
  if (hi == lo) return th;
  T tl = fm == 4 ? sampleNearest<T>(lv[lo], uv, wrap, border) : sampleBilinear<T>(lv[lo], uv, wrap, border);
  float f = lod - floorf(lod); /* glm::fract */
  return TexelOps<T>::mix(th, tl, f);
}

static uint32_t borderRGBA(const Texture &t) { return t.border == 1 ? 0xffffffffu : 0u; }
static float borderF(const Texture &t) { return t.border == 1 ? 1.f : 0.f; }

/* convertXYZ2UV, SamplerSoft.h:311-373 (later faces win ties) */
static void cubeFace(V3 c, int *face, V2 *uv) {
  float ax = fabsf(c.x), ay = fabsf(c.y), az = fabsf(c.z);
  bool px = c.x > 0, py = c.y > 0, pz = c.z > 0;
  float ma = 0, uc = 0, vc = 0;
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

/* generateMipmaps + sampleBufferBilinear, SamplerSoft.h:89-110, 240-252 */
template <typename T>
static void buildMips(Texture &t, int layer) {
  std::vector<Level> &lv = t.layers[layer];
  lv.resize(1);
  int w = t.w, h = t.h;
  int count = (int)floor(log2((double)std::max(w, h))) + 1;
  for (int i = 1; i < count; i++) {
    Level l;
    l.w = std::max(1, w >> i);
    l.h = std::max(1, h >> i);
    l.hasSingle = true;
    const Level &src = lv[i - 1];
    float rx = (float)src.w / (float)l.w, ry = (float)src.h / (float)l.h;
    float dx = 0.5f * rx, dy = 0.5f * ry;
    std::vector<T> out((size_t)l.w * l.h);
    for (int y = 0; y < l.h; y++)
      for (int x = 0; x < l.w; x++)
        out[(size_t)y * l.w + x] = bilinearAt<T>(src, fmaf((float)x, rx, dx), fmaf((float)y, ry, dy), 2,
                                                 TexelOps<T>::zero());
    if (std::is_same<T, uint32_t>::value)
      l.rgba.assign((const uint32_t *)out.data(), (const uint32_t *)out.data() + out.size());
    else
      l.f.assign((const float *)out.data(), (const float *)out.data() + out.size());
    lv.push_back(std::move(l));
  }
}

/* ------------------------------------------------------------------ geometry / programs */

struct VAO {
  int stride = 0;
  int nverts = 0;
  std::vector<float> verts;
  std::vector<int32_t> idx;
};

struct FBO {
  int color = -1, colorLayer = 0, colorLevel = 0, depth = -1;
};

struct VertexHolder {
  bool discard = false;
  std::vector<float> raw; /* attribute floats (own copy; RendererSoft.cpp:1058-1070 for clip vertices) */
  const float *rawPtr = nullptr;
  std::vector<float> vary;
  V4 clip{0, 0, 0, 0};
  int mask = 0;
  V4 frag{0, 0, 0, 0};
};

struct Prim {
  bool discard = false;
  bool front = true;
  size_t i[3] = {0, 0, 0};
};

/* Varying counts including GLM padding (BlinnPhongSoft.h:58-68 = 128 B, PbrSoft.h:63-72 = 112 B). */
static int varyingCount(int shading) {
  switch (shading) {
    case RHS_SHADING_BLINN_PHONG:
    case RHS_SHADING_BLINN_PHONG_DIR:
      return 32;
    case RHS_SHADING_PBR:
      return 28;
    case RHS_SHADING_VERTEX_COLOR:
      return 4;
    case RHS_SHADING_SKYBOX:
    case RHS_SHADING_IBL_IRRADIANCE:
    case RHS_SHADING_IBL_PREFILTER:
      return 3; /* v_worldPos */
    case RHS_SHADING_FXAA:
      return 2; /* v_texCoord */
    default:
      return 0;
  }
}

struct DrawCtx {
  int shading;
  uint32_t defines;
  Uniforms u;
  const Texture *tex[RHS_SLOT_COUNT] = {};
  RhsStates st;
  int nvary = 0;
};

/* Vertex shaders: BasicSoft.h:66-69, BlinnPhongSoft.h:103-121, PbrSoft.h:110-127 */
static void runVS(const DrawCtx &c, const float *a, VertexHolder &vh) {
  const Uniforms &u = c.u;
  V3 pos{a[0], a[1], a[2]};
  vh.clip = m4p(u.mvp, pos.x, pos.y, pos.z);
  float *v = vh.vary.data();
  if (c.shading == RHS_SHADING_BLINN_PHONG || c.shading == RHS_SHADING_BLINN_PHONG_DIR ||
      c.shading == RHS_SHADING_PBR) {
    bool bp = c.shading != RHS_SHADING_PBR;
    v[0] = a[4];
    v[1] = a[5];
    V4 wp4 = m4p(u.model, pos.x, pos.y, pos.z);
    V3 wp{wp4.x, wp4.y, wp4.z};
    V3 nrm{a[8], a[9], a[10]};
    V3 nv = m3v(u.model, 4, nrm);
    v[4] = nv.x, v[5] = nv.y, v[6] = nv.z;
    v[8] = wp.x, v[9] = wp.y, v[10] = wp.z;
    V3 cd = sub3(u.camPos, wp);
    v[12] = cd.x, v[13] = cd.y, v[14] = cd.z;
    V3 ld = c.shading == RHS_SHADING_BLINN_PHONG_DIR ? u.lightPos : sub3(u.lightPos, wp);
    v[16] = ld.x, v[17] = ld.y, v[18] = ld.z;
    if (bp) {
      V4 s = m4p(u.shadowMvp, pos.x, pos.y, pos.z);
      v[20] = s.x, v[21] = s.y, v[22] = s.z, v[23] = s.w;
    }
    if (c.defines & RHS_DEF_NORMAL_MAP) {
      V3 tg{a[12], a[13], a[14]};
      /* as compiled: x-first normalisations, T - dot(T, N) N fused (PbrSoft.h:96-101, BlinnPhongSoft.h:115-120) */
      V3 N = normalize3b(m3v(u.invT, 4, nrm));
      V3 T = normalize3b(m3v(u.invT, 4, tg));
      const float d = dot3b(T, N);
      V3 Tn = normalize3b({fmaf(-d, N.x, T.x), fmaf(-d, N.y, T.y), fmaf(-d, N.z, T.z)});
      int o = bp ? 24 : 20;
      v[o] = N.x, v[o + 1] = N.y, v[o + 2] = N.z;
      v[o + 4] = Tn.x, v[o + 5] = Tn.y, v[o + 6] = Tn.z;
    }
  } else if (c.shading == RHS_SHADING_VERTEX_COLOR) {
    v[0] = a[4], v[1] = a[5], v[2] = a[6], v[3] = a[7];
  } else if (c.shading == RHS_SHADING_SKYBOX) {
    /* SkyboxSoft.h:66-75 */
    vh.clip = {vh.clip.x, vh.clip.y, vh.clip.w, vh.clip.w};
    if (u.reverseZ) vh.clip.z = 0.f;
    v[0] = a[0], v[1] = a[1], v[2] = a[2];
  } else if (c.shading == RHS_SHADING_IBL_IRRADIANCE || c.shading == RHS_SHADING_IBL_PREFILTER) {
    /* IBLIrradianceSoft.h:62-67, IBLPrefilterSoft.h:66-71 */
    vh.clip.z = vh.clip.w;
    v[0] = a[0], v[1] = a[1], v[2] = a[2];
  } else if (c.shading == RHS_SHADING_FXAA) {
    /* FxaaSoft.h:72-75 */
    vh.clip = {a[0], a[1], a[2], 1.0f};
    v[0] = a[4], v[1] = a[5];
  }
}

/* countFrustumClipMask, RendererSoft.cpp:994-1003 */
static int clipMask(V4 c) {
  int m = 0;
  if (c.w < c.x) m |= 1;
  if (c.w < -c.x) m |= 2;
  if (c.w < c.y) m |= 4;
  if (c.w < -c.y) m |= 8;
  if (c.w < c.z) m |= 16;
  if (c.w < -c.z) m |= 32;
  return m;
}

/* dot(FrustumClipPlane[i], clipPos) via dpps: (p0*c0 + p1*c1) + (p2*c2 + p3*c3); Geometry.h:126-133 */
static float planeDist(int i, V4 c) {
  static const float P[6][4] = {{-1, 0, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 1}, {0, 1, 0, 1}, {0, 0, -1, 1}, {0, 0, 1, 1}};
  float a = P[i][0] * c.x, b = P[i][1] * c.y, d = P[i][2] * c.z, e = P[i][3] * c.w;
  return (a + b) + (d + e);
}

/* ------------------------------------------------------------------ fragment shaders */

struct FragIn {
  const float *v[4]; /* the quad's four pixels' varyings, v[k] = pixel k (p0 p1 p2 p3) */
  int self;          /* which of the four is being shaded */
  bool front;
};

static V4 tex2D(const DrawCtx &c, int slot, V2 uv, float lod) {
  const Texture *t = c.tex[slot];
  if (!t) return {0, 0, 0, 0};
  uint32_t p = textureImpl<uint32_t>(*t, 0, uv, lod, borderRGBA(*t));
  return {(float)chan(p, 0) / 255.f, (float)chan(p, 1) / 255.f, (float)chan(p, 2) / 255.f, (float)chan(p, 3) / 255.f};
}

static V4 texCube(const DrawCtx &c, int slot, V3 dir, float lod) {
  const Texture *t = c.tex[slot];
  if (!t) return {0, 0, 0, 0};
  int face;
  V2 uv;
  cubeFace(dir, &face, &uv);
  uint32_t p = textureImpl<uint32_t>(*t, face, uv, lod, borderRGBA(*t));
  return {(float)chan(p, 0) / 255.f, (float)chan(p, 1) / 255.f, (float)chan(p, 2) / 255.f, (float)chan(p, 3) / 255.f};
}

static float texShadow(const DrawCtx &c, V2 uv) {
  const Texture *t = c.tex[RHS_SLOT_SHADOW];
  if (!t) return 0.f;
  return textureImpl<float>(*t, 0, uv, 0.f, borderF(*t));
}

/* ShaderSoft::getSampler2DLod, ShaderSoft.h:117-133: coarse derivative from quad pixels 0,1,2,
 * texture size of level 0; the sampler consults it only when it filters with mipmaps. */
static float lodFor(const DrawCtx &c, int slot, const FragIn &f) {
  const Texture *t = c.tex[slot];
  if (!t || t->filterMin <= 1) return 0.f;
  float w = (float)t->w, h = (float)t->h;
  const float *c0 = f.v[0], *c1 = f.v[1], *c2 = f.v[2];
  float dxx = (c1[0] - c0[0]) * w, dyx = (c2[0] - c0[0]) * w;
  float dxy = (c1[1] - c0[1]) * h, dyy = (c2[1] - c0[1]) * h;
  float ddx = fmaf(dxy, dxy, dxx * dxx);
  float ddy = fmaf(dyy, dyy, dyx * dyx);
  float d = ddy < ddx ? ddx : ddy; /* vmaxss */
  float lod = 0.5f * m_log2f(d);
  return 0.f < lod ? lod : 0.f;
}

/* GetNormalFromMap of the Blinn-Phong shader (BlinnPhongSoft.h:150-163) as the reference build compiles
 * it: x-first normalisations, T - dot(T, N) N fused, and TBN * tn with x and y in a vector register,
 * fma(N, tn.z, fma(B, tn.y, T tn.x)), z as a scalar, fma(N.z, tn.z, fma(T.z, tn.x, B.z tn.y)). */
static V3 normalFromMap(const DrawCtx &c, const FragIn &f, int offN, int offT, float lod) {
  const float *v = f.v[f.self];
  if (c.defines & RHS_DEF_NORMAL_MAP) {
    V3 N = normalize3b({v[offN], v[offN + 1], v[offN + 2]});
    V3 T = normalize3b({v[offT], v[offT + 1], v[offT + 2]});
    const float d = dot3b(T, N);
    T = normalize3b({fmaf(-d, N.x, T.x), fmaf(-d, N.y, T.y), fmaf(-d, N.z, T.z)});
    V3 B = cross3(T, N);
    V4 tn4 = tex2D(c, RHS_SLOT_NORMAL, {v[0], v[1]}, lod);
    V3 tn{tn4.x * 2.0f - 1.0f, tn4.y * 2.0f - 1.0f, tn4.z * 2.0f - 1.0f};
    V3 r{fmaf(N.x, tn.z, fmaf(B.x, tn.y, T.x * tn.x)), fmaf(N.y, tn.z, fmaf(B.y, tn.y, T.y * tn.x)),
         fmaf(N.z, tn.z, fmaf(T.z, tn.x, B.z * tn.y))};
    return normalize3b(r);
  }
  return normalize3b({v[4], v[5], v[6]});
}

/* BlinnPhongSoft.h:165-189 */
static float shadowCalc(const DrawCtx &c, V4 sp, V3 n, V3 lightDir) {
  V3 pc{sp.x / sp.w, sp.y / sp.w, sp.z / sp.w};
  float cur = pc.z;
  if (cur < 0.f || cur > 1.f) return 0.0f;
  float bias = maxf(0.00025f * (1.0f - dot3b(n, normalize3b(lightDir))), 0.00005f); /* x-first, as compiled */
  const Texture *t = c.tex[RHS_SLOT_SHADOW];
  float ox = t ? 1.0f / (float)t->w : 0.f, oy = t ? 1.0f / (float)t->h : 0.f;
  float s = 0.f;
  for (int x = -1; x <= 1; ++x)
    for (int y = -1; y <= 1; ++y) {
      float d = texShadow(c, {pc.x + (float)x * ox, pc.y + (float)y * oy});
      if (c.u.reverseZ)
        s += cur + bias < d ? 1.0f : 0.0f;
      else
        s += cur - bias > d ? 1.0f : 0.0f;
    }
  return s / 9.0f;
}

/* BlinnPhongSoft.h:191-242 (and the directional-light variant: constant light vector, no attenuation) */
static V4 fsBlinnPhong(const DrawCtx &c, const FragIn &f, bool directional) {
  const float *v = f.v[f.self];
  const Uniforms &u = c.u;
  V2 uv{v[0], v[1]};
  V4 base = (c.defines & RHS_DEF_ALBEDO_MAP) ? tex2D(c, RHS_SLOT_ALBEDO, uv, lodFor(c, RHS_SLOT_ALBEDO, f))
                                             : u.baseColor;
  V3 N = normalFromMap(c, f, 24, 28, lodFor(c, RHS_SLOT_NORMAL, f));
  float ao = (c.defines & RHS_DEF_AO_MAP) ? tex2D(c, RHS_SLOT_AO, uv, lodFor(c, RHS_SLOT_AO, f)).x : 1.f;
  V3 b3{base.x, base.y, base.z};
  V3 amb = scale3(mul3(b3, u.ambient), ao);
  V3 dif{0, 0, 0}, spe{0, 0, 0}, emi{0, 0, 0};
  if (u.enableLight) {
    /* the reference build sums every dot product x first, and computes L + normalize(C) with the
     * normalisation's multiply fused into the add */
    V3 ldv{v[16], v[17], v[18]};
    V3 L = normalize3b(ldv);
    float diffuse = maxf(dot3b(N, L), 0.0f);
    if (directional) {
      dif = scale3(mul3(u.lightColor, b3), diffuse);
    } else {
      V3 l = scale3(ldv, 1.0f / 5.f);
      float att = clampf(1.0f - dot3b(l, l), 0.0f, 1.0f);
      dif = scale3(scale3(mul3(u.lightColor, b3), diffuse), att);
    }
    V3 cr{v[12], v[13], v[14]};
    const float ic = 1.0f / sqrtf(dot3b(cr, cr));
    V3 H = normalize3b({fmaf(cr.x, ic, L.x), fmaf(cr.y, ic, L.y), fmaf(cr.z, ic, L.z)});
    float sa = maxf(dot3b(N, H), 0.0f);
    float sp = u.kSpecular * m_powf(sa, 128.f);
    spe = {sp, sp, sp};
    if (u.enableShadow) {
      float s = 1.0f - shadowCalc(c, {v[20], v[21], v[22], v[23]}, N, ldv);
      dif = scale3(dif, s);
      spe = scale3(spe, s);
    }
  }
  if (c.defines & RHS_DEF_EMISSIVE_MAP) {
    V4 e = tex2D(c, RHS_SLOT_EMISSIVE, uv, lodFor(c, RHS_SLOT_EMISSIVE, f));
    emi = {e.x, e.y, e.z};
  }
  V3 col = add3(add3(add3(amb, dif), spe), emi);
  return {col.x, col.y, col.z, base.w};
}

/* GetNormalFromMap of the PBR shader (PbrSoft.h:153-166) as the reference build compiles it: x-first
 * normalisations and T - dot(T, N) N as one multiply-add per component (the Blinn-Phong shader's copy of
 * this function compiles differently: normalFromMap). */
static V3 normalFromMapPbr(const DrawCtx &c, const FragIn &f, float lod) {
  const float *v = f.v[f.self];
  if (c.defines & RHS_DEF_NORMAL_MAP) {
    V3 N = normalize3b({v[20], v[21], v[22]});
    V3 T = normalize3b({v[24], v[25], v[26]});
    const float d = dot3b(T, N);
    T = normalize3b({fmaf(-d, N.x, T.x), fmaf(-d, N.y, T.y), fmaf(-d, N.z, T.z)});
    V3 B = cross3(T, N);
    V4 tn4 = tex2D(c, RHS_SLOT_NORMAL, {v[0], v[1]}, lod);
    V3 tn{tn4.x * 2.0f - 1.0f, tn4.y * 2.0f - 1.0f, tn4.z * 2.0f - 1.0f};
    V3 r{fmaf(N.x, tn.z, fmaf(T.x, tn.x, B.x * tn.y)), fmaf(N.y, tn.z, fmaf(T.y, tn.x, B.y * tn.y)),
         fmaf(N.z, tn.z, fmaf(T.z, tn.x, B.z * tn.y))};
    return normalize3b(r);
  }
  return normalize3b({v[4], v[5], v[6]});
}

/* PbrSoft.h:168-223 */
static const float kPI = 3.14159265359f;
static float distGGX(V3 N, V3 H, float r) {
  float a = r * r, a2 = a * a;
  float nh = maxf(dot3b(N, H), 0.0f);
  float nh2 = nh * nh;
  float den = fmaf(a2 - 1.0f, nh2, 1.0f); /* the reference build fuses NdotH2 * (a2 - 1) + 1 */
  den = kPI * den * den;
  return a2 / den;
}
static float schlickGGX(float nv, float r) {
  float k = ((r + 1.0f) * (r + 1.0f)) / 8.0f;
  return nv / (nv * (1.0f - k) + k);
}
/* every multiply-add of EnvBRDFApprox is fused in the reference build */
static V3 envBRDFApprox(V3 spec, float r, float nv) {
  float rx = fmaf(r, -1.f, 1.f), ry = fmaf(r, -0.0275f, 0.0425f), rz = fmaf(r, -0.572f, 1.04f),
        rw = fmaf(r, 0.022f, -0.04f);
  float a004 = fmaf(minf(rx * rx, m_exp2f(-9.28f * nv)), rx, ry);
  float ax = fmaf(-1.04f, a004, rz), ay = fmaf(1.04f, a004, rw);
  ay *= maxf(0.f, minf(1.f, 50.0f * spec.y));
  return {fmaf(spec.x, ax, ay), fmaf(spec.y, ax, ay), fmaf(spec.z, ax, ay)};
}

/* PbrSoft.h:225-322 */
static V4 fsPbr(const DrawCtx &c, const FragIn &f) {
  const float *v = f.v[f.self];
  const Uniforms &u = c.u;
  V2 uv{v[0], v[1]};
  V4 albedoRGBA = (c.defines & RHS_DEF_ALBEDO_MAP) ? tex2D(c, RHS_SLOT_ALBEDO, uv, lodFor(c, RHS_SLOT_ALBEDO, f))
                                                   : u.baseColor;
  V3 albedo{m_powf(albedoRGBA.x, 2.2f), m_powf(albedoRGBA.y, 2.2f), m_powf(albedoRGBA.z, 2.2f)};
  float metallic = 0.0f, rough = 1.0f;
  if (c.defines & RHS_DEF_METALROUGHNESS_MAP) {
    /* no LOD callback is ever installed on this sampler (PbrSoft.h:138-151): lod 0 */
    V4 mr = tex2D(c, RHS_SLOT_METAL_ROUGHNESS, uv, 0.f);
    metallic = mr.z;
    rough = mr.y;
  }
  float ao = (c.defines & RHS_DEF_AO_MAP) ? tex2D(c, RHS_SLOT_AO, uv, lodFor(c, RHS_SLOT_AO, f)).x : 1.f;
  V3 N = normalFromMapPbr(c, f, lodFor(c, RHS_SLOT_NORMAL, f));
  V3 V = normalize3b({v[12], v[13], v[14]});
  /* glm::reflect(-V, N) = I - 2 dot(N, I) N; the reference build rounds the products of dot(N, I) on
   * their own and reuses them, negated, for dot(N, V) */
  V3 I{-V.x, -V.y, -V.z};
  float dNI = dot3s(N, I);
  const float dNV = -dNI;
  V3 R = sub3(I, scale3(N, dNI * 2.0f));
  V3 F0{mixF(0.04f, albedo.x, metallic), mixF(0.04f, albedo.y, metallic), mixF(0.04f, albedo.z, metallic)};
  V3 Lo{0, 0, 0};
  if (u.enableLight) {
    V3 ldv{v[16], v[17], v[18]};
    V3 L = normalize3b(ldv);
    V3 H = normalize3b(add3(V, L));
    V3 ld = scale3(ldv, 1.0f / 5.f);
    float att = clampf(1.0f - dot3b(ld, ld), 0.0f, 1.0f);
    V3 radiance = scale3(u.lightColor, att);
    float NDF = distGGX(N, H, rough);
    float nv = maxf(dNV, 0.0f), nl = maxf(dot3b(N, L), 0.0f);
    float G = schlickGGX(nl, rough) * schlickGGX(nv, rough);
    float ct = maxf(dot3b(H, V), 0.0f);
    float p5 = m_powf(clampf(1.0f - ct, 0.0f, 1.0f), 5.0f);
    V3 F{fmaf(1.0f - F0.x, p5, F0.x), fmaf(1.0f - F0.y, p5, F0.y), fmaf(1.0f - F0.z, p5, F0.z)};
    float ng = NDF * G;
    V3 num = scale3(F, ng);
    float den = fmaf(4.0f * nv, nl, 0.0001f);
    V3 spec{num.x / den, num.y / den, num.z / den};
    V3 kD{(1.0f - F.x) * (1.0f - metallic), (1.0f - F.y) * (1.0f - metallic), (1.0f - F.z) * (1.0f - metallic)};
    V3 t{kD.x * albedo.x / kPI + spec.x, kD.y * albedo.y / kPI + spec.y, kD.z * albedo.z / kPI + spec.z};
    Lo = scale3(mul3(t, radiance), nl);
  }
  V3 ambient;
  if (u.enableIBL) {
    float nv = maxf(dNV, 0.0f);
    float p5 = m_powf(clampf(1.0f - nv, 0.0f, 1.0f), 5.0f);
    float omr = 1.0f - rough;
    V3 F{fmaf(maxf(omr, F0.x) - F0.x, p5, F0.x), fmaf(maxf(omr, F0.y) - F0.y, p5, F0.y),
         fmaf(maxf(omr, F0.z) - F0.z, p5, F0.z)};
    V3 kD{(1.0f - F.x) * (1.0f - metallic), (1.0f - F.y) * (1.0f - metallic), (1.0f - F.z) * (1.0f - metallic)};
    V4 irr = texCube(c, RHS_SLOT_IRRADIANCE, N, 0.f);
    V3 diffuse{irr.x * albedo.x, irr.y * albedo.y, irr.z * albedo.z};
    V4 pf = texCube(c, RHS_SLOT_PREFILTER, R, rough * 4.0f);
    V3 brdf = envBRDFApprox(F, rough, nv);
    /* (kD * diffuse + prefiltered * brdf) with the second product fused into the add */
    V3 pre{fmaf(brdf.x, pf.x, kD.x * diffuse.x), fmaf(brdf.y, pf.y, kD.y * diffuse.y),
           fmaf(brdf.z, pf.z, kD.z * diffuse.z)};
    ambient = scale3(pre, ao);
  } else {
    ambient = scale3(mul3(u.ambient, albedo), ao);
  }
  V3 col = add3(ambient, Lo);
  col = {m_powf(col.x, 1.0f / 2.2f), m_powf(col.y, 1.0f / 2.2f), m_powf(col.z, 1.0f / 2.2f)};
  if (c.defines & RHS_DEF_EMISSIVE_MAP) {
    V4 e = tex2D(c, RHS_SLOT_EMISSIVE, uv, lodFor(c, RHS_SLOT_EMISSIVE, f));
    col = add3(col, {e.x, e.y, e.z});
  }
  return {col.x, col.y, col.z, albedoRGBA.w};
}

/* ------------------------------------------------------------------ skybox, IBL precompute, FXAA */

/* textureLodOffset, SamplerSoft.h:216-238 (offset after scaling: to the texel-space coordinate for bilinear,
 * to the floored texel for nearest). The reference build fuses the bilinear `uv * size + offset` into one
 * multiply-add, so a coordinate that rounds to a texel centre unshifted may land just past it shifted. */
static uint32_t sampleOff(const Level &l, V2 uv, bool linear, int wrap, uint32_t border, int ox, int oy) {
  if (linear)
    return bilinearAt<uint32_t>(l, fmaf(uv.x, (float)l.w, (float)ox), fmaf(uv.y, (float)l.h, (float)oy), wrap,
                                border);
  float tx = uv.x * (float)l.w, ty = uv.y * (float)l.h;
  return pixelWrap<uint32_t>(l, (int)floorf(tx) + ox, (int)floorf(ty) + oy, wrap, border);
}
static V4 tex2DOffset(const DrawCtx &c, int slot, V2 uv, float lod, int ox, int oy) {
  const Texture *t = c.tex[slot];
  if (!t || t->layers[0].empty()) return {0, 0, 0, 0};
  const std::vector<Level> &lv = t->layers[0];
  const uint32_t b = borderRGBA(*t);
  const int fm = t->filterMin, wrap = t->wrapS, maxLevel = (int)lv.size() - 1;
  uint32_t p;
  if (fm <= 1) {
    p = sampleOff(lv[0], uv, fm == 1, wrap, b, ox, oy);
  } else if (fm == 2 || fm == 3) {
    int level = std::min(std::max((int)ceilf(lod + 0.5f) - 1, 0), maxLevel);
    p = sampleOff(lv[level], uv, fm == 3, wrap, b, ox, oy);
  } else {
    int hi = std::min(std::max((int)floorf(lod), 0), maxLevel), lo = std::min(std::max(hi + 1, 0), maxLevel);
    p = sampleOff(lv[hi], uv, fm == 5, wrap, b, ox, oy);
    if (hi != lo) p = mixRGBA(p, sampleOff(lv[lo], uv, fm == 5, wrap, b, ox, oy), lod - floorf(lod));
  }
  return {(float)chan(p, 0) / 255.f, (float)chan(p, 1) / 255.f, (float)chan(p, 2) / 255.f, (float)chan(p, 3) / 255.f};
}


/* SkyboxSoft.h:78-96: the host's atan2f, asin in double as the reference calls it; the reference build fuses
 * `uv * scale + 0.5`. */
static V4 fsSkybox(const DrawCtx &c, const FragIn &f) {
  const float *v = f.v[f.self];
  V3 wp{v[0], v[1], v[2]};
  if (c.defines & RHS_DEF_EQUIRECTANGULAR_MAP) {
    V3 d = normalize3b(wp);
    float ax = m_atan2f(d.z, d.x), ay = (float)asin((double)-d.y);
    V2 uv{fmaf(ax, 0.1591f, 0.5f), fmaf(ay, 0.3183f, 0.5f)};
    return tex2D(c, RHS_SLOT_EQUIRECT, uv, 0.f); /* the skybox sampler has no LOD callback */
  }
  return texCube(c, RHS_SLOT_CUBE, wp, 0.f);
}

/* IBLIrradianceSoft.h:74-101 */
static V4 fsIrradiance(const DrawCtx &c, const FragIn &f) {
  const float *v = f.v[f.self];
  /* as compiled: x-first normalisations, the sample vector and the sum as multiply-adds */
  V3 N = normalize3b({v[0], v[1], v[2]});
  V3 irr{0.f, 0.f, 0.f};
  V3 up{0.0f, 1.0f, 0.0f};
  V3 right = normalize3b(cross3(up, N));
  up = normalize3b(cross3(N, right));
  const float sampleDelta = 0.025f;
  float nr = 0.0f;
  for (float phi = 0.0f; phi < 2.0f * kPI; phi += sampleDelta) {
    for (float theta = 0.0f; theta < 0.5f * kPI; theta += sampleDelta) {
      V3 ts{m_sinf(theta) * m_cosf(phi), m_sinf(theta) * m_sinf(phi), m_cosf(theta)};
      V3 sv{fmaf(N.x, ts.z, fmaf(up.x, ts.y, right.x * ts.x)), fmaf(N.y, ts.z, fmaf(up.y, ts.y, right.y * ts.x)),
            fmaf(N.z, ts.z, fmaf(up.z, ts.y, right.z * ts.x))};
      V4 t = texCube(c, RHS_SLOT_CUBE, sv, 0.f);
      float ct = m_cosf(theta), st = m_sinf(theta);
      irr = {fmaf(t.x * ct, st, irr.x), fmaf(t.y * ct, st, irr.y), fmaf(t.z * ct, st, irr.z)};
      nr++;
    }
  }
  irr = scale3(scale3(irr, kPI), 1.0f / nr);
  return {irr.x, irr.y, irr.z, 1.0f};
}

/* IBLPrefilterSoft.h:88-168 */
static uint32_t radicalInverseBits(uint32_t bits) {
  bits = (bits << 16u) | (bits >> 16u);
  bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
  bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
  bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
  bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
  return bits;
}
static V4 fsPrefilter(const DrawCtx &c, const FragIn &f) {
  const float *v = f.v[f.self];
  V3 N = normalize3b({v[0], v[1], v[2]});
  V3 R = N, V = R;
  const float roughness = c.u.aux[1], resolution = c.u.aux[0];
  const uint32_t SAMPLE_COUNT = 1024u;
  V3 pre{0.f, 0.f, 0.f};
  float totalWeight = 0.0f;
  for (uint32_t i = 0u; i < SAMPLE_COUNT; ++i) {
    V2 Xi{(float)i / (float)SAMPLE_COUNT, (float)((double)radicalInverseBits(i) * 2.3283064365386963e-10)};
    float a = roughness * roughness;
    float phi = 2.0f * kPI * Xi.x;
    /* the reference build contracts: (a^2 - 1) Xi.y + 1, 1 - cos^2, the tangent-space sum, 2 (V.H) H - V,
     * the GGX denominator, 1024 pdf + 1e-4 and the colour sum into fused multiply-adds */
    float cosTheta = sqrtf((1.0f - Xi.y) / fmaf(a * a - 1.0f, Xi.y, 1.0f));
    float sinTheta = sqrtf(fmaf(-cosTheta, cosTheta, 1.0f));
    V3 Ht{m_cosf(phi) * sinTheta, m_sinf(phi) * sinTheta, cosTheta};
    V3 upv = (double)fabsf(N.z) < 0.999 ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    V3 tangent = normalize3b(cross3(upv, N));
    V3 bitangent = cross3(N, tangent);
    V3 sv{fmaf(N.x, Ht.z, fmaf(bitangent.x, Ht.y, tangent.x * Ht.x)),
          fmaf(N.y, Ht.z, fmaf(bitangent.y, Ht.y, tangent.y * Ht.x)),
          fmaf(N.z, Ht.z, fmaf(bitangent.z, Ht.y, tangent.z * Ht.x))};
    V3 H = normalize3b(sv);
    float vh2 = 2.0f * dot3b(V, H);
    V3 L = normalize3b({fmaf(H.x, vh2, -V.x), fmaf(H.y, vh2, -V.y), fmaf(H.z, vh2, -V.z)});
    float NdotL = maxf(dot3b(N, L), 0.0f);
    if (NdotL > 0.0f) {
      float ra = roughness * roughness, ra2 = ra * ra;
      float nh = maxf(dot3b(N, H), 0.0f);
      float den = fmaf(nh * nh, ra2 - 1.0f, 1.0f);
      float D = ra2 / (kPI * den * den);
      float NdotH = maxf(dot3b(N, H), 0.0f);
      float HdotV = maxf(dot3b(H, V), 0.0f);
      float pdf = D * NdotH / (4.0f * HdotV) + 0.0001f;
      float saTexel = 4.0f * kPI / (6.0f * resolution * resolution);
      float saSample = 1.0f / fmaf((float)SAMPLE_COUNT, pdf, 0.0001f);
      float mipLevel = roughness == 0.0f ? 0.0f : 0.5f * m_log2f(saSample / saTexel);
      V4 t = texCube(c, RHS_SLOT_CUBE, L, mipLevel);
      pre = {fmaf(t.x, NdotL, pre.x), fmaf(t.y, NdotL, pre.y), fmaf(t.z, NdotL, pre.z)};
      totalWeight += NdotL;
    }
  }
  return {pre.x / totalWeight, pre.y / totalWeight, pre.z / totalWeight, 1.0f};
}

/* FxaaSoft.h:81-266 */
/* the reference build contracts: the luma dot product red first, -2 l + corners, |.| * 2 + |.|,
 * 2 (DU + LR) + LC and (average - centre) into fused multiply-adds */
static float rgb2luma(V4 c) { return dot3b({c.x, c.y, c.z}, {0.299f, 0.587f, 0.114f}); }
static float fxaaQuality(float q) { return q < 5.f ? 1.0f : (q > 5.f ? (q < 10.f ? 2.0f : (q < 11.f ? 4.0f : 8.0f)) : 1.5f); }
static V4 fsFxaa(const DrawCtx &c, const FragIn &f) {
  const float *v = f.v[f.self];
  const int S = RHS_SLOT_SCREEN;
  const float lod = 0.f; /* the FXAA sampler has no LOD callback */
  V2 tc{v[0], v[1]};
  V2 inv{1.0f / c.u.aux[0], 1.0f / c.u.aux[1]};
  V4 colorCenter = tex2D(c, S, tc, lod);
  float lumaCenter = rgb2luma(colorCenter);
  float lumaDown = rgb2luma(tex2DOffset(c, S, tc, 0.f, 0, -1));
  float lumaUp = rgb2luma(tex2DOffset(c, S, tc, 0.f, 0, 1));
  float lumaLeft = rgb2luma(tex2DOffset(c, S, tc, 0.f, -1, 0));
  float lumaRight = rgb2luma(tex2DOffset(c, S, tc, 0.f, 1, 0));
  float lumaMin = minf(lumaCenter, minf(minf(lumaDown, lumaUp), minf(lumaLeft, lumaRight)));
  float lumaMax = maxf(lumaCenter, maxf(maxf(lumaDown, lumaUp), maxf(lumaLeft, lumaRight)));
  float lumaRange = lumaMax - lumaMin;
  if (lumaRange < maxf(0.0312f, lumaMax * 0.125f)) return {colorCenter.x, colorCenter.y, colorCenter.z, 1.0f};
  float lumaDownLeft = rgb2luma(tex2DOffset(c, S, tc, 0.f, -1, -1));
  float lumaUpRight = rgb2luma(tex2DOffset(c, S, tc, 0.f, 1, 1));
  float lumaUpLeft = rgb2luma(tex2DOffset(c, S, tc, 0.f, -1, 1));
  float lumaDownRight = rgb2luma(tex2DOffset(c, S, tc, 0.f, 1, -1));
  float lumaDownUp = lumaDown + lumaUp, lumaLeftRight = lumaLeft + lumaRight;
  float lumaLeftCorners = lumaDownLeft + lumaUpLeft, lumaDownCorners = lumaDownLeft + lumaDownRight;
  float lumaRightCorners = lumaDownRight + lumaUpRight, lumaUpCorners = lumaUpRight + lumaUpLeft;
  float edgeHorizontal = fmaf(fabsf(fmaf(-2.0f, lumaCenter, lumaDownUp)), 2.0f,
                              fabsf(fmaf(-2.0f, lumaLeft, lumaLeftCorners))) +
                         fabsf(fmaf(-2.0f, lumaRight, lumaRightCorners));
  float edgeVertical = fmaf(fabsf(fmaf(-2.0f, lumaCenter, lumaLeftRight)), 2.0f,
                            fabsf(fmaf(-2.0f, lumaUp, lumaUpCorners))) +
                       fabsf(fmaf(-2.0f, lumaDown, lumaDownCorners));
  bool isHorizontal = edgeHorizontal >= edgeVertical;
  float stepLength = isHorizontal ? inv.y : inv.x;
  float luma1 = isHorizontal ? lumaDown : lumaLeft, luma2 = isHorizontal ? lumaUp : lumaRight;
  float gradient1 = luma1 - lumaCenter, gradient2 = luma2 - lumaCenter;
  bool is1Steepest = fabsf(gradient1) >= fabsf(gradient2);
  float gradientScaled = 0.25f * maxf(fabsf(gradient1), fabsf(gradient2));
  float lumaLocalAverage;
  if (is1Steepest) {
    stepLength = -stepLength;
    lumaLocalAverage = 0.5f * (luma1 + lumaCenter);
  } else {
    lumaLocalAverage = 0.5f * (luma2 + lumaCenter);
  }
  V2 currentUv = tc;
  if (isHorizontal)
    currentUv.y += stepLength * 0.5f;
  else
    currentUv.x += stepLength * 0.5f;
  V2 offset = isHorizontal ? V2{inv.x, 0.0f} : V2{0.0f, inv.y};
  V2 uv1{currentUv.x - offset.x * fxaaQuality(0.f), currentUv.y - offset.y * fxaaQuality(0.f)};
  V2 uv2{currentUv.x + offset.x * fxaaQuality(0.f), currentUv.y + offset.y * fxaaQuality(0.f)};
  float lumaEnd1 = 0.f, lumaEnd2 = 0.f;
  bool reached1 = false, reached2 = false, reachedBoth = false;
  for (int i = 1; i < 12; i++) {
    if (!reached1) {
      lumaEnd1 = rgb2luma(tex2D(c, S, uv1, lod)) - lumaLocalAverage;
      reached1 = fabsf(lumaEnd1) >= gradientScaled;
    }
    if (!reached2) {
      lumaEnd2 = rgb2luma(tex2D(c, S, uv2, lod)) - lumaLocalAverage;
      reached2 = fabsf(lumaEnd2) >= gradientScaled;
    }
    reachedBoth = reached1 && reached2;
    if (!reached1) uv1 = {uv1.x - offset.x * fxaaQuality((float)i), uv1.y - offset.y * fxaaQuality((float)i)};
    if (!reached2) uv2 = {uv2.x + offset.x * fxaaQuality((float)i), uv2.y + offset.y * fxaaQuality((float)i)};
    if (reachedBoth) break;
  }
  float distance1 = isHorizontal ? (tc.x - uv1.x) : (tc.y - uv1.y);
  float distance2 = isHorizontal ? (uv2.x - tc.x) : (uv2.y - tc.y);
  bool isDirection1 = distance1 < distance2;
  float distanceFinal = minf(distance1, distance2);
  bool isLumaCenterSmaller = lumaCenter < lumaLocalAverage;
  bool correctVariation1 = (lumaEnd1 < 0.0f) != isLumaCenterSmaller;
  bool correctVariation2 = (lumaEnd2 < 0.0f) != isLumaCenterSmaller;
  bool correctVariation = isDirection1 ? correctVariation1 : correctVariation2;
  float edgeLength = distance1 + distance2;
  float pixelOffset = -distanceFinal / edgeLength + 0.5f;
  float finalOffset = correctVariation ? pixelOffset : 0.0f;
  float lumaSum = fmaf(2.0f, lumaDownUp + lumaLeftRight, lumaLeftCorners) + lumaRightCorners;
  float subPixelOffset1 = clampf(fabsf(fmaf(lumaSum, 1.0f / 12.0f, -lumaCenter)) / lumaRange, 0.0f, 1.0f);
  float subPixelOffset2 = fmaf(-2.0f, subPixelOffset1, 3.0f) * subPixelOffset1 * subPixelOffset1;
  float subPixelOffsetFinal = subPixelOffset2 * subPixelOffset2 * 0.75f;
  finalOffset = maxf(finalOffset, subPixelOffsetFinal);
  V2 finalUv = tc;
  if (isHorizontal)
    finalUv.y += finalOffset * stepLength;
  else
    finalUv.x += finalOffset * stepLength;
  V4 fc = tex2D(c, S, finalUv, lod);
  return {fc.x, fc.y, fc.z, 1.0f};
}

static V4 runFS(const DrawCtx &c, const FragIn &f) {
  switch (c.shading) {
    case RHS_SHADING_BASIC:
      return c.u.baseColor;
    case RHS_SHADING_BLINN_PHONG:
      return fsBlinnPhong(c, f, false);
    case RHS_SHADING_BLINN_PHONG_DIR:
      return fsBlinnPhong(c, f, true);
    case RHS_SHADING_PBR:
      return fsPbr(c, f);
    case RHS_SHADING_VERTEX_COLOR: {
      const float *v = f.v[f.self];
      return {v[0], v[1], v[2], v[3]};
    }
    case RHS_SHADING_SKYBOX:
      return fsSkybox(c, f);
    case RHS_SHADING_IBL_IRRADIANCE:
      return fsIrradiance(c, f);
    case RHS_SHADING_IBL_PREFILTER:
      return fsPrefilter(c, f);
    case RHS_SHADING_FXAA:
      return fsFxaa(c, f);
  }
  return {0, 0, 0, 0};
}

/* ------------------------------------------------------------------ blending (BlendSoft.h) */

static float blendFactorF(int f, float s, float sa, float d, float da) {
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
static float blendFuncF(int fn, float s, float d) {
  switch (fn) {
    case 0: return s + d;
    case 1: return s - d;
    case 2: return d - s;
    case 3: return minf(s, d);
    case 4: return maxf(s, d);
  }
  return s + d;
}
static V4 blendColor(V4 s, V4 d, const RhsStates &st) {
  float sc[3] = {s.x, s.y, s.z}, dc[3] = {d.x, d.y, d.z}, r[4];
  for (int i = 0; i < 3; i++) {
    float fs = blendFactorF(st.blendSrcRgb, sc[i], s.w, dc[i], d.w);
    float fd = blendFactorF(st.blendDstRgb, sc[i], s.w, dc[i], d.w);
    r[i] = blendFuncF(st.blendFuncRgb, sc[i] * fs, dc[i] * fd);
  }
  float fsa = blendFactorF(st.blendSrcAlpha, s.w, s.w, d.w, d.w);
  float fda = blendFactorF(st.blendDstAlpha, s.w, s.w, d.w, d.w);
  r[3] = blendFuncF(st.blendFuncAlpha, s.w * fsa, d.w * fda);
  return {r[0], r[1], r[2], r[3]};
}

/* DepthSoft.h:13-25 */
static bool depthFuncPass(float a, float b, int fn) {
  const float eps = 1.19209290e-07f;
  switch (fn) {
    case 0: return false;
    case 1: return a < b;
    case 2: return fabsf(a - b) <= eps;
    case 3: return a <= b;
    case 4: return a > b;
    case 5: return fabsf(a - b) > eps;
    case 6: return a >= b;
    case 7: return true;
  }
  return a < b;
}

/* ------------------------------------------------------------------ the renderer */

struct Sample {
  bool inside = false;
  int fx = 0, fy = 0;
  V4 pos{0, 0, 0, 0};
  V4 bc{0, 0, 0, 0};
};

struct Pixel {
  bool inside = false;
  Sample s[5];
  int nsamp = 1;
  int shading = 0; /* index into s[] */
  std::vector<float> vary;
};

struct Renderer {
  std::map<int, Texture> tex;
  std::map<int, VAO> vaos;
  std::map<int, FBO> fbos;
  FBO *fbo = nullptr;
  float vpX = 0, vpY = 0, vpW = 0, vpH = 0;
  float oX = 0, oY = 0, pX = 0, pY = 0; /* innerO / innerP, RendererSoft.cpp:92-113 */

  Level *colorLevel() {
    if (!fbo || fbo->color < 0) return nullptr;
    return &tex[fbo->color].layers[fbo->colorLayer][fbo->colorLevel];
  }
  Level *depthLevel() {
    if (!fbo || fbo->depth < 0) return nullptr;
    return &tex[fbo->depth].layers[0][0];
  }

  /* RendererSoft::beginRenderPass, :61-90 */
  void begin(const uint32_t *p) {
    auto it = fbos.find((int)p[0]);
    fbo = it == fbos.end() ? nullptr : &it->second;
    if (!fbo) return;
    Level *c = colorLevel(), *d = depthLevel();
    if (p[1] && c) {
      uint32_t px = packRGBA(u8trunc(rhs_f32(p[3]) * 255), u8trunc(rhs_f32(p[4]) * 255), u8trunc(rhs_f32(p[5]) * 255),
                             u8trunc(rhs_f32(p[6]) * 255));
      if (c->hasMs)
        std::fill(c->rgbaMs.begin(), c->rgbaMs.end(), px);
      else
        std::fill(c->rgba.begin(), c->rgba.end(), px);
    }
    if (p[2] && d) {
      float z = rhs_f32(p[7]);
      if (d->hasMs)
        std::fill(d->fMs.begin(), d->fMs.end(), z);
      else
        std::fill(d->f.begin(), d->f.end(), z);
    }
  }

  void viewport(const uint32_t *p) {
    vpX = (float)(int)p[0];
    vpY = (float)(int)p[1];
    vpW = (float)(int)p[2];
    vpH = (float)(int)p[3];
    oX = vpX + vpW / 2.f;
    oY = vpY + vpH / 2.f;
    pX = vpW / 2.f;
    pY = vpH / 2.f;
  }

  /* per-draw state */
  DrawCtx ctx;
  std::vector<VertexHolder> verts;
  std::vector<Prim> prims;
  const VAO *vao = nullptr;
  Level *cl = nullptr, *dl = nullptr;
  int samples = 1;

  size_t newClipVertex(size_t i0, size_t i1, float t) {
    /* clippingNewVertex + interpolateVertex, RendererSoft.cpp:956-970, 1058-1070 */
    VertexHolder vh;
    int nf = vao->stride / 4;
    vh.raw.resize(nf);
    const float *a = verts[i0].rawPtr, *b = verts[i1].rawPtr;
    float omt = 1.0f - t;
    for (int k = 0; k < nf; k++) vh.raw[k] = fmaf(a[k], omt, b[k] * t);
    vh.vary.assign(std::max(ctx.nvary, 1), 0.f);
    verts.push_back(std::move(vh));
    VertexHolder &nv = verts.back();
    nv.rawPtr = nv.raw.data();
    runVS(ctx, nv.rawPtr, nv);
    nv.mask = clipMask(nv.clip);
    /* re-point the raw pointers of vertices that own their attribute copy */
    for (auto &x : verts)
      if (!x.raw.empty()) x.rawPtr = x.raw.data();
    return verts.size() - 1;
  }

  /* clippingTriangle, RendererSoft.cpp:489-559 */
  void clipTriangle(Prim &tri, std::vector<Prim> &append) {
    int mask = verts[tri.i[0]].mask | verts[tri.i[1]].mask | verts[tri.i[2]].mask;
    if (mask == 0) return;
    g_stats[3]++;
    bool full = false;
    std::vector<size_t> in{tri.i[0], tri.i[1], tri.i[2]}, out;
    for (int pl = 0; pl < 6; pl++) {
      if (!(mask & (1 << pl))) continue;
      if (in.size() < 3) {
        full = true;
        break;
      }
      out.clear();
      size_t ip = in[0];
      float dp = planeDist(pl, verts[ip].clip);
      in.push_back(ip);
      for (size_t k = 1; k < in.size(); k++) {
        size_t ic = in[k];
        float d = planeDist(pl, verts[ic].clip);
        if (dp >= 0) out.push_back(ip);
        if (std::signbit(dp) != std::signbit(d)) {
          float t = d < 0 ? dp / (dp - d) : -dp / (d - dp);
          out.push_back(newClipVertex(ip, ic, t));
        }
        ip = ic;
        dp = d;
      }
      std::swap(in, out);
    }
    if (full || in.empty()) {
      tri.discard = true;
      return;
    }
    tri.i[0] = in[0];
    tri.i[1] = in[1];
    tri.i[2] = in[2];
    for (size_t k = 3; k < in.size(); k++) {
      Prim p;
      p.i[0] = in[0];
      p.i[1] = in[k - 1];
      p.i[2] = in[k];
      p.front = tri.front;
      append.push_back(p);
    }
  }

  float *depthPtr(int x, int y, int s) {
    if (!dl || (size_t)x >= (size_t)dl->w || (size_t)y >= (size_t)dl->h) return nullptr;
    size_t i = (size_t)y * dl->w + x;
    return dl->hasMs ? &dl->fMs[i * 4 + s] : &dl->f[i];
  }
  uint32_t *colorPtr(int x, int y, int s) {
    if (!cl || (size_t)x >= (size_t)cl->w || (size_t)y >= (size_t)cl->h) return nullptr;
    size_t i = (size_t)y * cl->w + x;
    return cl->hasMs ? &cl->rgbaMs[i * 4 + s] : &cl->rgba[i];
  }

  /* processDepthTest, RendererSoft.cpp:377-395 */
  bool depthTest(int x, int y, float z, int s, bool skipWrite) {
    if (!ctx.st.depthTest || !dl) return true;
    z = clampf(z, 0.f, 1.f);
    float *zp = depthPtr(x, y, s);
    if (zp && depthFuncPass(z, *zp, ctx.st.depthFunc)) {
      if (!skipWrite && ctx.st.depthMask) *zp = z;
      return true;
    }
    return false;
  }

  /* processPerSampleOperations, RendererSoft.cpp:358-375 */
  void perSample(int x, int y, float z, V4 col, int s, long long *st) {
    if (!depthTest(x, y, z, s, false)) return;
    if (!cl) return;
    V4 c{clampf(col.x, 0.f, 1.f), clampf(col.y, 0.f, 1.f), clampf(col.z, 0.f, 1.f), clampf(col.w, 0.f, 1.f)};
    uint32_t *cp = colorPtr(x, y, s);
    if (ctx.st.blend) {
      V4 d{0, 0, 0, 0};
      if (cp) d = {(float)chan(*cp, 0) / 255.f, (float)chan(*cp, 1) / 255.f, (float)chan(*cp, 2) / 255.f,
                   (float)chan(*cp, 3) / 255.f};
      c = blendColor(c, d, ctx.st);
    }
    if (cp) {
      *cp = packRGBA(u8trunc(c.x * 255.f), u8trunc(c.y * 255.f), u8trunc(c.z * 255.f), u8trunc(c.w * 255.f));
      st[1]++;
    }
  }

  /* barycentric, RendererSoft.cpp:1021-1056 (SSE path: vmulps + vfmsub231ps, vdivps) */
  static bool bary(const V4 &vx, const V4 &vy, V4 v0, float px, float py, V4 &bc) {
    float a0 = vx.x - v0.x, a1 = vx.y - v0.x, a2 = vx.z - px;
    float b0 = vy.x - v0.y, b1 = vy.y - v0.y, b2 = vy.z - py;
    float ux = fmaf(a1, b2, -(a2 * b1));
    float uy = fmaf(a2, b0, -(a0 * b2));
    float uz = fmaf(a0, b1, -(a1 * b0));
    if (fabsf(uz) < 1.19209290e-07f) return false;
    ux = ux / uz;
    uy = uy / uz;
    bc = {1.f - (ux + uy), uy, ux, 0.f};
    if (bc.x < 0 || bc.y < 0 || bc.z < 0) return false;
    return true;
  }

  /* dpps with imm 0xff: (b0*v0 + b1*v1) + (b2*v2 + b3*0) */
  static inline float dpps3(V4 bc, float v0, float v1, float v2) {
    float p0 = bc.x * v0, p1 = bc.y * v1, p2 = bc.z * v2, p3 = bc.w * 0.f;
    return (p0 + p1) + (p2 + p3);
  }

  /* rasterizationTriangle + rasterizationPixelQuad + earlyZTest, RendererSoft.cpp:720-878.
   * Rows outside [rowLo, rowHi) belong to another worker and are neither tested nor written. */
  void rasterTriangle(const Prim &tri, int rowLo, int rowHi, long long *st) {
    const VertexHolder *v[3] = {&verts[tri.i[0]], &verts[tri.i[1]], &verts[tri.i[2]]};
    V4 sp[3] = {v[0]->frag, v[1]->frag, v[2]->frag};
    float minX = std::min(std::min(sp[0].x, sp[1].x), sp[2].x);
    float minY = std::min(std::min(sp[0].y, sp[1].y), sp[2].y);
    float maxX = std::max(std::max(sp[0].x, sp[1].x), sp[2].x);
    float maxY = std::max(std::max(sp[0].y, sp[1].y), sp[2].y);
    minX = std::max(minX - 0.5f, 0.f);
    minY = std::max(minY - 0.5f, 0.f);
    maxX = std::min(maxX + 0.5f, vpW - 1.f);
    maxY = std::min(maxY + 0.5f, vpH - 1.f);
    minX -= 1.f;
    minY -= 1.f;
    const int bs = 32;
    int bcx = (int)((maxX - minX + bs - 1.f) / bs);
    int bcy = (int)((maxY - minY + bs - 1.f) / bs);
    V4 vx{sp[2].x, sp[1].x, sp[0].x, 0.f}, vy{sp[2].y, sp[1].y, sp[0].y, 0.f};
    V4 vw{sp[0].w, sp[1].w, sp[2].w, 1.f};
    static const float loc[4][2] = {{0.375f, 0.875f}, {0.875f, 0.625f}, {0.125f, 0.375f}, {0.625f, 0.125f}};
    Pixel px[4];
    for (auto &p : px) p.vary.assign(std::max(ctx.nvary, 1), 0.f);
    const float *vv[3] = {v[0]->vary.data(), v[1]->vary.data(), v[2]->vary.data()};
    for (int by = 0; by < bcy; by++) {
      for (int bx = 0; bx < bcx; bx++) {
        int sx = (int)(minX + (float)(bx * bs));
        int sy = (int)(minY + (float)(by * bs));
        for (int y = sy + 1; y < sy + bs && (float)y <= maxY; y += 2) {
          if (y + 1 < rowLo || y >= rowHi) continue;
          for (int x = sx + 1; x < sx + bs && (float)x <= maxX; x += 2) {
            st[5]++;
            /* PixelQuadContext::Init */
            for (int k = 0; k < 4; k++) {
              Pixel &p = px[k];
              float fx = (float)(x + (k & 1)), fy = (float)(y + (k >> 1));
              p.inside = false;
              if (samples == 4) {
                p.nsamp = 5;
                for (int s = 0; s < 4; s++) {
                  p.s[s].fx = (int)fx;
                  p.s[s].fy = (int)fy;
                  p.s[s].pos = {loc[s][0] + fx, loc[s][1] + fy, 0.f, 0.f};
                }
                p.s[4].fx = (int)fx;
                p.s[4].fy = (int)fy;
                p.s[4].pos = {fx + 0.5f, fy + 0.5f, 0.f, 0.f};
                p.shading = 4;
              } else {
                p.nsamp = 1;
                p.s[0].fx = (int)fx;
                p.s[0].fy = (int)fy;
                p.s[0].pos = {fx + 0.5f, fy + 0.5f, 0.f, 0.f};
                p.shading = 0;
              }
            }
            bool any = false;
            for (int k = 0; k < 4; k++) {
              Pixel &p = px[k];
              for (int s = 0; s < p.nsamp; s++) p.s[s].inside = bary(vx, vy, sp[0], p.s[s].pos.x, p.s[s].pos.y, p.s[s].bc);
              if (samples == 4) {
                int cov = 0;
                for (int s = 0; s < 4; s++) cov += p.s[s].inside ? 1 : 0;
                p.inside = cov > 0;
              } else {
                p.inside = p.s[0].inside;
              }
              if (!p.s[p.shading].inside)
                for (int s = 0; s < p.nsamp; s++)
                  if (p.s[s].inside) {
                    p.shading = s;
                    break;
                  }
              bool mine = (y + (k >> 1)) >= rowLo && (y + (k >> 1)) < rowHi;
              any = any || (p.inside && mine);
            }
            if (!any) continue;
            for (int k = 0; k < 4; k++)
              for (int s = 0; s < px[k].nsamp; s++) {
                Sample &q = px[k].s[s];
                if (!q.inside) continue;
                q.pos.z = dpps3(q.bc, sp[0].z, sp[1].z, sp[2].z);
                q.pos.w = dpps3(q.bc, sp[0].w, sp[1].w, sp[2].w);
                if (q.pos.z < 0.f || q.pos.z > 1.f) q.inside = false;
                float iw = 1.f / q.pos.w;
                q.bc = {q.bc.x * (iw * vw.x), q.bc.y * (iw * vw.y), q.bc.z * (iw * vw.z), q.bc.w * (iw * vw.w)};
              }
            if (ctx.st.depthTest) {
              bool anyIn = false;
              for (int k = 0; k < 4; k++) {
                Pixel &p = px[k];
                bool mine = (y + (k >> 1)) >= rowLo && (y + (k >> 1)) < rowHi;
                if (!p.inside) continue;
                if (!mine) {
                  p.inside = false;
                  continue;
                }
                if (samples == 4) {
                  bool in = false;
                  for (int s = 0; s < 4; s++) {
                    Sample &q = p.s[s];
                    if (!q.inside) continue;
                    q.inside = depthTest(q.fx, q.fy, q.pos.z, s, true);
                    in = in || q.inside;
                  }
                  p.inside = in;
                } else {
                  Sample &q = p.s[p.shading];
                  q.inside = depthTest(q.fx, q.fy, q.pos.z, 0, true);
                  p.inside = q.inside;
                }
                anyIn = anyIn || p.inside;
              }
              if (!anyIn) continue;
            }
            /* varyings for all four pixels at their shading sample (RendererSoft.cpp:813-820) */
            for (int k = 0; k < 4; k++) {
              V4 b = px[k].s[px[k].shading].bc;
              float *o = px[k].vary.data();
              for (int i = 0; i < ctx.nvary; i++) o[i] = fmaf(vv[2][i], b.z, fmaf(vv[1][i], b.y, vv[0][i] * b.x));
            }
            if (g_trace_x >= 0)
              for (int k = 0; k < 4; k++)
                if (x + (k & 1) == g_trace_x && y + (k >> 1) == g_trace_y)
                  fprintf(stderr, "trace tri v(%zu,%zu,%zu) quad (%d,%d) block (%d,%d) bb %.9g %.9g %.9g %.9g bc %d %d k=%d inside=%d shade=%d s_in=%d%d%d%d\n",
                          tri.i[0], tri.i[1], tri.i[2], x, y, bx, by, minX, minY, maxX, maxY, bcx, bcy, k, (int)px[k].inside, px[k].shading,
                          (int)px[k].s[0].inside, (int)px[k].s[1].inside, (int)px[k].s[2].inside, (int)px[k].s[3].inside);
            FragIn fi;
            for (int k = 0; k < 4; k++) fi.v[k] = px[k].vary.data();
            fi.front = tri.front;
            for (int k = 0; k < 4; k++) {
              Pixel &p = px[k];
              bool mine = (y + (k >> 1)) >= rowLo && (y + (k >> 1)) < rowHi;
              if (!p.inside || !mine) continue;
              V4 col{0, 0, 0, 0};
              if (cl) {
                fi.self = k;
                col = runFS(ctx, fi);
                st[0]++;
              }
              if (samples == 4) {
                for (int s = 0; s < 4; s++) {
                  Sample &q = p.s[s];
                  if (!q.inside) continue;
                  perSample(q.fx, q.fy, q.pos.z, col, s, st);
                }
              } else {
                Sample &q = p.s[p.shading];
                perSample(q.fx, q.fy, q.pos.z, col, 0, st);
              }
            }
          }
        }
      }
    }
  }

  /* ---------------------------------------------------------------- points and lines (RendererSoft.cpp:575-718) */

  /* gl->PointSize after the vertex shader: BasicSoft.h:68 sets u_pointSize; the others leave the default 1 */
  float pointSize() const { return ctx.shading == RHS_SHADING_BASIC ? ctx.u.pointSize : 1.0f; }

  /* the fragment shader of a point or line fragment (BasicSoft.h:78-80; the vertex colour shader: its varying) */
  V4 shadeFlat(const std::vector<float> &vary) const {
    if (ctx.shading == RHS_SHADING_BASIC) return ctx.u.baseColor;
    return {vary[0], vary[1], vary[2], vary[3]};
  }

  /* perspectiveDivideImpl + viewportTransformImpl, RendererSoft.cpp:981-992 */
  void toScreen(VertexHolder &v) {
    float iw = 1.f / v.clip.w;
    V4 f{v.clip.x * iw, v.clip.y * iw, v.clip.z * iw, v.clip.w * iw};
    f.w = iw;
    V4 g{f.x * pX, f.y * pY, f.z * 1.f, f.w * 1.f};
    v.frag = {g.x + oX, g.y + oY, g.z + 0.f, g.w + 0.f};
  }

  /* clippingLine, RendererSoft.cpp:451-487. The second endpoint is interpolated from the line's first index
   * after that index may already have been replaced by the first clip vertex, with t1 of the original line. */
  bool clipLine(size_t &i0, size_t &i1, bool post) {
    const int m0 = verts[i0].mask, m1 = verts[i1].mask;
    const V4 c0 = verts[i0].clip, c1 = verts[i1].clip;
    float t0 = 0.f, t1 = 1.f;
    const int mask = m0 | m1;
    if (mask != 0) {
      for (int pl = 0; pl < 6; pl++) {
        if (!(mask & (1 << pl))) continue;
        float d0 = planeDist(pl, c0), d1 = planeDist(pl, c1);
        if (d0 < 0 && d1 < 0) return false;
        if (d0 < 0) {
          float t = -d0 / (d1 - d0);
          t0 = t0 < t ? t : t0;
        } else {
          float t = d0 / (d0 - d1);
          t1 = t < t1 ? t : t1;
        }
      }
    }
    if (m0) {
      i0 = newClipVertex(i0, i1, t0);
      if (post) toScreen(verts[i0]);
    }
    if (m1) {
      i1 = newClipVertex(i0, i1, t1);
      if (post) toScreen(verts[i1]);
    }
    return true;
  }

  /* rasterizationPoint, RendererSoft.cpp:637-661. FragCoord is the integer pixel (x, y), the depth the
   * point's z; every sample gets the fragment. The loop writes (x, y) into the point's own screen position, so
   * a vertex rasterized again starts from where the last one ended. No colour attachment: nothing at all. */
  /* screen coordinates a point or line may start from (the device's rule): finite and below 2^24; a vertex
   * inside the frustum can still sit at w = 0, and the reference's int conversion of infinity is undefined */
  static bool sane(V4 p) { return fabsf(p.x) < 16777216.f && fabsf(p.y) < 16777216.f; }

  void rasterPoint(V4 &frag, const std::vector<float> &vary, float size) {
    if (!cl) return;
    float left = frag.x - size / 2.f + 0.5f, right = left + size;
    float top = frag.y - size / 2.f + 0.5f, bottom = top + size;
    const V4 col = shadeFlat(vary);
    long long st[8] = {};
    for (int x = (int)left; x < (int)right; x++)
      for (int y = (int)top; y < (int)bottom; y++) {
        frag.x = (float)x;
        frag.y = (float)y;
        for (int smp = 0; smp < samples; smp++) perSample(x, y, frag.z, col, smp, st);
      }
  }

  /* rasterizationLine, RendererSoft.cpp:663-718: Bresenham over the major axis of the truncated endpoints,
   * z, w and the varyings interpolated linearly in t = (x - x0) / dx (NaN for a line within one column),
   * every step a point of the line width. */
  void rasterLine(size_t ia, size_t ib, float width) {
    const VertexHolder *va = &verts[ia], *vb = &verts[ib];
    int x0 = (int)va->frag.x, y0 = (int)va->frag.y, x1 = (int)vb->frag.x, y1 = (int)vb->frag.y;
    float z0 = va->frag.z, z1 = vb->frag.z, w0 = va->frag.w, w1 = vb->frag.w;
    bool steep = false;
    if (std::abs(x0 - x1) < std::abs(y0 - y1)) {
      std::swap(x0, y0);
      std::swap(x1, y1);
      steep = true;
    }
    const float *vin[2] = {va->vary.data(), vb->vary.data()};
    if (x0 > x1) {
      std::swap(x0, x1);
      std::swap(y0, y1);
      std::swap(z0, z1);
      std::swap(w0, w1);
      std::swap(vin[0], vin[1]);
    }
    const int dx = x1 - x0, dy = y1 - y0;
    int error = 0;
    const int dError = 2 * std::abs(dy);
    int y = y0;
    std::vector<float> vary(std::max(ctx.nvary, 1), 0.f);
    for (int x = x0; x <= x1; x++) {
      const float t = (float)(x - x0) / (float)dx;
      const float omt = 1.0f - t;
      /* glm::mix as compiled there and in interpolateLinear: fma(a, 1 - t, b t) */
      V4 frag{(float)x, (float)y, fmaf(z0, omt, z1 * t), fmaf(w0, omt, w1 * t)};
      if (steep) std::swap(frag.x, frag.y);
      for (int k = 0; k < ctx.nvary; k++) vary[k] = fmaf(vin[0][k], omt, vin[1][k] * t);
      rasterPoint(frag, vary, width);
      error += dError;
      if (error > dx) {
        y += (y1 > y0 ? 1 : -1);
        error -= 2 * dx;
      }
    }
  }

  /* RendererSoft::draw for points, lines and triangles in point or line mode (RendererSoft.cpp:136-164:
   * processPrimitiveAssembly :192-204, processClipping :206-257, processPerspectiveDivide/ViewportTransform
   * :259-275, processFaceCulling :277-299, processRasterization :301-321 and rasterizationPolygons :560-626) */
  void drawPointsLines() {
    const int type = ctx.st.primitiveType, mode = ctx.st.polygonMode;
    const int per = type == 0 ? 1 : (type == 1 ? 2 : 3);
    const size_t np = vao->idx.size() / per;
    prims.assign(np, Prim());
    for (size_t t = 0; t < np; t++)
      for (int k = 0; k < per; k++) prims[t].i[k] = (size_t)vao->idx[t * per + k];
    for (auto &pr : prims) {
      if (type == 0)
        pr.discard = verts[pr.i[0]].mask != 0;
      else if (type == 1)
        pr.discard = !clipLine(pr.i[0], pr.i[1], false);
      /* triangles in point or line mode are not clipped here */
    }
    for (auto &v : verts) v.discard = true;
    for (auto &pr : prims)
      if (!pr.discard)
        for (int k = 0; k < per; k++) verts[pr.i[k]].discard = false;
    for (auto &v : verts)
      if (!v.discard) toScreen(v);
    if (type == 2)
      for (auto &pr : prims) {
        if (pr.discard) continue;
        V4 a = verts[pr.i[0]].frag, b = verts[pr.i[1]].frag, c = verts[pr.i[2]].frag;
        float e0x = b.x - a.x, e0y = b.y - a.y, e1x = c.x - a.x, e1y = c.y - a.y;
        float nz = fmaf(e0x, e1y, -(e0y * e1x));
        pr.front = nz > 0;
        if (ctx.st.cullFace) pr.discard = !pr.front;
      }
    const float psize = pointSize(), lwidth = ctx.st.lineWidth;
    for (size_t t = 0; t < prims.size(); t++) {
      if (prims[t].discard) continue;
      if (type == 0) {
        VertexHolder &v = verts[prims[t].i[0]];
        if (sane(v.frag)) rasterPoint(v.frag, v.vary, psize);
      } else if (type == 1) {
        if (sane(verts[prims[t].i[0]].frag) && sane(verts[prims[t].i[1]].frag))
          rasterLine(prims[t].i[0], prims[t].i[1], lwidth);
      } else if (mode == 0) {
        for (int k = 0; k < 3; k++) {
          const size_t vi = prims[t].i[k];
          if (verts[vi].mask != 0) continue; /* clippingPoint */
          if (sane(verts[vi].frag)) rasterPoint(verts[vi].frag, verts[vi].vary, psize);
        }
      } else {
        for (int k = 0; k < 3; k++) {
          size_t a = prims[t].i[k], b = prims[t].i[(k + 1) % 3];
          if (!clipLine(a, b, true)) continue;
          if (sane(verts[a].frag) && sane(verts[b].frag)) rasterLine(a, b, lwidth);
        }
      }
    }
  }

  /* multiSampleResolve, RendererSoft.cpp:880-912 */
  void resolve() {
    Level *c = cl;
    if (!c->hasSingle) {
      c->rgba.assign((size_t)c->w * c->h, 0);
      c->hasSingle = true;
    }
    for (size_t i = 0; i < (size_t)c->w * c->h; i++) {
      int r[4];
      for (int ch = 0; ch < 4; ch++) {
        float s = 0.f;
        for (int k = 0; k < 4; k++) s += (float)chan(c->rgbaMs[i * 4 + k], ch);
        r[ch] = u8trunc(s / 4.f);
      }
      c->rgba[i] = packRGBA(r[0], r[1], r[2], r[3]);
    }
  }

  /* RendererSoft::draw, RendererSoft.cpp:136-164 (triangles, fill mode) */
  int draw(const uint32_t *p, uint32_t n) {
    RhsDraw d;
    if (rhs_decode_draw(p, n, &d)) return -1;
    if (!fbo) return 0;
    auto vit = vaos.find(d.vao);
    if (vit == vaos.end()) return -1;
    vao = &vit->second;
    ctx = DrawCtx();
    ctx.shading = d.shading;
    if (d.shading < 0 || d.shading > RHS_SHADING_FXAA || d.shading == RHS_SHADING_NEURAL_PBR) return -1;
    ctx.defines = d.defines;
    ctx.st = d.states;
    parseUniforms(d, ctx.u);
    for (int i = 0; i < d.nsamplers; i++) {
      auto it = tex.find(d.tex[i]);
      if (it == tex.end() || d.slot[i] < 0 || d.slot[i] >= RHS_SLOT_COUNT) return -1;
      ctx.tex[d.slot[i]] = &it->second;
    }
    const bool fill = ctx.st.primitiveType == 2 && ctx.st.polygonMode == 2;
    /* points and lines shade without a derivative context; the reference's Viewer draws them with the base
     * colour shader only (Viewer.cpp:264-275, ModelLoader.cpp:60-85) */
    if (!fill && d.shading != RHS_SHADING_BASIC && d.shading != RHS_SHADING_VERTEX_COLOR) return -1;
    if (ctx.st.primitiveType < 0 || ctx.st.primitiveType > 2 || ctx.st.polygonMode < 0 || ctx.st.polygonMode > 2)
      return -1;
    ctx.nvary = varyingCount(d.shading);
    g_stats[7]++;
    cl = colorLevel();
    dl = depthLevel();
    if (cl)
      samples = cl->hasMs ? 4 : 1;
    else if (dl)
      samples = dl->hasMs ? 4 : 1;
    else
      samples = 1;

    /* processVertexShader :170-190 */
    int nf = vao->stride / 4;
    verts.clear();
    verts.resize(vao->nverts);
    for (int i = 0; i < vao->nverts; i++) {
      VertexHolder &vh = verts[i];
      vh.rawPtr = vao->verts.data() + (size_t)i * nf;
      vh.vary.assign(std::max(ctx.nvary, 1), 0.f);
      runVS(ctx, vh.rawPtr, vh);
      vh.mask = clipMask(vh.clip);
    }
    if (!fill) {
      drawPointsLines();
      if (cl && cl->hasMs) resolve();
      return 0;
    }
    /* processPolygonAssembly :428-437 */
    size_t np = vao->idx.size() / 3;
    prims.assign(np, Prim());
    for (size_t t = 0; t < np; t++)
      for (int k = 0; k < 3; k++) prims[t].i[k] = (size_t)vao->idx[t * 3 + k];
    /* processClipping :206-257 */
    for (size_t t = 0; t < np; t++) {
      if (prims[t].discard) continue;
      std::vector<Prim> app;
      clipTriangle(prims[t], app);
      prims.insert(prims.end(), app.begin(), app.end());
    }
    for (auto &v : verts) v.discard = true;
    for (auto &pr : prims)
      if (!pr.discard)
        for (int k = 0; k < 3; k++) verts[pr.i[k]].discard = false;
    /* perspective divide + viewport :259-275, 981-992 */
    for (auto &v : verts) {
      if (v.discard) continue;
      float iw = 1.f / v.clip.w;
      V4 f{v.clip.x * iw, v.clip.y * iw, v.clip.z * iw, v.clip.w * iw};
      f.w = iw;
      V4 g{f.x * pX, f.y * pY, f.z * 1.f, f.w * 1.f};
      v.frag = {g.x + oX, g.y + oY, g.z + 0.f, g.w + 0.f};
    }
    /* processFaceCulling :277-299 */
    for (auto &pr : prims) {
      if (pr.discard) continue;
      V4 a = verts[pr.i[0]].frag, b = verts[pr.i[1]].frag, c = verts[pr.i[2]].frag;
      float e0x = b.x - a.x, e0y = b.y - a.y, e1x = c.x - a.x, e1y = c.y - a.y;
      float nz = fmaf(e0x, e1y, -(e0y * e1x));
      pr.front = nz > 0;
      if (ctx.st.cullFace) {
        pr.discard = !pr.front;
        if (pr.discard) g_stats[4]++;
      }
    }
    /* processRasterization :301-340 */
    int H = (int)vpH;
    int nt = g_touch ? 1 : std::max(1, std::min(g_threads, H / 32));
    if (g_touch)
      for (int s = 0; s < RHS_SLOT_COUNT; s++)
        if (ctx.tex[s])
          for (auto &layer : ctx.tex[s]->layers)
            for (auto &l : layer) l.touched.clear();
    std::vector<long long> st((size_t)nt * 8, 0);
    auto work = [&](int w) {
      int lo = (H * w / nt) & ~1, hi = w == nt - 1 ? 1 << 30 : (H * (w + 1) / nt) & ~1;
      if (w == 0) lo = -(1 << 30);
      for (auto &pr : prims) {
        if (pr.discard) continue;
        if (w == 0) st[w * 8 + 2]++;
        rasterTriangle(pr, lo, hi, &st[w * 8]);
      }
    };
    if (nt == 1) {
      work(0);
    } else {
      std::vector<std::thread> th;
      for (int w = 0; w < nt; w++) th.emplace_back(work, w);
      for (auto &t : th) t.join();
    }
    for (int w = 0; w < nt; w++)
      for (int k = 0; k < 6; k++) g_stats[k] += st[w * 8 + k];
    if (g_touch) {
      long long texels = 0;
      std::vector<const Texture *> seen;
      for (int s = 0; s < RHS_SLOT_COUNT; s++) {
        const Texture *t = ctx.tex[s];
        if (!t || std::find(seen.begin(), seen.end(), t) != seen.end()) continue;
        seen.push_back(t);
        for (auto &layer : t->layers)
          for (auto &l : layer)
            for (uint8_t b : l.touched) texels += b;
      }
      g_draw_texel_bytes.push_back(4 * texels);
    }
    if (cl && cl->hasMs) resolve();
    return 0;
  }

  int exec(int op, const uint32_t *p, uint32_t n, std::vector<uint8_t> &out) {
    switch (op) {
      case RHS_TEXTURE: {
        Texture t;
        t.type = (int)p[1];
        t.format = (int)p[2];
        t.w = (int)p[3];
        t.h = (int)p[4];
        t.usage = (int)p[5];
        t.mips = p[6] != 0;
        t.msaa = p[7] != 0;
        t.filterMin = (int)p[8];
        t.filterMag = (int)p[9];
        t.wrapS = (int)p[10];
        t.border = (int)p[13];
        t.layers.resize(t.type == 1 ? 6 : 1);
        tex[(int)p[0]] = std::move(t);
        return 0;
      }
      case RHS_TEX_INIT: {
        /* TextureSoft::initImageData :136-144 (mip levels allocated, not filled) */
        Texture &t = tex.at((int)p[0]);
        for (auto &lv : t.layers) {
          lv.assign(1, Level());
          Level &l = lv[0];
          l.w = t.w;
          l.h = t.h;
          if (t.msaa) {
            l.hasMs = true;
            if (t.format == 0) l.rgbaMs.assign((size_t)t.w * t.h * 4, 0);
            else l.fMs.assign((size_t)t.w * t.h * 4, 0.f);
          } else {
            l.hasSingle = true;
            if (t.format == 0) l.rgba.assign((size_t)t.w * t.h, 0);
            else l.f.assign((size_t)t.w * t.h, 0.f);
          }
          if (t.mips) {
            int cnt = (int)floor(log2((double)std::max(t.w, t.h))) + 1;
            for (int i = 1; i < cnt; i++) {
              Level m;
              m.w = std::max(1, t.w >> i);
              m.h = std::max(1, t.h >> i);
              m.hasSingle = true;
              if (t.format == 0) m.rgba.assign((size_t)m.w * m.h, 0);
              else m.f.assign((size_t)m.w * m.h, 0.f);
              lv.push_back(std::move(m));
            }
          }
        }
        return 0;
      }
      case RHS_TEX_DATA: {
        /* TextureSoft::setImageData :115-134 */
        Texture &t = tex.at((int)p[0]);
        int layers = (int)p[1];
        if (layers != (int)t.layers.size()) return -1;
        const uint8_t *src = (const uint8_t *)(p + 2);
        size_t per = (size_t)t.w * t.h * 4;
        if ((size_t)(n - 2) * 4 < per * layers) return -1;
        for (int l = 0; l < layers; l++) {
          auto &lv = t.layers[l];
          lv.assign(1, Level());
          lv[0].w = t.w;
          lv[0].h = t.h;
          lv[0].hasSingle = true;
          if (t.format == 0) {
            lv[0].rgba.resize((size_t)t.w * t.h);
            memcpy(lv[0].rgba.data(), src + per * l, per);
            if (t.mips) buildMips<uint32_t>(t, l);
          } else {
            lv[0].f.resize((size_t)t.w * t.h);
            memcpy(lv[0].f.data(), src + per * l, per);
            if (t.mips) buildMips<float>(t, l);
          }
        }
        return 0;
      }
      case RHS_TEX_LEVEL: {
        Texture &t = tex.at((int)p[0]);
        int layer = (int)p[1], level = (int)p[2];
        if (layer >= (int)t.layers.size() || level >= (int)t.layers[layer].size()) return -1;
        Level &l = t.layers[layer][level];
        size_t bytes = (size_t)l.w * l.h * 4;
        if ((size_t)(n - 3) * 4 < bytes) return -1;
        if (t.format == 0) memcpy(l.rgba.data(), p + 3, bytes);
        else memcpy(l.f.data(), p + 3, bytes);
        return 0;
      }
      case RHS_VAO: {
        VAO v;
        v.stride = (int)p[1];
        v.nverts = (int)p[2];
        size_t ni = p[3];
        size_t vw = ((size_t)v.nverts * v.stride + 3) / 4;
        if (v.stride % 4 || 4 + vw + ni > n) return -1;
        v.verts.assign((const float *)(p + 4), (const float *)(p + 4) + (size_t)v.nverts * v.stride / 4);
        v.idx.assign((const int32_t *)(p + 4 + vw), (const int32_t *)(p + 4 + vw) + ni);
        for (int32_t i : v.idx)
          if (i < 0 || i >= v.nverts) return -1;
        vaos[(int)p[0]] = std::move(v);
        return 0;
      }
      case RHS_FBO: {
        FBO f;
        f.color = (int)p[1];
        f.colorLayer = (int)p[2];
        f.colorLevel = (int)p[3];
        f.depth = (int)p[4];
        fbos[(int)p[0]] = f;
        return 0;
      }
      case RHS_BEGIN:
        begin(p);
        return 0;
      case RHS_VIEWPORT:
        viewport(p);
        return 0;
      case RHS_DRAW:
        return draw(p, n);
      case RHS_END:
      case RHS_FRAME:
        return 0;
      case RHS_READ: {
        auto it = tex.find((int)p[0]);
        if (it == tex.end()) return -1;
        Texture &t = it->second;
        Level &l = t.layers.at(p[1]).at(p[2]);
        const void *src;
        size_t bytes;
        if (p[3] == 1) {
          src = t.format == 0 ? (const void *)l.rgbaMs.data() : (const void *)l.fMs.data();
          bytes = (size_t)l.w * l.h * 16;
        } else {
          src = t.format == 0 ? (const void *)l.rgba.data() : (const void *)l.f.data();
          bytes = (size_t)l.w * l.h * 4;
          if (!l.hasSingle) return -1;
        }
        size_t off = out.size();
        out.resize(off + bytes);
        memcpy(out.data() + off, src, bytes);
        return 0;
      }
    }
    return -1;
  }
};

}  // namespace cref

extern "C" {

void cpu_ref_set_math(int correctly_rounded) { cref::g_cr_math = correctly_rounded != 0; }

void cpu_ref_set_threads(int n) {
  cref::g_threads = n < 1 ? 1 : n;
  const char *t = getenv("CPU_REF_TRACE");
  if (t) sscanf(t, "%d,%d", &cref::g_trace_x, &cref::g_trace_y);
}

void cpu_ref_stats(long long *out8) { memcpy(out8, cref::g_stats, sizeof(cref::g_stats)); }

/* Unique-texel counting for the next replays (the raster then runs on one thread). */
void cpu_ref_count_texels(int on) { cref::g_touch = on != 0; }

/* Bytes of distinct texels each draw of the last frame read (4 per texel), in draw order; returns the number
 * of draws, or -1 when `cap` is too small. */
int cpu_ref_draw_texel_bytes(long long *out, int cap) {
  const int n = (int)cref::g_draw_texel_bytes.size();
  if (n > cap) return -1;
  for (int i = 0; i < n; i++) out[i] = cref::g_draw_texel_bytes[i];
  return n;
}

long long cpu_ref_replay(const void *stream, size_t bytes, void *out, size_t out_cap, int repeats, double *frame_ms) {
  RhsReader rd;
  if (rhs_open(&rd, stream, bytes)) return -1;
  std::unique_ptr<cref::Renderer> r(new cref::Renderer());
  std::vector<uint8_t> result;
  struct Cmd {
    int op;
    const uint32_t *p;
    uint32_t n;
  };
  std::vector<Cmd> frame, reads;
  bool inFrame = false;
  const uint32_t *p;
  uint32_t n;
  int op;
  while ((op = rhs_next(&rd, &p, &n)) > 0) {
    if (op == RHS_FRAME) {
      inFrame = true;
      continue;
    }
    if (op == RHS_READ) {
      reads.push_back({op, p, n});
      continue;
    }
    if (inFrame)
      frame.push_back({op, p, n});
    else if (r->exec(op, p, n, result))
      return -1;
  }
  if (op < 0) return -1;
  if (repeats < 1) repeats = 1;
  for (int i = 0; i < repeats; i++) {
    memset(cref::g_stats, 0, sizeof(cref::g_stats));
    cref::g_draw_texel_bytes.clear();
    auto t0 = std::chrono::steady_clock::now();
    for (auto &c : frame)
      if (r->exec(c.op, c.p, c.n, result)) return -1;
    auto t1 = std::chrono::steady_clock::now();
    if (frame_ms) frame_ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  for (auto &c : reads)
    if (r->exec(c.op, c.p, c.n, result)) return -1;
  if (result.size() > out_cap) return -2;
  memcpy(out, result.data(), result.size());
  return (long long)result.size();
}
}
