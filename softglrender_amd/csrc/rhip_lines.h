/*
 * rhip_lines.h -- points, lines, and triangles drawn in point or line mode
 * (RendererSoft.cpp:192-321 for those primitive types, :560-718).
 *
 * The reference rasterizes these one primitive after another, in draw order:
 * every point is a size x size block of fragments, every line a Bresenham walk
 * whose steps are points of the line's width, and each fragment goes through
 * the per-sample operations on every sample (no coverage: "TODO MSAA").
 * Here the work is split so that it runs in parallel and still lands in that
 * order on every pixel:
 *
 *   k_pl_units   one thread per point or line the reference would rasterize
 *                ("unit"): assembly, clipping (new vertices re-run the vertex
 *                shader), face culling of triangles in point/line mode; a line
 *                counts its fragments here
 *   k_pl_points  after the point units are sorted by vertex (stable, so in
 *                draw order): the reference writes each rasterized point's last
 *                pixel back into its vertex's screen position, so the k-th
 *                rasterization of a vertex starts where the (k-1)-th ended;
 *                each point unit replays that chain and counts its fragments
 *   (scan)       fragment offsets per unit: the fragments' draw order
 *   k_pl_emit    every unit writes its fragments (pixel key, depth, colour) at
 *                its offset, in the reference's loop order
 *   (sort)       a stable radix sort of the fragments by pixel: each pixel's
 *                fragments stay in draw order
 *   k_pl_apply   one thread per pixel runs its fragments through depth test,
 *                blending and the colour write, sample by sample
 *
 * Only the base colour and vertex colour shaders draw points and lines (they
 * need no derivatives; the reference's Viewer draws points and lines with the
 * base colour shader only: Viewer.cpp:264-275, ModelLoader.cpp:60-85).
 */
#pragma once

#include "rhip_kernels.h"

namespace rhip {

/* One rasterized point or line. */
struct PlUnit {
  float4 a, b;   /* screen positions (x, y, z, w): the point, or the line's endpoints */
  float4 ca, cb; /* fragment shader colour at a and b (vertex colour: the varyings) */
  int kind;      /* 0 nothing, 1 point, 2 line */
  int vertex;    /* point: the vertex whose screen position it starts from */
  int pad[2];
};

constexpr int kPlNone = 0, kPlPoint = 1, kPlLine = 2;

/* Points, lines, or triangles in point or line mode: the units a draw rasterizes. */
__host__ __device__ inline long long plUnitCount(int type, int nidx) {
  return type == 0 ? nidx : (type == 1 ? nidx / 2 : (long long)(nidx / 3) * 3);
}

/* The fragment shader's colour from a vertex's varyings (BasicSoft.h:78-80: the base colour). */
__device__ __forceinline__ float4 plColour(const DrawParams &P, const float *v) {
  if (P.shading == RHS_SHADING_BASIC)
    return make_float4(P.u.base_color[0], P.u.base_color[1], P.u.base_color[2], P.u.base_color[3]);
  return make_float4(v[0], v[1], v[2], v[3]);
}

__device__ __forceinline__ float4 plVertexColour(const DrawParams &P, int i) {
  float v[4] = {0.f, 0.f, 0.f, 0.f};
  if (P.vary_stride >= 4) {
    const float4 q = *(const float4 *)(P.vary + (size_t)i * P.vary_stride);
    v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
  }
  return plColour(P, v);
}

/* The raw attributes of original vertex i (up to 16 floats). */
__device__ __forceinline__ void plRaw(const DrawParams &P, int i, float *a) {
  const float *src = P.verts + (size_t)i * P.stride_f;
  for (int k = 0; k < 16; k++) a[k] = k < P.stride_f ? src[k] : 0.f;
}

/* clippingNewVertex + interpolateVertex (RendererSoft.cpp:956-970, :1058-1070): the raw attributes
 * interpolated as fma(a, 1 - t, b t), then the vertex shader, perspective divide and viewport. */
__device__ __forceinline__ void plNewVertex(const DrawParams &P, float *ra, const float *rb, float t, float4 &frag,
                                            float4 &col) {
  const float omt = 1.0f - t;
  for (int k = 0; k < 16; k++) ra[k] = k < P.stride_f ? fmaf(ra[k], omt, rb[k] * t) : 0.f;
  float v[kVaryStride];
  for (int k = 0; k < kVaryStride; k++) v[k] = 0.f;
  const float4 c = runVS(P, ra, v);
  frag = toScreen(P, c);
  col = plColour(P, v);
}

/* clippingLine (RendererSoft.cpp:451-487) of the line (i0, i1). Past the plane tests, endpoint 0 is
 * replaced when it was outside and endpoint 1 is then interpolated between the new endpoint 0 and
 * vertex i1 with t1 of the original line, as the reference does. */
__device__ bool plClipLine(const DrawParams &P, int i0, int i1, PlUnit &U) {
  const int m0 = P.mask[i0], m1 = P.mask[i1];
  const float4 c0 = P.clip[i0], c1 = P.clip[i1];
  float t0 = 0.f, t1 = 1.f;
  const int mask = m0 | m1;
  if (mask) {
    for (int pl = 0; pl < 6; pl++) {
      if (!((mask >> pl) & 1)) continue;
      const float d0 = planeDist(pl, c0), d1 = planeDist(pl, c1);
      if (d0 < 0 && d1 < 0) return false;
      if (d0 < 0) {
        const float t = -d0 / (d1 - d0);
        t0 = t0 < t ? t : t0;
      } else {
        const float t = d0 / (d0 - d1);
        t1 = t < t1 ? t : t1;
      }
    }
  }
  U.a = P.frag[i0];
  U.b = P.frag[i1];
  U.ca = plVertexColour(P, i0);
  U.cb = plVertexColour(P, i1);
  if (m0 | m1) {
    float ra[16], rb[16];
    plRaw(P, i0, ra);
    plRaw(P, i1, rb);
    if (m0) plNewVertex(P, ra, rb, t0, U.a, U.ca);
    if (m1) plNewVertex(P, ra, rb, t1, U.b, U.cb); /* ra: the new endpoint 0 when it was replaced */
  }
  return true;
}

/* rasterizationPoint's block (RendererSoft.cpp:642-645): columns [x0, x1) and rows [y0, y1). */
__device__ __forceinline__ void plBox(float px, float py, float size, int &x0, int &x1, int &y0, int &y1) {
  const float left = px - size / 2.f + 0.5f, right = left + size;
  const float top = py - size / 2.f + 0.5f, bottom = top + size;
  x0 = (int)left;
  x1 = (int)right;
  y0 = (int)top;
  y1 = (int)bottom;
}

__device__ __forceinline__ long long plBoxCount(int x0, int x1, int y0, int y1) {
  return (x1 > x0 && y1 > y0) ? (long long)(x1 - x0) * (y1 - y0) : 0;
}

/* The Bresenham walk of rasterizationLine (RendererSoft.cpp:663-718). */
struct PlWalk {
  int x0, x1, y0, y1, dx, dError, error, y, x;
  bool steep, swapped;
};

__device__ __forceinline__ PlWalk plWalkStart(const PlUnit &U) {
  PlWalk w;
  w.x0 = (int)U.a.x, w.y0 = (int)U.a.y, w.x1 = (int)U.b.x, w.y1 = (int)U.b.y;
  w.steep = false;
  if (abs(w.x0 - w.x1) < abs(w.y0 - w.y1)) {
    int t = w.x0;
    w.x0 = w.y0, w.y0 = t;
    t = w.x1;
    w.x1 = w.y1, w.y1 = t;
    w.steep = true;
  }
  w.swapped = w.x0 > w.x1;
  if (w.swapped) {
    int t = w.x0;
    w.x0 = w.x1, w.x1 = t;
    t = w.y0;
    w.y0 = w.y1, w.y1 = t;
  }
  w.dx = w.x1 - w.x0;
  w.dError = 2 * abs(w.y1 - w.y0);
  w.error = 0;
  w.y = w.y0;
  w.x = w.x0;
  return w;
}

__device__ __forceinline__ void plWalkStep(PlWalk &w) {
  w.error += w.dError;
  if (w.error > w.dx) {
    w.y += (w.y1 > w.y0 ? 1 : -1);
    w.error -= 2 * w.dx;
  }
  w.x++;
}

/* Screen coordinates a point or line may start from: finite and below 2^24 (a vertex inside the frustum
 * can still sit at w = 0 and divide to infinity; the reference's int conversion of that is undefined). */
__device__ __forceinline__ bool plSane(float4 p) { return fabsf(p.x) < 16777216.f && fabsf(p.y) < 16777216.f; }

__device__ void plUnitSetupRaw(const DrawParams &P, int u, PlUnit &U);

/* Assembly, clipping and culling of unit u; units at insane coordinates are dropped. */
__device__ void plUnitSetup(const DrawParams &P, int u, PlUnit &U) {
  plUnitSetupRaw(P, u, U);
  if ((U.kind == kPlPoint && !plSane(U.a)) || (U.kind == kPlLine && !(plSane(U.a) && plSane(U.b))))
    U.kind = kPlNone;
}

/* Assembly, clipping and culling of unit u (RendererSoft.cpp:192-299, :575-626). */
__device__ void plUnitSetupRaw(const DrawParams &P, int u, PlUnit &U) {
  U.kind = kPlNone;
  U.vertex = -1;
  const int type = P.st.primitiveType;
  if (type == 0) { /* processPointAssembly + clippingPoint */
    const int v = P.idx[u];
    if (P.mask[v] != 0) return;
    U.kind = kPlPoint;
    U.vertex = v;
    U.a = P.frag[v];
    U.ca = plVertexColour(P, v);
    return;
  }
  if (type == 1) { /* processLineAssembly + clippingLine before the perspective divide */
    if (plClipLine(P, P.idx[2 * u], P.idx[2 * u + 1], U)) U.kind = kPlLine;
    return;
  }
  /* triangles in point or line mode: no triangle clipping; processFaceCulling on screen positions */
  const int t = u / 3, k = u - 3 * t;
  const int i[3] = {P.idx[3 * t], P.idx[3 * t + 1], P.idx[3 * t + 2]};
  const float4 a = P.frag[i[0]], b = P.frag[i[1]], c = P.frag[i[2]];
  const float e0x = b.x - a.x, e0y = b.y - a.y, e1x = c.x - a.x, e1y = c.y - a.y;
  const bool front = fmaf(e0x, e1y, -(e0y * e1x)) > 0;
  if (P.st.cullFace && !front) return;
  if (P.st.polygonMode == 0) { /* rasterizationPolygonsPoint: clippingPoint on each index */
    const int v = i[k];
    if (P.mask[v] != 0) return;
    U.kind = kPlPoint;
    U.vertex = v;
    U.a = P.frag[v];
    U.ca = plVertexColour(P, v);
    return;
  }
  /* rasterizationPolygonsLine: edge (k, k+1), clipped after the perspective divide */
  if (plClipLine(P, i[k], i[(k + 1) % 3], U)) U.kind = kPlLine;
}

/* Fragments of a line: every step is a point of the line width. */
__device__ long long plLineCount(const PlUnit &U, float width) {
  PlWalk w = plWalkStart(U);
  long long n = 0;
  for (; w.x <= w.x1; plWalkStep(w)) {
    const float fx = w.steep ? (float)w.y : (float)w.x, fy = w.steep ? (float)w.x : (float)w.y;
    int x0, x1, y0, y1;
    plBox(fx, fy, width, x0, x1, y0, y1);
    n += plBoxCount(x0, x1, y0, y1);
  }
  return n;
}

static __global__ __launch_bounds__(256) void k_pl_units(DrawParams P, PlUnit *units, long long nunits,
                                                        unsigned long long *count, uint32_t *pkey, uint32_t *pval,
                                                        int has_colour) {
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nunits) return;
  PlUnit U;
  plUnitSetup(P, (int)u, U);
  units[u] = U;
  long long n = 0;
  if (has_colour && U.kind == kPlLine) n = plLineCount(U, P.st.lineWidth);
  count[u] = (unsigned long long)n;
  /* points sort by vertex, in unit order; everything else goes last */
  pkey[u] = U.kind == kPlPoint ? (uint32_t)U.vertex : 0xffffffffu;
  pval[u] = (uint32_t)u;
}

/* After the stable sort of point units by vertex: unit sorted[j] is the r-th rasterization of its vertex,
 * r = j - (first j of that vertex). Each rasterization moves the vertex's screen position to the last pixel
 * it wrote (RendererSoft.cpp:637-661 writes x, y into v->fragPos), so unit r starts r moves on. */
static __global__ __launch_bounds__(256) void k_pl_points(DrawParams P, PlUnit *units, const uint32_t *skey,
                                                         const uint32_t *sval, long long nunits,
                                                         unsigned long long *count, int has_colour) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nunits) return;
  const uint32_t key = skey[j];
  if (key == 0xffffffffu) return;
  long long first = j;
  while (first > 0 && skey[first - 1] == key) first--;
  const long long r = j - first;
  PlUnit &U = units[sval[j]];
  float px = U.a.x, py = U.a.y;
  const float size = P.pl_point_size;
  int x0, x1, y0, y1;
  if (has_colour)
    for (long long k = 0; k < r; k++) {
      plBox(px, py, size, x0, x1, y0, y1);
      if (x0 < x1 && y0 < y1) {
        px = (float)(x1 - 1);
        py = (float)(y1 - 1);
      }
    }
  U.a.x = px;
  U.a.y = py;
  plBox(px, py, size, x0, x1, y0, y1);
  count[sval[j]] = has_colour ? (unsigned long long)plBoxCount(x0, x1, y0, y1) : 0ull;
}

/* Fragment record of draw-order index f: pixel key (npix: outside the framebuffer), depth and colour. */
__device__ __forceinline__ void plPut(const DrawParams &P, long long f, int x, int y, float z, float4 c,
                                      uint32_t *key, uint32_t *val, float *zf, float4 *cf) {
  const bool in = (unsigned)x < (unsigned)P.fb_w && (unsigned)y < (unsigned)P.fb_h;
  key[f] = in ? (uint32_t)y * (uint32_t)P.fb_w + (uint32_t)x : (uint32_t)P.fb_w * (uint32_t)P.fb_h;
  val[f] = (uint32_t)f;
  zf[f] = z;
  cf[f] = c;
}

static __global__ __launch_bounds__(256) void k_pl_emit(DrawParams P, const PlUnit *units, long long nunits,
                                                       const unsigned long long *offset, uint32_t *key,
                                                       uint32_t *val, float *zf, float4 *cf) {
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nunits) return;
  const PlUnit U = units[u];
  long long f = (long long)offset[u];
  if (U.kind == kPlPoint) {
    int x0, x1, y0, y1;
    plBox(U.a.x, U.a.y, P.pl_point_size, x0, x1, y0, y1);
    for (int x = x0; x < x1; x++)
      for (int y = y0; y < y1; y++) plPut(P, f++, x, y, U.a.z, U.ca, key, val, zf, cf);
  } else if (U.kind == kPlLine) {
    PlWalk w = plWalkStart(U);
    const float z0 = w.swapped ? U.b.z : U.a.z, z1 = w.swapped ? U.a.z : U.b.z;
    const float4 c0 = w.swapped ? U.cb : U.ca, c1 = w.swapped ? U.ca : U.cb;
    const bool vc = P.shading != RHS_SHADING_BASIC;
    for (; w.x <= w.x1; plWalkStep(w)) {
      /* t = (x - x0) / dx: NaN for a line within one column, as in the reference */
      const float t = (float)(w.x - w.x0) / (float)w.dx, omt = 1.0f - t;
      const float z = fmaf(z0, omt, z1 * t); /* glm::mix as compiled there: fma(a, 1 - t, b t) */
      const float4 c = vc ? make_float4(fmaf(c0.x, omt, c1.x * t), fmaf(c0.y, omt, c1.y * t),
                                        fmaf(c0.z, omt, c1.z * t), fmaf(c0.w, omt, c1.w * t))
                          : c0;
      const float fx = w.steep ? (float)w.y : (float)w.x, fy = w.steep ? (float)w.x : (float)w.y;
      int x0, x1, y0, y1;
      plBox(fx, fy, P.st.lineWidth, x0, x1, y0, y1);
      for (int x = x0; x < x1; x++)
        for (int y = y0; y < y1; y++) plPut(P, f++, x, y, z, c, key, val, zf, cf);
    }
  }
}

/* Per pixel, its fragments in draw order through processPerSampleOperations (RendererSoft.cpp:358-407)
 * on every sample. */
template <int S>
static __global__ __launch_bounds__(256) void k_pl_apply(DrawParams P, const uint32_t *key, const uint32_t *val,
                                                        const float *zf, const float4 *cf, long long nfrag) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfrag) return;
  const uint32_t px = key[i];
  const uint32_t npix = (uint32_t)P.fb_w * (uint32_t)P.fb_h;
  if (px >= npix || (i > 0 && key[i - 1] == px)) return;
  const bool dtest = P.st.depthTest && P.depth != nullptr;
  const int dfunc = P.st.depthFunc;
  float dep[S];
  uint32_t col[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    dep[s] = dtest ? P.depth[(size_t)px * S + s] : 0.f;
    col[s] = P.color[(size_t)px * S + s];
  }
  for (long long j = i; j < nfrag && key[j] == px; j++) {
    const uint32_t f = val[j];
    const float z = gclamp(zf[f], 0.f, 1.f);
    const float4 c = cf[f];
#pragma unroll
    for (int s = 0; s < S; s++) {
      if (dtest) {
        if (!depthPass(z, dep[s], dfunc)) continue;
        if (P.st.depthMask) dep[s] = z;
      }
      RGBAf o = {gclamp(c.x, 0.f, 1.f), gclamp(c.y, 0.f, 1.f), gclamp(c.z, 0.f, 1.f), gclamp(c.w, 0.f, 1.f)};
      if (P.st.blend) o = blendColor(o, unpack01(col[s]), P.st);
      col[s] = pack4(u8t(o.r * 255.f), u8t(o.g * 255.f), u8t(o.b * 255.f), u8t(o.a * 255.f));
    }
  }
#pragma unroll
  for (int s = 0; s < S; s++) {
    if (dtest) P.depth[(size_t)px * S + s] = dep[s];
    P.color[(size_t)px * S + s] = col[s];
  }
}

}  // namespace rhip
