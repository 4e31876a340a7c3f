/*
 * rhip_kernels.h -- the gfx950 kernels of one RendererHIP draw.
 *
 * Pipeline (one draw; RendererSoft::draw, RendererSoft.cpp:136-164):
 *   k_vertex      vertex shader, clip mask, perspective divide, viewport   (:170-190, :971-992)
 *   k_setup       assembly, clip test, face culling, bounding box, tile count (:192-204, :277-299, :1005-1019)
 *   k_clip        Sutherland-Hodgman against the six planes, re-shading new vertices (:489-559, :956-970)
 *   k_clip_fans   the fan triangles a clipped polygon appends after all originals (:550-558, :227)
 *   k_emit        (tile, primitive) pairs; sorted stably by tile so every tile keeps API order
 *   k_tile_ranges start of each tile's run in the sorted pairs
 *   k_raster      one workgroup per 16x16 tile, one lane per pixel: quads anchored per
 *                 triangle bounding box, coverage, z/w, early-Z, interpolation, fragment
 *                 shading, depth test, blending, writes; then the MSAA resolve (:720-912)
 */
#pragma once
#include "rhip_shading.h"

namespace rhip {

/* --------------------------------------------------------------------------- clears */

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

/* --------------------------------------------------------------------------- vertex stage */

__global__ __launch_bounds__(256) void k_vertex(DrawParams P) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nverts) return;
  float a[16];
  const float *src = P.verts + (size_t)i * P.stride_f;
  for (int k = 0; k < 16; k++) a[k] = k < P.stride_f ? src[k] : 0.f;
  float v[kVaryStride];
#pragma unroll
  for (int k = 0; k < kVaryStride; k++) v[k] = 0.f;
  float4 c = runVS(P, a, v);
  P.clip[i] = c;
  P.mask[i] = clipMask(c);
  P.frag[i] = toScreen(P, c);
  float4 *o = (float4 *)(P.vary + (size_t)i * kVaryStride);
#pragma unroll
  for (int k = 0; k < kVaryStride / 4; k++) o[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
}

/* --------------------------------------------------------------------------- triangle setup */

/* One axis of the reference's block loop (RendererSoft.cpp:727-763): blocks of
 * 32 pixels start at (int)(min + 32*b); quads start at block start + 1 and step
 * by 2 while x < block start + 32 and x <= max. Returns the first and last quad
 * origin, or first > last when the triangle has no quads on this axis. */
__device__ __forceinline__ void axisRange(float mn, float mx, int bc, int *first, int *last) {
  *first = 1;
  *last = 0;
  if (bc <= 0) return;
  int fl = (int)floorf(mx);
  for (int b = 0; b < bc; b++) {
    int s = (int)(mn + (float)(32 * b));
    if (s + 1 <= min(fl, s + 31)) {
      *first = s + 1;
      break;
    }
  }
  for (int b = bc - 1; b >= 0; b--) {
    int s = (int)(mn + (float)(32 * b));
    int lim = min(fl, s + 31);
    if (s + 1 <= lim) {
      *last = s + 1 + 2 * ((lim - s - 1) >> 1);
      break;
    }
  }
}

/* Fills the setup record of primitive `id` from three screen-space vertices and
 * returns the number of screen tiles its quads touch (0 if culled or empty). */
__device__ int setupPrim(const DrawParams &P, int id, int i0, int i1, int i2) {
  PrimSetup s;
  s.v[0] = i0;
  s.v[1] = i1;
  s.v[2] = i2;
  float4 a = P.frag[i0], b = P.frag[i1], c = P.frag[i2];
  /* processFaceCulling: z of cross(v1 - v0, v2 - v0) as compiled (vfmsub) */
  float nz = fmaf(b.x - a.x, c.y - a.y, -((b.y - a.y) * (c.x - a.x)));
  bool front = nz > 0;
  /* the edge-function denominator of barycentric() depends only on the triangle
   * (RendererSoft.cpp:1025-1036): a degenerate one never covers a sample */
  float a0 = c.x - a.x, a1 = b.x - a.x, b0 = c.y - a.y, b1 = b.y - a.y;
  float uz = fmaf(a0, b1, -(a1 * b0));
  bool live = !(P.st.cullFace && !front) && !(fabsf(uz) < 1.19209290e-07f);
  /* triangleBoundingBox (RendererSoft.cpp:1005-1019) then min -= 1 (:725) */
  float mnx = fminf(fminf(a.x, b.x), c.x), mny = fminf(fminf(a.y, b.y), c.y);
  float mxx = fmaxf(fmaxf(a.x, b.x), c.x), mxy = fmaxf(fmaxf(a.y, b.y), c.y);
  mnx = fmaxf(mnx - 0.5f, 0.f);
  mny = fmaxf(mny - 0.5f, 0.f);
  mxx = fminf(mxx + 0.5f, P.vp_w - 1.f);
  mxy = fminf(mxy + 0.5f, P.vp_h - 1.f);
  mnx -= 1.f;
  mny -= 1.f;
  s.min_x = mnx;
  s.min_y = mny;
  s.max_x = mxx;
  s.max_y = mxy;
  s.bcx = (int)((mxx - mnx + 32 - 1.f) / 32);
  s.bcy = (int)((mxy - mny + 32 - 1.f) / 32);
  int fx, lx, fy, ly;
  axisRange(mnx, mxx, s.bcx, &fx, &lx);
  axisRange(mny, mxy, s.bcy, &fy, &ly);
  s.qx0 = fx;
  s.qx1 = lx + 1;
  s.qy0 = fy;
  s.qy1 = ly + 1;
  int count = 0;
  if (live && fx <= lx && fy <= ly) {
    int x0 = max(fx, 0), x1 = min(lx + 1, P.fb_w - 1);
    int y0 = max(fy, 0), y1 = min(ly + 1, P.fb_h - 1);
    if (x0 <= x1 && y0 <= y1) count = (x1 / kTile - x0 / kTile + 1) * (y1 / kTile - y0 / kTile + 1);
  }
  s.flags = (count > 0 ? 1 : 0) | (front ? 2 : 0);
  P.prims[id] = s;
  P.tile_count[id] = count;
  return count;
}

__global__ __launch_bounds__(256) void k_setup(DrawParams P, int *clip_flag, int *clip_total) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.ntris) return;
  int i0 = P.idx[3 * t], i1 = P.idx[3 * t + 1], i2 = P.idx[3 * t + 2];
  int m = P.mask[i0] | P.mask[i1] | P.mask[i2];
  if (m) {
    clip_flag[t] = 1;
    atomicAdd(clip_total, 1);
    P.prims[t].flags = 0;
    P.tile_count[t] = 0;
    return;
  }
  clip_flag[t] = 0;
  setupPrim(P, t, i0, i1, i2);
}

/* --------------------------------------------------------------------------- clipping */

/* dot(FrustumClipPlane[i], clipPos) as dpps computes it: (p0*c0 + p1*c1) + (p2*c2 + p3*c3) */
__device__ __forceinline__ float planeDist(int i, float4 c) {
  const float px = i == 0 ? -1.f : (i == 1 ? 1.f : 0.f);
  const float py = i == 2 ? -1.f : (i == 3 ? 1.f : 0.f);
  const float pz = i == 4 ? -1.f : (i == 5 ? 1.f : 0.f);
  return (px * c.x + py * c.y) + (pz * c.z + 1.f * c.w);
}

__device__ __forceinline__ const float *rawOf(const DrawParams &P, int v) {
  return v < P.nverts ? P.verts + (size_t)v * P.stride_f : P.clip_raw + (size_t)(v - P.nverts) * P.stride_f;
}

/* clippingTriangle (RendererSoft.cpp:489-559) for the k-th triangle that needs it. */
__global__ void k_clip(DrawParams P) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nclip) return;
  int t = P.clip_list[k];
  int i0 = P.idx[3 * t], i1 = P.idx[3 * t + 1], i2 = P.idx[3 * t + 2];
  int mask = P.mask[i0] | P.mask[i1] | P.mask[i2];
  int in[kMaxPoly + 1], out[kMaxPoly + 1];
  int nin = 3, nnew = 0;
  in[0] = i0;
  in[1] = i1;
  in[2] = i2;
  bool full = false;
  for (int pl = 0; pl < 6; pl++) {
    if (!(mask & (1 << pl))) continue;
    if (nin < 3) {
      full = true;
      break;
    }
    int nout = 0;
    int ip = in[0];
    float dp = planeDist(pl, P.clip[ip]);
    in[nin] = ip;
    for (int q = 1; q <= nin; q++) {
      int ic = in[q];
      float d = planeDist(pl, P.clip[ic]);
      if (dp >= 0) {
        if (nout >= kMaxPoly) {
          full = true;
          break;
        }
        out[nout++] = ip;
      }
      if (signbit(dp) != signbit(d)) {
        if (nout >= kMaxPoly || nnew >= kMaxClipVerts) {
          full = true;
          break;
        }
        float tt = d < 0 ? dp / (dp - d) : -dp / (d - dp);
        /* clippingNewVertex + interpolateVertex (:956-970, :1058-1070): lerp the raw
         * attributes as fma(a, 1-t, b*t), then run the vertex shader on the result */
        int nv = P.nverts + k * kMaxClipVerts + nnew;
        nnew++;
        const float *ra = rawOf(P, ip), *rb = rawOf(P, ic);
        float *rn = (float *)P.clip_raw + (size_t)(nv - P.nverts) * P.stride_f;
        float omt = 1.0f - tt;
        float a[16];
        for (int j = 0; j < 16; j++) {
          float val = j < P.stride_f ? fmaf(ra[j], omt, rb[j] * tt) : 0.f;
          if (j < P.stride_f) rn[j] = val;
          a[j] = val;
        }
        float v[kVaryStride];
        for (int j = 0; j < kVaryStride; j++) v[j] = 0.f;
        float4 c = runVS(P, a, v);
        P.clip[nv] = c;
        P.mask[nv] = clipMask(c);
        P.frag[nv] = toScreen(P, c);
        float *vo = P.vary + (size_t)nv * kVaryStride;
        for (int j = 0; j < kVaryStride; j++) vo[j] = v[j];
        out[nout++] = nv;
      }
      ip = ic;
      dp = d;
    }
    if (full) {
      /* more vertices than a convex polygon can have: rounding produced extra
       * plane crossings; the triangle is dropped and the overflow counted */
      atomicAdd(&P.counters[2], 1ull);
      break;
    }
    for (int q = 0; q < nout; q++) in[q] = out[q];
    nin = nout;
  }
  int *poly = P.clip_poly + k * (kMaxPoly + 1);
  if (full || nin < 3) { /* fewer than 3 left: the reference reads past its index list here */
    poly[kMaxPoly] = 0;
    P.prims[t].flags = 0;
    P.tile_count[t] = 0;
    return;
  }
  for (int q = 0; q < nin; q++) poly[q] = in[q];
  poly[kMaxPoly] = nin;
  setupPrim(P, t, in[0], in[1], in[2]);
}

/* fans appended after all original primitives, in clipped-triangle order (:550-558) */
__global__ void k_fan_count(DrawParams P, int *fan_count) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nclip) return;
  int n = P.clip_poly[k * (kMaxPoly + 1) + kMaxPoly];
  fan_count[k] = n > 3 ? n - 3 : 0;
}

__global__ void k_clip_fans(DrawParams P) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nclip) return;
  const int *poly = P.clip_poly + k * (kMaxPoly + 1);
  int n = poly[kMaxPoly];
  for (int i = 3; i < n; i++) setupPrim(P, P.ntris + P.fan_offset[k] + (i - 3), poly[0], poly[i - 1], poly[i]);
}

/* --------------------------------------------------------------------------- binning */

__global__ __launch_bounds__(256) void k_emit(DrawParams P) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P.nprims) return;
  int n = P.tile_count[id];
  if (n == 0) return;
  const PrimSetup &s = P.prims[id];
  int x0 = max(s.qx0, 0) / kTile, x1 = min(s.qx1, P.fb_w - 1) / kTile;
  int y0 = max(s.qy0, 0) / kTile, y1 = min(s.qy1, P.fb_h - 1) / kTile;
  int o = P.tile_offset[id];
  for (int ty = y0; ty <= y1; ty++)
    for (int tx = x0; tx <= x1; tx++) {
      P.pair_key[o] = (uint32_t)(ty * P.tiles_x + tx);
      P.pair_val[o] = id;
      o++;
    }
}

__global__ void k_tile_ranges(DrawParams P) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ntiles = P.tiles_x * P.tiles_y;
  if (t > ntiles) return;
  /* lower bound of tile t in the sorted keys */
  int lo = 0, hi = P.npairs;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (P.pair_key_sorted[mid] < (uint32_t)t) lo = mid + 1;
    else hi = mid;
  }
  P.tile_start[t] = lo;
}

/* --------------------------------------------------------------------------- fine raster */

/* barycentric (RendererSoft.cpp:1021-1056) as its SSE path computes it:
 * u = vmulps/vfmsub cross product, divided by u.z, bc = (1-(ux+uy), uy, ux).
 * bc is produced even when the sample is outside. */
struct Tri {
  float v0x, v0y, a0, a1, b0, b1, uz;
  float z[3], w[3];
};

__device__ __forceinline__ bool bary(const Tri &T, float px, float py, float4 &bc) {
  float a2 = T.v0x - px, b2 = T.v0y - py;
  float ux = fmaf(T.a1, b2, -(a2 * T.b1));
  float uy = fmaf(a2, T.b0, -(T.a0 * b2));
  float uz = T.uz;
  ux = ux / uz;
  uy = uy / uz;
  bc = make_float4(1.f - (ux + uy), uy, ux, 0.f);
  return !(bc.x < 0 || bc.y < 0 || bc.z < 0);
}

/* dpps with imm 0xff over (b0, b1, b2, 0) . (v0, v1, v2, 0) */
__device__ __forceinline__ float dpps3(float4 b, const float *v) {
  return (b.x * v[0] + b.y * v[1]) + (b.z * v[2] + b.w * 0.f);
}

/* Quad origins that contain pixel coordinate p on one axis, in the order the
 * reference visits them (increasing block). Returns 0..2. */
__device__ __forceinline__ int axisQuads(float mn, float mx, int bc, int p, int *q) {
  int n = 0;
  int fl = (int)floorf(mx);
  int s0 = (int)mn;
#pragma unroll
  for (int c = 0; c < 2; c++) {
    int x = p - 1 + c;
    if (x > fl) continue;
    int sb = 0;
    bool found = false;
    if (x >= s0 + 1 && x <= s0 + 31) {
      sb = s0;
      found = bc > 0;
    } else if (x > s0 + 31) {
      int be = (x - s0 - 1) >> 5;
      for (int b = max(be - 1, 1); b <= be + 1 && b < bc; b++) {
        int s = (int)(mn + (float)(32 * b));
        if (x >= s + 1 && x <= s + 31) {
          sb = s;
          found = true;
          break;
        }
      }
    }
    if (found && ((x - sb - 1) & 1) == 0) q[n++] = x;
  }
  return n;
}

struct FragOut {
  bool any;
};

/* Shading-sample barycentrics of one pixel of a quad, as the quad context
 * leaves them for interpolation (RendererSoft.cpp:776-820): the centre sample
 * if it is inside, else the first inside sample, perspective-corrected; a pixel
 * with no inside sample keeps the centre's screen-space barycentrics. */
template <int S>
__device__ __forceinline__ float4 shadingBc(const Tri &T, float fx, float fy) {
  const float lx[4] = {0.375f, 0.875f, 0.125f, 0.625f};
  const float ly[4] = {0.875f, 0.625f, 0.375f, 0.125f};
  float4 bc;
  bool in = bary(T, fx + 0.5f, fy + 0.5f, bc);
  if (S == 4 && !in) {
    float4 bcs;
    float4 centre = bc;
    bool found = false;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      if (!found && bary(T, lx[s] + fx, ly[s] + fy, bcs)) {
        bc = bcs;
        found = true;
      }
    }
    in = found;
    if (!found) bc = centre;
  }
  if (in) {
    float w = dpps3(bc, T.w);
    float iw = 1.f / w;
    bc = make_float4(bc.x * (iw * T.w[0]), bc.y * (iw * T.w[1]), bc.z * (iw * T.w[2]), bc.w * (iw * 1.f));
  }
  return bc;
}

__device__ __forceinline__ float interp(float4 b, float v0, float v1, float v2) {
  return fmaf(v2, b.z, fmaf(v1, b.y, v0 * b.x));
}

template <int SHADING>
__device__ __forceinline__ RGBAf runFS(const DrawParams &P, const Frag &f) {
  if (SHADING == RHS_SHADING_BASIC) {
    return {P.u.base_color[0], P.u.base_color[1], P.u.base_color[2], P.u.base_color[3]};
  } else if (SHADING == RHS_SHADING_BLINN_PHONG) {
    return fsBlinnPhong(P, f, false);
  } else if (SHADING == RHS_SHADING_BLINN_PHONG_DIR) {
    return fsBlinnPhong(P, f, true);
  } else if (SHADING == RHS_SHADING_PBR) {
    return fsPbr(P, f);
  } else {
    return {f.v[0], f.v[1], f.v[2], f.v[3]};
  }
}

/* Varyings a fragment shader reads (the others are never interpolated). */
template <int SHADING>
__device__ __forceinline__ bool varyUsed(const DrawParams &P, int j) {
  if (SHADING == RHS_SHADING_VERTEX_COLOR) return j < 4;
  if (SHADING == RHS_SHADING_BASIC) return false;
  bool nm = (P.defines & RHS_DEF_NORMAL_MAP) != 0;
  if (j < 2) return true;
  if (j >= 4 && j <= 6) return !nm;
  if (j >= 12 && j <= 14) return true;
  if (j >= 16 && j <= 18) return true;
  if (SHADING == RHS_SHADING_PBR) return nm && ((j >= 20 && j <= 22) || (j >= 24 && j <= 26));
  if (j >= 20 && j <= 23) return P.u.enable_shadow != 0;
  return nm && ((j >= 24 && j <= 26) || (j >= 28 && j <= 30));
}

template <int S, int SHADING>
__global__ __launch_bounds__(256) void k_raster(DrawParams P, int need_deriv) {
  const int tile = blockIdx.x;
  const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wx0 = tx * kTile + (wave & 1) * kSub, wy0 = ty * kTile + (wave >> 1) * kSub;
  const int px = wx0 + (lane & 7), py = wy0 + (lane >> 3);
  const bool inFb = px < P.fb_w && py < P.fb_h;
  const size_t pix = (size_t)py * P.fb_w + px;
  const bool hasC = P.color != nullptr, hasD = P.depth != nullptr;

  float dep[S];
  uint32_t col[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    dep[s] = 0.f;
    col[s] = 0u;
  }
  if (inFb) {
    if (hasD)
      for (int s = 0; s < S; s++) dep[s] = P.depth[pix * S + s];
    if (hasC)
      for (int s = 0; s < S; s++) col[s] = P.color[pix * S + s];
  }

  const int start = P.tile_start[tile], end = P.tile_start[tile + 1];
  const float lx[4] = {0.375f, 0.875f, 0.125f, 0.625f};
  const float ly[4] = {0.875f, 0.625f, 0.375f, 0.125f};
  unsigned long long nfs = 0, nwr = 0;
  const bool dtest = P.st.depthTest && hasD;

  for (int it = start; it < end; it++) {
    const int id = P.pair_val_sorted[it];
    const PrimSetup &ps = P.prims[id];
    if (ps.qx1 < wx0 || ps.qx0 > wx0 + kSub - 1 || ps.qy1 < wy0 || ps.qy0 > wy0 + kSub - 1) continue;
    if (!inFb) continue;
    int qxs[2], qys[2];
    int nqx = axisQuads(ps.min_x, ps.max_x, ps.bcx, px, qxs);
    if (nqx == 0) continue;
    int nqy = axisQuads(ps.min_y, ps.max_y, ps.bcy, py, qys);
    if (nqy == 0) continue;

    const float4 f0 = P.frag[ps.v[0]], f1 = P.frag[ps.v[1]], f2 = P.frag[ps.v[2]];
    Tri T;
    T.v0x = f0.x;
    T.v0y = f0.y;
    T.a0 = f2.x - f0.x;
    T.a1 = f1.x - f0.x;
    T.b0 = f2.y - f0.y;
    T.b1 = f1.y - f0.y;
    T.uz = fmaf(T.a0, T.b1, -(T.a1 * T.b0));
    T.z[0] = f0.z, T.z[1] = f1.z, T.z[2] = f2.z;
    T.w[0] = f0.w, T.w[1] = f1.w, T.w[2] = f2.w;

    for (int iy = 0; iy < nqy; iy++)
      for (int ix = 0; ix < nqx; ix++) {
        const int qx = qxs[ix], qy = qys[iy];
        const float fx = (float)px, fy = (float)py;
        /* own pixel: samples 0..3 (+ centre) for 4x, the centre for 1x */
        bool sin[5];
        float4 sbc[5];
        float sz[5];
        int shade;
        bool pin;
        if (S == 4) {
#pragma unroll
          for (int s = 0; s < 4; s++) sin[s] = bary(T, lx[s] + fx, ly[s] + fy, sbc[s]);
          sin[4] = bary(T, fx + 0.5f, fy + 0.5f, sbc[4]);
          pin = sin[0] || sin[1] || sin[2] || sin[3];
          shade = 4;
          if (!sin[4]) {
#pragma unroll
            for (int s = 3; s >= 0; s--)
              if (sin[s]) shade = s;
          }
        } else {
          sin[0] = bary(T, fx + 0.5f, fy + 0.5f, sbc[0]);
          pin = sin[0];
          shade = 0;
        }
        if (!pin) continue;
        const int NS = S == 4 ? 5 : 1;
#pragma unroll
        for (int s = 0; s < NS; s++) {
          sz[s] = 0.f;
          if (!sin[s]) continue;
          float z = dpps3(sbc[s], T.z);
          float w = dpps3(sbc[s], T.w);
          sz[s] = z;
          if (z < 0.f || z > 1.f) sin[s] = false;
          float iw = 1.f / w;
          sbc[s] = make_float4(sbc[s].x * (iw * T.w[0]), sbc[s].y * (iw * T.w[1]), sbc[s].z * (iw * T.w[2]),
                               sbc[s].w * (iw * 1.f));
        }
        /* early-Z (RendererSoft.cpp:853-878): test without writing */
        if (P.st.depthTest) {
          if (S == 4) {
            bool any = false;
#pragma unroll
            for (int s = 0; s < 4; s++) {
              if (!sin[s]) continue;
              if (hasD) sin[s] = depthPass(gclamp(sz[s], 0.f, 1.f), dep[s], P.st.depthFunc);
              any = any || sin[s];
            }
            pin = any;
          } else {
            bool ok = hasD ? depthPass(gclamp(sz[0], 0.f, 1.f), dep[0], P.st.depthFunc) : true;
            sin[0] = ok;
            pin = ok;
          }
          if (!pin) continue;
        }
        RGBAf c = {0.f, 0.f, 0.f, 0.f};
        if (hasC) {
          /* interpolate the varyings the shader reads at the shading sample */
          float vv[kVaryStride];
          const float *va = P.vary + (size_t)ps.v[0] * kVaryStride;
          const float *vb = P.vary + (size_t)ps.v[1] * kVaryStride;
          const float *vc = P.vary + (size_t)ps.v[2] * kVaryStride;
          const float4 b = sbc[shade];
#pragma unroll
          for (int j = 0; j < kVaryStride; j++) vv[j] = varyUsed<SHADING>(P, j) ? interp(b, va[j], vb[j], vc[j]) : 0.f;
          Frag fr;
          fr.v = vv;
          if (need_deriv) {
            const int k = (px - qx) + 2 * (py - qy);
#pragma unroll
            for (int j = 0; j < 3; j++) {
              const int jx = qx + (j & 1), jy = qy + (j >> 1);
              if (j == k) {
                fr.tc[j] = {vv[0], vv[1]};
              } else {
                float4 pb = shadingBc<S>(T, (float)jx, (float)jy);
                fr.tc[j] = {interp(pb, va[0], vb[0], vc[0]), interp(pb, va[1], vb[1], vc[1])};
              }
            }
          } else {
            fr.tc[0] = fr.tc[1] = fr.tc[2] = {0.f, 0.f};
          }
          c = runFS<SHADING>(P, fr);
          nfs++;
        }
        /* per-sample operations (RendererSoft.cpp:358-407) */
#pragma unroll
        for (int s = 0; s < S; s++) {
          const int si = S == 4 ? s : shade;
          if (S == 4 && !sin[s]) continue;
          if (dtest) {
            float z = gclamp(sz[si], 0.f, 1.f);
            if (!depthPass(z, dep[s], P.st.depthFunc)) continue;
            if (P.st.depthMask) dep[s] = z;
          }
          if (!hasC) continue;
          RGBAf o = {gclamp(c.r, 0.f, 1.f), gclamp(c.g, 0.f, 1.f), gclamp(c.b, 0.f, 1.f), gclamp(c.a, 0.f, 1.f)};
          if (P.st.blend) o = blendColor(o, unpack01(col[s]), P.st);
          col[s] = pack4(u8t(o.r * 255.f), u8t(o.g * 255.f), u8t(o.b * 255.f), u8t(o.a * 255.f));
          nwr++;
        }
      }
  }

  if (inFb) {
    if (hasD && start < end)
      for (int s = 0; s < S; s++) P.depth[pix * S + s] = dep[s];
    if (hasC) {
      if (start < end)
        for (int s = 0; s < S; s++) P.color[pix * S + s] = col[s];
      if (S == 4) {
        /* multiSampleResolve (RendererSoft.cpp:880-912): u8(sum / 4) */
        int r[4];
#pragma unroll
        for (int ch = 0; ch < 4; ch++) r[ch] = (chan(col[0], ch) + chan(col[1], ch) + chan(col[2], ch) + chan(col[3], ch)) >> 2;
        P.color_resolved[pix] = pack4(r[0], r[1], r[2], r[3]);
      }
    }
  }
  if (nfs) atomicAdd(&P.counters[0], nfs);
  if (nwr) atomicAdd(&P.counters[1], nwr);
}

/* --------------------------------------------------------------------------- two-phase raster
 *
 * For draws without blending whose depth test writes depth with LESS, LEQUAL,
 * GREATER or GEQUAL, the colour a sample ends with is the fragment shader
 * output of the last fragment that passed its depth test, and which fragment
 * that is follows from depths and API order alone: the minimum (or maximum)
 * depth, ties going to the earliest fragment for the strict tests and the
 * latest for the others. Phase 1 finds that winner per sample with 64-bit LDS
 * atomics on key = (monotonic depth bits << 32) | order word, one lane per
 * triangle of the tile; phase 2 shades each pixel's winners once, one lane per
 * pixel. The fragment shader's inputs depend only on the triangle, the pixel
 * and the 2x2 quad it is shaded in, so the result equals the sequential one. */
enum VisMode { VIS_LESS = 0, VIS_LEQUAL = 1, VIS_GREATER = 2, VIS_GEQUAL = 3 };

template <int MODE>
__device__ __forceinline__ constexpr bool visMax() { return MODE == VIS_GREATER || MODE == VIS_GEQUAL; }
/* strict tests keep the first of equal depths; LEQUAL / GEQUAL the last */
template <int MODE>
__device__ __forceinline__ constexpr bool visFirst() { return MODE == VIS_LESS || MODE == VIS_GREATER; }
/* the low word grows with the order for LESS and GEQUAL and shrinks for the others */
template <int MODE>
__device__ __forceinline__ uint32_t visLow(uint32_t order) {
  return (MODE == VIS_LESS || MODE == VIS_GEQUAL) ? order : ~order;
}

/* Unsigned key that orders like the float (both zeros equal). */
__device__ __forceinline__ uint32_t depthKey(float z) {
  uint32_t b = __float_as_uint(z);
  if ((b << 1) == 0) b = 0;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

template <int MODE>
__device__ __forceinline__ unsigned long long visInit(float dep, bool open) {
  /* a NaN in the buffer fails every comparison: no fragment may win there */
  if (!open || dep != dep) return ((unsigned long long)(visMax<MODE>() ? 0xffffffffu : 0u) << 32) | visLow<MODE>(0);
  return ((unsigned long long)depthKey(dep) << 32) | visLow<MODE>(0);
}

__device__ __forceinline__ Tri makeTri(const float4 &f0, const float4 &f1, const float4 &f2) {
  Tri T;
  T.v0x = f0.x;
  T.v0y = f0.y;
  T.a0 = f2.x - f0.x;
  T.a1 = f1.x - f0.x;
  T.b0 = f2.y - f0.y;
  T.b1 = f1.y - f0.y;
  T.uz = fmaf(T.a0, T.b1, -(T.a1 * T.b0));
  T.z[0] = f0.z, T.z[1] = f1.z, T.z[2] = f2.z;
  T.w[0] = f0.w, T.w[1] = f1.w, T.w[2] = f2.w;
  return T;
}

template <int S, int SHADING, int MODE>
__global__ __launch_bounds__(256) void k_raster_vis(DrawParams P, int need_deriv) {
  constexpr int NPIX = kTile * kTile;
  __shared__ unsigned long long keys[NPIX * S];
  const int tile = blockIdx.x;
  const int tx0 = (tile % P.tiles_x) * kTile, ty0 = (tile / P.tiles_x) * kTile;
  const int lp = threadIdx.x;
  const int px = tx0 + (lp & (kTile - 1)), py = ty0 + lp / kTile;
  const bool inFb = px < P.fb_w && py < P.fb_h;
  const size_t pix = (size_t)py * P.fb_w + px;
  const bool hasC = P.color != nullptr;
  const int start = P.tile_start[tile], end = P.tile_start[tile + 1];
  const float lx[4] = {0.375f, 0.875f, 0.125f, 0.625f};
  const float ly[4] = {0.875f, 0.625f, 0.375f, 0.125f};

  float dep[S];
  uint32_t col[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    dep[s] = 0.f;
    col[s] = 0u;
  }
  if (inFb) {
    for (int s = 0; s < S; s++) dep[s] = P.depth[pix * S + s];
    if (hasC)
      for (int s = 0; s < S; s++) col[s] = P.color[pix * S + s];
  }
  if (start == end) {
    if (inFb && hasC && S == 4) {
      int r[4];
#pragma unroll
      for (int ch = 0; ch < 4; ch++) r[ch] = (chan(col[0], ch) + chan(col[1], ch) + chan(col[2], ch) + chan(col[3], ch)) >> 2;
      P.color_resolved[pix] = pack4(r[0], r[1], r[2], r[3]);
    }
    return;
  }
#pragma unroll
  for (int s = 0; s < S; s++) keys[lp * S + s] = visInit<MODE>(dep[s], inFb);
  __syncthreads();

  /* phase 1: visibility, one lane per triangle of the tile */
  for (int it = start + lp; it < end; it += blockDim.x) {
    const int id = P.pair_val_sorted[it];
    const PrimSetup ps = P.prims[id];
    const int x0 = max(ps.qx0, tx0), x1 = min(min(ps.qx1, tx0 + kTile - 1), P.fb_w - 1);
    const int y0 = max(ps.qy0, ty0), y1 = min(min(ps.qy1, ty0 + kTile - 1), P.fb_h - 1);
    if (x0 > x1 || y0 > y1) continue;
    const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
    for (int y = y0; y <= y1; y++) {
      int qys[2];
      const int nqy = axisQuads(ps.min_y, ps.max_y, ps.bcy, y, qys);
      if (!nqy) continue;
      for (int x = x0; x <= x1; x++) {
        int qxs[2];
        const int nqx = axisQuads(ps.min_x, ps.max_x, ps.bcx, x, qxs);
        if (!nqx) continue;
        /* of several quads of one triangle holding this pixel, the test keeps the first or the last */
        const uint32_t inst = visFirst<MODE>() ? 0u : (uint32_t)((nqy - 1) * 2 + (nqx - 1));
        const uint32_t low = visLow<MODE>((((uint32_t)id << 2) | inst) + 1u);
        unsigned long long *kp = &keys[((y - ty0) * kTile + (x - tx0)) * S];
        const float fx = (float)x, fy = (float)y;
        if (S == 4) {
#pragma unroll
          for (int s = 0; s < 4; s++) {
            float4 bc;
            if (!bary(T, lx[s] + fx, ly[s] + fy, bc)) continue;
            const float z = dpps3(bc, T.z);
            if (z < 0.f || z > 1.f) continue; /* multisampled depth clip (RendererSoft.cpp:796-799) */
            const unsigned long long key = ((unsigned long long)depthKey(z) << 32) | low;
            if (visMax<MODE>())
              atomicMax(kp + s, key);
            else
              atomicMin(kp + s, key);
          }
        } else {
          float4 bc;
          if (!bary(T, fx + 0.5f, fy + 0.5f, bc)) continue;
          const float z = gclamp(dpps3(bc, T.z), 0.f, 1.f);
          if (z != z) continue;
          const unsigned long long key = ((unsigned long long)depthKey(z) << 32) | low;
          if (visMax<MODE>())
            atomicMax(kp, key);
          else
            atomicMin(kp, key);
        }
      }
    }
  }
  __syncthreads();

  /* phase 2: shade each pixel's winners, one lane per pixel */
  if (!inFb) return;
  uint32_t win[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    const uint32_t low = (uint32_t)keys[lp * S + s];
    const uint32_t ord = (MODE == VIS_LESS || MODE == VIS_GEQUAL) ? low : ~low;
    win[s] = ord ? ord - 1u : 0xffffffffu;
  }
  unsigned long long nfs = 0, nwr = 0;
#pragma unroll
  for (int s = 0; s < S; s++) {
    const uint32_t w = win[s];
    if (w == 0xffffffffu) continue;
    bool seen = false;
#pragma unroll
    for (int t = 0; t < s; t++) seen = seen || win[t] == w;
    if (seen) continue;
    const int id = (int)(w >> 2), inst = (int)(w & 3u);
    const PrimSetup ps = P.prims[id];
    const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
    const float fx = (float)px, fy = (float)py;
    /* raw barycentrics of the samples and the shading sample (RendererSoft.cpp:776-802) */
    float4 sbc[5];
    bool sin[5];
    int shade;
    if (S == 4) {
#pragma unroll
      for (int t = 0; t < 4; t++) sin[t] = bary(T, lx[t] + fx, ly[t] + fy, sbc[t]);
      sin[4] = bary(T, fx + 0.5f, fy + 0.5f, sbc[4]);
      shade = 4;
      if (!sin[4]) {
#pragma unroll
        for (int t = 3; t >= 0; t--)
          if (sin[t]) shade = t;
      }
    } else {
      sin[0] = bary(T, fx + 0.5f, fy + 0.5f, sbc[0]);
      shade = 0;
    }
    RGBAf c = {0.f, 0.f, 0.f, 0.f};
    if (hasC) {
      float4 b = sbc[shade];
      const float w = dpps3(b, T.w);
      const float iw = 1.f / w;
      b = make_float4(b.x * (iw * T.w[0]), b.y * (iw * T.w[1]), b.z * (iw * T.w[2]), b.w * (iw * 1.f));
      float vv[kVaryStride];
      const float *va = P.vary + (size_t)ps.v[0] * kVaryStride;
      const float *vb = P.vary + (size_t)ps.v[1] * kVaryStride;
      const float *vc = P.vary + (size_t)ps.v[2] * kVaryStride;
#pragma unroll
      for (int j = 0; j < kVaryStride; j++) vv[j] = varyUsed<SHADING>(P, j) ? interp(b, va[j], vb[j], vc[j]) : 0.f;
      Frag fr;
      fr.v = vv;
      fr.tc[0] = fr.tc[1] = fr.tc[2] = {0.f, 0.f};
      if (need_deriv) {
        int qxs[2], qys[2];
        axisQuads(ps.min_x, ps.max_x, ps.bcx, px, qxs);
        axisQuads(ps.min_y, ps.max_y, ps.bcy, py, qys);
        const int qx = qxs[inst & 1], qy = qys[inst >> 1];
        const int k = (px - qx) + 2 * (py - qy);
#pragma unroll
        for (int j = 0; j < 3; j++) {
          if (j == k) {
            fr.tc[j] = {vv[0], vv[1]};
          } else {
            const float4 pb = shadingBc<S>(T, (float)(qx + (j & 1)), (float)(qy + (j >> 1)));
            fr.tc[j] = {interp(pb, va[0], vb[0], vc[0]), interp(pb, va[1], vb[1], vc[1])};
          }
        }
      }
      c = runFS<SHADING>(P, fr);
      c = {gclamp(c.r, 0.f, 1.f), gclamp(c.g, 0.f, 1.f), gclamp(c.b, 0.f, 1.f), gclamp(c.a, 0.f, 1.f)};
      nfs++;
    }
    const uint32_t packed = pack4(u8t(c.r * 255.f), u8t(c.g * 255.f), u8t(c.b * 255.f), u8t(c.a * 255.f));
#pragma unroll
    for (int t = 0; t < S; t++) {
      if (win[t] != w) continue;
      dep[t] = gclamp(dpps3(sbc[S == 4 ? t : 0], T.z), 0.f, 1.f);
      if (hasC) {
        col[t] = packed;
        nwr++;
      }
    }
  }
  for (int s = 0; s < S; s++) P.depth[pix * S + s] = dep[s];
  if (hasC) {
    for (int s = 0; s < S; s++) P.color[pix * S + s] = col[s];
    if (S == 4) {
      int r[4];
#pragma unroll
      for (int ch = 0; ch < 4; ch++) r[ch] = (chan(col[0], ch) + chan(col[1], ch) + chan(col[2], ch) + chan(col[3], ch)) >> 2;
      P.color_resolved[pix] = pack4(r[0], r[1], r[2], r[3]);
    }
  }
  if (nfs) atomicAdd(&P.counters[0], nfs);
  if (nwr) atomicAdd(&P.counters[1], nwr);
}

/* --------------------------------------------------------------------------- textures */

/* generateMipmaps / sampleBufferBilinear (SamplerSoft.h:89-110, 240-252): level i is a
 * clamp-to-edge bilinear resample of level i-1 at (x*r + r/2, y*r + r/2). */
template <typename T>
__global__ void k_mipgen(const TexDesc *d, int layer, int level) {
  const TexLevel &src = d->lv[layer][level - 1];
  const TexLevel &dst = d->lv[layer][level];
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= dst.w || y >= dst.h) return;
  float rx = (float)src.w / (float)dst.w, ry = (float)src.h / (float)dst.h;
  float dx = 0.5f * rx, dy = 0.5f * ry;
  T v = bilinear<T>(*d, src, fmaf((float)x, rx, dx), fmaf((float)y, ry, dy), 2);
  T *o = (T *)(d->base + dst.offset);
  o[texelIndex(dst, d->layout, x, y)] = v;
}

}  // namespace rhip
