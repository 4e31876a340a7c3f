/*
 * rhip_kernels.h -- the gfx950 kernels of one RendererHIP draw.
 *
 * Pipeline (one draw; RendererSoft::draw, RendererSoft.cpp:136-164), queued on
 * one stream with no host round trip:
 *   k_vertex      vertex shader, clip mask, perspective divide, viewport   (:170-190, :971-992)
 *   k_setup       assembly, clip test, face culling, bounding box, quad range, bin counts
 *                 (:192-204, :277-299, :1005-1019)
 *   k_clip        Sutherland-Hodgman against the six planes, re-shading new vertices (:489-559, :956-970)
 *   k_clip_fans   the fan triangles of clipped polygons, after all originals in clipped-triangle
 *                 order (:550-558, :227)
 *   k_bin_scan    primitive count, start of every bin
 *   k_bin_emit    primitive ids into their bins
 *   k_raster_vis  draws whose result follows from depths and order: visibility, then one
 *                 shading per visible fragment (k_shade)
 *   k_ord_list    draws that write no depth: each pixel's fragments listed in API order,
 *                 then shaded all at once (k_ord_shade) and written in order (k_ord_apply);
 *                 tiles that do not fit take the ordered raster (k_ord_rest)
 *   k_raster_ord  every other draw: fragments in API order per pixel, shading deferred and
 *                 batched per wavefront, per-sample operations in order (:720-912)
 */
#pragma once
#include "rhip_shading.h"

namespace rhip {

/* --------------------------------------------------------------------------- shading tables */

/* --------------------------------------------------------------------------- XCD placement
 *
 * The MI355X deals a launch's workgroups to its eight XCDs in turn (workgroup i
 * runs on XCD i mod 8), and each XCD has its own L2. Neighbouring blocks of work
 * share data (triangles share vertices, tiles share big triangles, fragments
 * share texels), so the kernels whose work items cost about the same (setup,
 * clipping, shading) renumber their workgroups to give each XCD one contiguous
 * range of the work; the rasters keep the round-robin order, which spreads tiles
 * of uneven cost over all XCDs. The renumbering is a bijection: which hardware
 * runs what changes, results do not. */
constexpr int kXcds = 8;

/* Logical block of workgroup b of n: XCD x gets one contiguous range. */
__device__ __forceinline__ int xcdBlock(int b, int n) {
  const int per = n / kXcds, extra = n % kXcds, x = b % kXcds, k = b / kXcds;
  return x * per + min(x, extra) + k;
}

/* For grid-stride loops over n items: this workgroup's XCD owns items [lo, hi),
 * a share in proportion to the workgroups it holds (none when the grid has fewer
 * than one per XCD); the workgroup is number `idx` of the `cnt` that XCD holds. */
struct XcdShare {
  int lo, hi, idx, cnt;
};
__device__ __forceinline__ XcdShare xcdShare(int n) {
  const int g = (int)gridDim.x, x = blockIdx.x % kXcds;
  const int first = xcdBlock(x, g); /* logical number of this XCD's first workgroup */
  XcdShare r;
  r.idx = blockIdx.x / kXcds;
  r.cnt = (g - x + kXcds - 1) / kXcds;
  r.lo = (int)((long long)n * first / g);
  r.hi = (int)((long long)n * (first + r.cnt) / g);
  return r;
}

/* --------------------------------------------------------------------------- clears */

/* Words [0, n) of p set to v by thread i of `stride`, 16 bytes a store when p is 16-byte aligned. */
__device__ __forceinline__ void fillWords(uint32_t *p, size_t n, uint32_t v, size_t i, size_t stride) {
  if (((uintptr_t)p & 15u) == 0) {
    const size_t n4 = n >> 2;
    const uint4 w = make_uint4(v, v, v, v);
    for (size_t j = i; j < n4; j += stride) reinterpret_cast<uint4 *>(p)[j] = w;
    for (size_t j = (n4 << 2) + i; j < n; j += stride) p[j] = v;
  } else {
    for (size_t j = i; j < n; j += stride) p[j] = v;
  }
}

/* A render pass's clears in one launch: colour words [0, n0) of p0 to v0 and depth words [0, n1) of p1 to v1. */
static __global__ void k_fill2_u32(uint32_t *p0, size_t n0, uint32_t v0, uint32_t *p1, size_t n1, uint32_t v1) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
  fillWords(p0, n0, v0, i, stride);
  fillWords(p1, n1, v1, i, stride);
}

/* --------------------------------------------------------------------------- vertex stage */

/* Float offset of stored varying group g within a vertex's record. */
__device__ __forceinline__ int varyGroupOffset(const DrawParams &P, int g) {
  return 4 * __popc(P.vary_groups & ((1u << g) - 1u));
}

/* Stores the groups of vertex i's varyings that a fragment shader reads. */
__device__ __forceinline__ void storeVaryings(const DrawParams &P, int i, const float *v) {
  float *o = P.vary + (size_t)i * P.vary_stride;
#pragma unroll
  for (int g = 0; g < kVaryStride / 4; g++)
    if ((P.vary_groups >> g) & 1u)
      *(float4 *)(o + varyGroupOffset(P, g)) = make_float4(v[4 * g], v[4 * g + 1], v[4 * g + 2], v[4 * g + 3]);
}

__device__ __forceinline__ void vertexBody(const DrawParams &P, int blk) {
  int i = blk * blockDim.x + threadIdx.x;
  if (i >= P.nverts) return;
  float a[16];
  const float *src = P.verts + (size_t)i * P.stride_f;
  if ((P.stride_f & 3) == 0) { /* 16-B aligned records: four-float loads */
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const float4 q = 4 * k < P.stride_f ? ((const float4 *)src)[k] : make_float4(0.f, 0.f, 0.f, 0.f);
      a[4 * k] = q.x, a[4 * k + 1] = q.y, a[4 * k + 2] = q.z, a[4 * k + 3] = q.w;
    }
  } else {
    for (int k = 0; k < 16; k++) a[k] = k < P.stride_f ? src[k] : 0.f;
  }
  float v[kVaryStride];
#pragma unroll
  for (int k = 0; k < kVaryStride; k++) v[k] = 0.f;
  float4 c = runVS(P, a, v);
  P.clip[i] = c;
  P.mask[i] = clipMask(c);
  P.frag[i] = toScreen(P, c);
  storeVaryings(P, i, v);
}

/* The vertex stage of one draw; it also clears the draw's bin counts, which k_setup fills next. */
static __global__ __launch_bounds__(256) void k_vertex(DrawParams P) {
  vertexBody(P, blockIdx.x);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.nbins; i += gridDim.x * blockDim.x) P.bin_count[i] = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < fanChunks(P.nblk); i += gridDim.x * blockDim.x)
    P.chunk_fans[i] = 0;
}

/* --------------------------------------------------------------------------- block helpers */

/* Inclusive prefix sum over the N threads of a block (N = blockDim.x). */
template <int N>
__device__ __forceinline__ int blockInclusiveScan(int v, int *part) {
  part[threadIdx.x] = v;
  __syncthreads();
#pragma unroll 1
  for (int off = 1; off < N; off <<= 1) {
    const int t = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  const int r = part[threadIdx.x];
  __syncthreads();
  return r;
}

/* Inclusive prefix sum over the 64 lanes of a wavefront. */
__device__ __forceinline__ int waveInclusiveScan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t waveMaxU(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t waveMinU(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

/* This workgroup's copy of statistics counter i. */
__device__ __forceinline__ unsigned long long *counterAt(const DrawParams &P, int i) {
  return P.counters + (blockIdx.x & (kCounterSlots - 1)) * kCounterStride + i;
}

/* Adds the lanes' v (each below 2^26) to counter i, one atomic per wavefront;
 * every lane of the wavefront calls it. */
__device__ __forceinline__ void waveAddCounter(const DrawParams &P, int i, unsigned long long v) {
  uint32_t w = (uint32_t)v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) w += (uint32_t)__shfl_xor((int)w, o, 64);
  if ((threadIdx.x & 63) == 0 && w) atomicAdd(counterAt(P, i), (unsigned long long)w);
}

/* Orders LDS accesses of the lanes of one wavefront around this point. */
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Exclusive prefix sum of in[0..n) into out[0..n) by one block of 1024 threads,
 * 8192 entries at a time with coalesced loads and stores (the next chunk's loads
 * issued before this chunk's scan); also clears clear[0..n) when given. Returns
 * the total (to every thread). */
__device__ int blockScanArray(const int *in, int *out, int n, int *clear) {
  __shared__ int buf[8192];
  __shared__ int wsum[16];
  __shared__ int s_carry;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t == 0) s_carry = 0;
  int nxt[8];
#pragma unroll
  for (int k = 0; k < 8; k++) nxt[k] = k * 1024 + t < n ? in[k * 1024 + t] : 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += 8192) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      buf[k * 1024 + t] = nxt[k];
      const int i = c0 + 8192 + k * 1024 + t;
      nxt[k] = i < n ? in[i] : 0;
    }
    __syncthreads();
    int v[8], sum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      v[k] = buf[t * 8 + k];
      sum += v[k];
    }
    const int inc = waveInclusiveScan(sum);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int run = s_carry + inc - sum;
    for (int w = 0; w < wave; w++) run += wsum[w];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      buf[t * 8 + k] = run;
      run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int i = c0 + k * 1024 + t;
      if (i < n) {
        out[i] = buf[k * 1024 + t];
        if (clear) clear[i] = 0;
      }
    }
    if (t == 1023) s_carry = run;
    __syncthreads();
  }
  return s_carry;
}

/* Ascending bitonic sort of n (a power of two) words in LDS by the whole block. */
__device__ void blockSortU32(uint32_t *a, int n) {
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const uint32_t x = a[i], y = a[l];
          if ((x > y) == ((i & k) == 0)) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
}

/* Ascending order of the m distinct ids in win[0..m) (LDS, at least a power of
 * two above m long), by the whole block: each id's rank is the count of smaller
 * ones when every id has a thread, a bitonic sort beyond. */
__device__ void sortIds(uint32_t *win, int m) {
  if (m <= (int)blockDim.x) {
    const uint32_t v = (int)threadIdx.x < m ? win[threadIdx.x] : 0u;
    int rk = 0;
    for (int i = 0; i < m; i++) rk += win[i] < v ? 1 : 0;
    __syncthreads();
    if ((int)threadIdx.x < m) win[rk] = v;
    __syncthreads();
    return;
  }
  int mp = 1;
  while (mp < m) mp <<= 1;
  for (int i = m + (int)threadIdx.x; i < mp; i += blockDim.x) win[i] = 0xffffffffu;
  __syncthreads();
  blockSortU32(win, mp);
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

/* Tiles touched by a live primitive's quads, clipped to the framebuffer. */
__device__ __forceinline__ void tileRect(const DrawParams &P, const PrimSetup &s, int &tx0, int &ty0, int &tx1,
                                         int &ty1) {
  tx0 = max(s.qx0, 0) / P.tile;
  tx1 = min(s.qx1, P.fb_w - 1) / P.tile;
  ty0 = max(s.qy0, 0) / P.tile;
  ty1 = min(s.qy1, P.fb_h - 1) / P.tile;
}

/* Which bins list a primitive: its tiles (0), its coarse bins, or the global bin. */
__device__ __forceinline__ int binClass(const DrawParams &P, int tx0, int ty0, int tx1, int ty1) {
  if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) <= kBinFine) return 0;
  const int c = P.coarse_tiles;
  if ((tx1 / c - tx0 / c + 1) * (ty1 / c - ty0 / c + 1) <= kBinCoarse) return kPrimCoarse;
  return kPrimGlobal;
}

/* Counts (EMIT false) or lists (EMIT true) primitive `id` in each of its bins. */
template <bool EMIT>
__device__ __forceinline__ void binVisit(const DrawParams &P, int id, int cls, int tx0, int ty0, int tx1, int ty1) {
  auto put = [&](int b) {
    if (EMIT) {
      const int slot = atomicAdd(&P.bin_cursor[b], 1);
      P.bin_items[P.bin_start[b] + slot] = id;
    } else {
      atomicAdd(&P.bin_count[b], 1);
    }
  };
  if (cls == kPrimGlobal) {
    put(P.nbins - 1);
  } else if (cls == kPrimCoarse) {
    for (int cy = ty0 / P.coarse_tiles; cy <= ty1 / P.coarse_tiles; cy++)
      for (int cx = tx0 / P.coarse_tiles; cx <= tx1 / P.coarse_tiles; cx++) put(P.ntiles + cy * P.coarse_x + cx);
  } else {
    for (int ty = ty0; ty <= ty1; ty++)
      for (int tx = tx0; tx <= tx1; tx++) put(ty * P.tiles_x + tx);
  }
}

/* The bins of a live primitive: class (-1 none, 0 tiles, kPrimCoarse,
 * kPrimGlobal) and its tile rectangle. */
struct BinRect {
  int cls, tx0, ty0, tx1, ty1;
};

/* Fills the setup record of primitive `id` from three screen-space vertices;
 * returns the bins it goes to (the caller counts it there). */
__device__ BinRect setupPrim(const DrawParams &P, int id, int i0, int i1, int i2, int src = -1) {
  PrimSetup s;
  s.src = src;
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
  /* every depth a sample can take: a convex combination of the vertex depths,
   * widened by the rounding of the barycentrics and of the dot product */
  if (a.z != a.z || b.z != b.z || c.z != c.z) {
    s.zlo = -INFINITY;
    s.zhi = INFINITY;
  } else {
    const float sl = 16.f * 1.19209290e-07f * (fabsf(a.z) + fabsf(b.z) + fabsf(c.z)) + 1e-7f;
    s.zlo = fminf(fminf(a.z, b.z), c.z) - sl;
    s.zhi = fmaxf(fmaxf(a.z, b.z), c.z) + sl;
  }
  int flags = front ? kPrimFront : 0;
  BinRect r = {-1, 0, 0, 0, 0};
  if (live && fx <= lx && fy <= ly) {
    const int x0 = max(fx, 0), x1 = min(lx + 1, P.fb_w - 1);
    const int y0 = max(fy, 0), y1 = min(ly + 1, P.fb_h - 1);
    if (x0 <= x1 && y0 <= y1) {
      r = {0, x0 / P.tile, y0 / P.tile, x1 / P.tile, y1 / P.tile};
      r.cls = binClass(P, r.tx0, r.ty0, r.tx1, r.ty1);
      flags |= kPrimLive | r.cls;
    }
  }
  s.flags = flags;
  if (flags & kPrimLive)
    P.prims[id] = s;
  else
    P.prims[id].flags = flags; /* nothing reads more of a primitive that is not live */
  return r;
}

/* Counts a primitive in its bins with plain atomics (for divergent callers). */
__device__ __forceinline__ void binCount(const DrawParams &P, const BinRect &r) {
  if (r.cls >= 0) binVisit<false>(P, 0, r.cls, r.tx0, r.ty0, r.tx1, r.ty1);
}

/* Bins of a primitive's BinRect: how many, and the k-th (row-major). */
__device__ __forceinline__ int binsOf(const DrawParams &P, const BinRect &r) {
  if (r.cls < 0) return 0;
  if (r.cls == kPrimGlobal) return 1;
  if (r.cls == kPrimCoarse) {
    const int c = P.coarse_tiles;
    return (r.tx1 / c - r.tx0 / c + 1) * (r.ty1 / c - r.ty0 / c + 1);
  }
  return (r.tx1 - r.tx0 + 1) * (r.ty1 - r.ty0 + 1);
}
__device__ __forceinline__ int binAt(const DrawParams &P, const BinRect &r, int k) {
  if (r.cls == kPrimGlobal) return P.nbins - 1;
  if (r.cls == kPrimCoarse) {
    const int c = P.coarse_tiles, cx0 = r.tx0 / c, w = r.tx1 / c - cx0 + 1;
    return P.ntiles + (r.ty0 / c + k / w) * P.coarse_x + cx0 + k % w;
  }
  const int w = r.tx1 - r.tx0 + 1;
  return (r.ty0 + k / w) * P.tiles_x + r.tx0 + k % w;
}

/* A workgroup's bins in LDS: open addressing, a bin's slot found within
 * kBinProbe probes or not at all (slots only ever go from empty to a bin, so a
 * bin one thread could not place no other thread can either). */
constexpr int kBinHash = 512, kBinProbe = 16;
struct BinHash {
  int bin[kBinHash];
  int cnt[kBinHash];
  int base[kBinHash];
};

__device__ __forceinline__ int binSlot(BinHash &H, int b, bool insert) {
  const int h = (int)(((uint32_t)b * 2654435761u) >> 23);
  for (int p = 0; p < kBinProbe; p++) {
    const int e = (h + p) & (kBinHash - 1);
    const int cur = __hip_atomic_load(&H.bin[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == b) return e;
    if (cur == -1) {
      if (!insert) return -1;
      const int old = atomicCAS(&H.bin[e], -1, b);
      if (old == -1 || old == b) return e;
    }
  }
  return -1;
}

/* Counts (EMIT false) or lists (EMIT true) every thread's primitive in its bins
 * with one global atomic per distinct bin of the workgroup: pairs are counted
 * per bin in LDS, each bin reserves its range once, and the pairs fill it. A
 * bin the table cannot hold takes a global atomic per pair. Every thread of
 * the workgroup must call it (r.cls < 0: no bins). */
template <bool EMIT>
__device__ void binBlock(const DrawParams &P, int id, const BinRect &r) {
  __shared__ BinHash H;
  for (int i = threadIdx.x; i < kBinHash; i += blockDim.x) {
    H.bin[i] = -1;
    H.cnt[i] = 0;
  }
  __syncthreads();
  const int nb = binsOf(P, r);
  for (int k = 0; k < nb; k++) {
    const int b = binAt(P, r, k);
    const int e = binSlot(H, b, true);
    if (e >= 0) {
      atomicAdd(&H.cnt[e], 1);
    } else if (EMIT) {
      P.bin_items[P.bin_start[b] + atomicAdd(&P.bin_cursor[b], 1)] = id;
    } else {
      atomicAdd(&P.bin_count[b], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBinHash; i += blockDim.x) {
    const int c = H.cnt[i];
    if (!c) continue;
    if (EMIT) {
      H.base[i] = P.bin_start[H.bin[i]] + atomicAdd(&P.bin_cursor[H.bin[i]], c);
      H.cnt[i] = 0; /* now the fill position within the reserved range */
    } else {
      atomicAdd(&P.bin_count[H.bin[i]], c);
    }
  }
  if (EMIT) {
    __syncthreads();
    for (int k = 0; k < nb; k++) {
      const int e = binSlot(H, binAt(P, r, k), false);
      if (e >= 0) P.bin_items[H.base[e] + atomicAdd(&H.cnt[e], 1)] = id;
    }
  }
  __syncthreads(); /* the table is free for the next call */
}

__device__ __forceinline__ bool validTri(const DrawParams &P, int i0, int i1, int i2) {
  return (unsigned)i0 < (unsigned)P.nverts && (unsigned)i1 < (unsigned)P.nverts && (unsigned)i2 < (unsigned)P.nverts;
}

/* processPolygonAssembly (RendererSoft.cpp:428-437) + the clip test: triangles
 * with a vertex outside the frustum are set up after clipping. Returns the
 * block's count of those (block-uniform). */
__device__ __forceinline__ int setupBody(const DrawParams &P, int blk) {
  const int t = blk * blockDim.x + threadIdx.x;
  int needs = 0;
  BinRect r = {-1, 0, 0, 0, 0};
  if (t < P.ntris) {
    const int i0 = P.idx[3 * t], i1 = P.idx[3 * t + 1], i2 = P.idx[3 * t + 2];
    if (!validTri(P, i0, i1, i2)) {
      P.prims[t].flags = 0;
    } else if (P.mask[i0] | P.mask[i1] | P.mask[i2]) {
      needs = 1;
      P.prims[t].flags = 0;
    } else {
      r = setupPrim(P, t, i0, i1, i2);
    }
  }
  binBlock<false>(P, t, r);
  const int n = __syncthreads_count(needs);
  if (threadIdx.x == 0) P.blk_clip[blk] = n;
  return n;
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

/* clippingTriangle (RendererSoft.cpp:489-559) for triangle t; its new vertices
 * take slots nverts + t * kMaxClipVerts onward. Returns the fan triangles it appends. */
__device__ int clipTriangle(const DrawParams &P, int t) {
  int *poly = P.clip_poly + (size_t)t * (kMaxPoly + 1);
  const int i0 = P.idx[3 * t], i1 = P.idx[3 * t + 1], i2 = P.idx[3 * t + 2];
  const int mask = validTri(P, i0, i1, i2) ? (P.mask[i0] | P.mask[i1] | P.mask[i2]) : 0;
  if (!mask) {
    poly[kMaxPoly] = 0;
    return 0;
  }
  atomicAdd(counterAt(P, 5), 1ull);
  const int k = t;
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
        storeVaryings(P, nv, v);
        out[nout++] = nv;
      }
      ip = ic;
      dp = d;
    }
    if (full) {
      /* more vertices than a convex polygon can have: rounding produced extra
       * plane crossings; the triangle is dropped and the overflow counted */
      atomicAdd(counterAt(P, 2), 1ull);
      P.errors[kErrClip] = 1u;
      break;
    }
    for (int q = 0; q < nout; q++) in[q] = out[q];
    nin = nout;
  }
  if (full || nin < 3) { /* fewer than 3 left: the reference reads past its index list here */
    poly[kMaxPoly] = 0;
    P.prims[t].flags = 0;
    return 0;
  }
  for (int q = 0; q < nin; q++) poly[q] = in[q];
  poly[kMaxPoly] = nin;
  binCount(P, setupPrim(P, t, in[0], in[1], in[2], t));
  return nin - 3;
}

/* One block per 256-triangle block of k_setup; blocks with nothing to clip return at once. */
__device__ __forceinline__ void clipBody(const DrawParams &P, int b) {
  __shared__ int fans;
  if (P.blk_clip[b] == 0) {
    if (threadIdx.x == 0) P.blk_fans[b] = 0;
    return;
  }
  if (threadIdx.x == 0) fans = 0;
  __syncthreads();
  const int t = b * blockDim.x + threadIdx.x;
  const int n = t < P.ntris ? clipTriangle(P, t) : 0;
  if (n) atomicAdd(&fans, n);
  __syncthreads();
  if (threadIdx.x == 0) P.blk_fans[b] = fans;
}

/* A single draw's setup: the block's triangles set up and binned, then those
 * that need it clipped by the same block (one launch instead of two); the
 * block's fan count also goes into its 64-block chunk's sum. */
static __global__ __launch_bounds__(256) void k_setup(DrawParams P) {
  __shared__ int fans;
  const int b = xcdBlock(blockIdx.x, gridDim.x);
  if (setupBody(P, b) == 0) {
    if (threadIdx.x == 0) P.blk_fans[b] = 0;
    return;
  }
  if (threadIdx.x == 0) fans = 0;
  __syncthreads();
  const int t = b * blockDim.x + threadIdx.x;
  const int n = t < P.ntris ? clipTriangle(P, t) : 0;
  if (n) atomicAdd(&fans, n);
  __syncthreads();
  if (threadIdx.x == 0) {
    P.blk_fans[b] = fans;
    if (fans) atomicAdd(&P.chunk_fans[b >> 6], fans);
  }
}

/* One block of 1024: exclusive offsets of every block's fan triangles, and the
 * primitive count (the batched views' stage; a single draw's blocks take their
 * own offsets in clipFansBody and its totals in fanTotalsBody). */
__device__ __forceinline__ void scanFansBody(const DrawParams &P) {
  const int total = blockScanArray(P.blk_fans, P.blk_fan_off, P.nblk, nullptr);
  if (threadIdx.x == 0) {
    P.dcount[0] = total;
    P.dcount[1] = P.ntris + total;
    P.dcount[2] = 0;
    P.dcount[3] = 0;
    P.dcount[7] = 0;
    P.dcount[8] = 0;
  }
}


/* fans appended after all original primitives, in clipped-triangle order (:550-558).
 * OWN_OFFSET: a block with fans sums the earlier blocks' fan counts itself, from
 * the sums of whole 64-block chunks and its own chunk's earlier blocks, instead
 * of reading scanFansBody's offsets. */
template <bool OWN_OFFSET>
__device__ __forceinline__ void clipFansBody(const DrawParams &P, int b) {
  __shared__ int part[256];
  __shared__ int s_off;
  if (P.blk_fans[b] == 0) return;
  int off;
  if (OWN_OFFSET) {
    const int c = b >> 6;
    int acc = threadIdx.x < (b & 63) ? P.blk_fans[(c << 6) + threadIdx.x] : 0;
    for (int i = threadIdx.x; i < c; i += blockDim.x) acc += P.chunk_fans[i];
    acc = blockInclusiveScan<256>(acc, part);
    if (threadIdx.x == 255) s_off = acc;
    __syncthreads();
    off = s_off;
  } else {
    off = P.blk_fan_off[b];
  }
  const int t = b * blockDim.x + threadIdx.x;
  const int *poly = P.clip_poly + (size_t)t * (kMaxPoly + 1);
  const int n = t < P.ntris ? poly[kMaxPoly] : 0;
  const int nf = n > 3 ? n - 3 : 0;
  const int inc = blockInclusiveScan<256>(nf, part);
  const int base = P.ntris + off + inc - nf;
  if (n > 3) P.fan_first[t] = base;
  for (int i = 3; i < n; i++) binCount(P, setupPrim(P, base + i - 3, poly[0], poly[i - 1], poly[i], t));
}

static __global__ __launch_bounds__(256) void k_clip_fans(DrawParams P) {
  clipFansBody<true>(P, xcdBlock(blockIdx.x, gridDim.x));
}

/* --------------------------------------------------------------------------- binning */

/* One block of 1024: bin starts from the counts setup made; cursors cleared. */
__device__ __forceinline__ void binScanBody(const DrawParams &P) {
  const int total = blockScanArray(P.bin_count, P.bin_start, P.nbins, P.bin_cursor);
  if (threadIdx.x == 0) {
    P.bin_start[P.nbins] = total;
    atomicAdd(counterAt(P, 3), (unsigned long long)total);
    atomicAdd(counterAt(P, 4), (unsigned long long)P.dcount[1]);
  }
}

/* What scanFansBody writes besides the per-block offsets: the fan total, the
 * primitive count and the cleared raster counters. */
__device__ __forceinline__ void fanTotalsBody(const DrawParams &P) {
  __shared__ int wsum[16];
  int acc = 0;
  for (int i = threadIdx.x; i < fanChunks(P.nblk); i += blockDim.x) acc += P.chunk_fans[i];
  const int inc = waveInclusiveScan(acc);
  if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) total += wsum[w];
    P.dcount[0] = total;
    P.dcount[1] = P.ntris + total;
    P.dcount[2] = 0;
    P.dcount[3] = 0;
    P.dcount[7] = 0;
    P.dcount[8] = 0;
  }
}

/* One block of 1024 after k_clip_fans: the fan totals, then the bin starts. */
static __global__ __launch_bounds__(1024) void k_bin_scan(DrawParams P) {
  fanTotalsBody(P);
  binScanBody(P);
}

/* Lists every live primitive in its bins (wavefronts walk the ids together). */
__device__ __forceinline__ void binEmitBody(const DrawParams &P, int blk, int nblocks) {
  const int n = P.dcount[1];
  for (int b0 = blk * blockDim.x; b0 < n; b0 += nblocks * blockDim.x) {
    const int id = b0 + threadIdx.x;
    BinRect r = {-1, 0, 0, 0, 0};
    if (id < n) {
      const PrimSetup &ps = P.prims[id];
      if (ps.flags & kPrimLive) {
        tileRect(P, ps, r.tx0, r.ty0, r.tx1, r.ty1);
        r.cls = ps.flags & (kPrimCoarse | kPrimGlobal);
      }
    }
    binBlock<true>(P, id, r);
  }
}

static __global__ __launch_bounds__(256) void k_bin_emit(DrawParams P) { binEmitBody(P, blockIdx.x, gridDim.x); }

/* The entries of one tile: its own bin, its coarse bin, the global bin. */
struct TileBins {
  int s[3], n[3];
  __device__ __forceinline__ int total() const { return n[0] + n[1] + n[2]; }
  __device__ __forceinline__ int at(const DrawParams &P, int i) const {
    if (i < n[0]) return P.bin_items[s[0] + i];
    i -= n[0];
    if (i < n[1]) return P.bin_items[s[1] + i];
    return P.bin_items[s[2] + i - n[1]];
  }
};

__device__ __forceinline__ TileBins tileBins(const DrawParams &P, int tile) {
  const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
  const int c = P.ntiles + (ty / P.coarse_tiles) * P.coarse_x + tx / P.coarse_tiles;
  const int g = P.nbins - 1;
  TileBins b;
  b.s[0] = P.bin_start[tile];
  b.n[0] = P.bin_start[tile + 1] - b.s[0];
  b.s[1] = P.bin_start[c];
  b.n[1] = P.bin_start[c + 1] - b.s[1];
  b.s[2] = P.bin_start[g];
  b.n[2] = P.bin_start[g + 1] - b.s[2];
  return b;
}

/* --------------------------------------------------------------------------- fine raster helpers */

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

/* True only for a sample bary() finds outside, decided without its divisions:
 * from the same numerators, ux = nx/uz or uy = ny/uz is negative and too large
 * to round to zero (a live primitive's |uz| is below 2^31 and nonzero), or
 * their sum exceeds 1 by far more than the rounding of the two quotients, their
 * sum and 1 - sum can take back. Samples near an edge and NaNs give false. */
__device__ __forceinline__ bool sampleSurelyOut(const Tri &T, float px, float py) {
  const float a2 = T.v0x - px, b2 = T.v0y - py;
  const float nx = fmaf(T.a1, b2, -(a2 * T.b1));
  const float ny = fmaf(a2, T.b0, -(T.a0 * b2));
  const float sg = T.uz < 0.f ? -1.f : 1.f;
  const float sx = nx * sg, sy = ny * sg, au = fabsf(T.uz);
  if (sx <= -0x1p-90f || sy <= -0x1p-90f) return true;
  return (sx + sy) - au > 0x1p-18f * (fabsf(sx) + fabsf(sy) + au);
}

/* A pixel none of whose samples (the centre without MSAA) any triangle test can cover. */
template <int S>
__device__ __forceinline__ bool pixelSurelyOut(const Tri &T, int x, int y) {
  const float fx = (float)x, fy = (float)y;
  if (S == 4)
    return sampleSurelyOut(T, 0.375f + fx, 0.875f + fy) && sampleSurelyOut(T, 0.875f + fx, 0.625f + fy) &&
           sampleSurelyOut(T, 0.125f + fx, 0.375f + fy) && sampleSurelyOut(T, 0.625f + fx, 0.125f + fy);
  return sampleSurelyOut(T, fx + 0.5f, fy + 0.5f);
}

/* Position of MSAA-4x sample s in its pixel (RendererInternal.h:59-186) */
__device__ __forceinline__ float sampleX(int s) { return s == 0 ? 0.375f : s == 1 ? 0.875f : s == 2 ? 0.125f : 0.625f; }
__device__ __forceinline__ float sampleY(int s) { return s == 0 ? 0.875f : s == 1 ? 0.625f : s == 2 ? 0.375f : 0.125f; }

/* dpps with imm 0xff over (b0, b1, b2, 0) . (v0, v1, v2, 0) */
__device__ __forceinline__ float dpps3(float4 b, const float *v) {
  return (b.x * v[0] + b.y * v[1]) + (b.z * v[2] + b.w * 0.f);
}

/* Quad origins that contain pixel coordinate p on one axis, in the order the
 * reference visits them (increasing block): q0, then q1. Returns 0..2. */
__device__ __forceinline__ int axisQuads(float mn, float mx, int bc, int p, int &q0, int &q1) {
  int n = 0;
  q0 = q1 = 0;
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
    const bool hit = found && ((x - sb - 1) & 1) == 0;
    q1 = (hit && n == 1) ? x : q1;
    q0 = (hit && n == 0) ? x : q0;
    n += hit ? 1 : 0;
  }
  return n;
}

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

template <int SHADING, bool U8OUT = false>
__device__ __forceinline__ RGBAf runFS(const DrawParams &P, const Frag &f) {
  if (SHADING == RHS_SHADING_BASIC) {
    return {P.u.base_color[0], P.u.base_color[1], P.u.base_color[2], P.u.base_color[3]};
  } else if (SHADING == RHS_SHADING_BLINN_PHONG) {
    return fsBlinnPhong(P, f, false);
  } else if (SHADING == RHS_SHADING_BLINN_PHONG_DIR) {
    return fsBlinnPhong(P, f, true);
  } else if (SHADING == RHS_SHADING_PBR) {
    return fsPbr<U8OUT>(P, f);
  } else if (SHADING == kShadingEnv) {
    return fsEnv(P, f);
  } else {
    return {f.v[0], f.v[1], f.v[2], f.v[3]};
  }
}

/* Varyings a fragment shader reads (the others are never interpolated). */
template <int SHADING>
__device__ __forceinline__ bool varyUsed(const DrawParams &P, int j) {
  return varySlotUsed(SHADING == kShadingEnv ? P.shading : SHADING, P.defines, P.u.enable_shadow, j);
}

/* Shading of one fragment: `b` is the raw (screen-space) barycentric of its shading
 * sample, `inst` which of the triangle's quads holding the pixel it was produced in
 * (low bit: x, high bit: y, in the reference's visiting order). RendererSoft.cpp:800-851. */
template <int S, int SHADING>
__device__ __forceinline__ Frag makeFrag(const DrawParams &P, const PrimSetup &ps, const Tri &T, int px, int py,
                                         float4 b, int inst, int need_deriv) {
  const float w = dpps3(b, T.w);
  const float iw = 1.f / w;
  b = make_float4(b.x * (iw * T.w[0]), b.y * (iw * T.w[1]), b.z * (iw * T.w[2]), b.w * (iw * 1.f));
  const float *va = P.vary + (size_t)ps.v[0] * P.vary_stride;
  const float *vb = P.vary + (size_t)ps.v[1] * P.vary_stride;
  const float *vc = P.vary + (size_t)ps.v[2] * P.vary_stride;
  Frag fr;
  float4 g0a = make_float4(0.f, 0.f, 0.f, 0.f), g0b = g0a, g0c = g0a; /* group 0 (texture coordinates) */
#pragma unroll
  for (int g = 0; g < kVaryStride / 4; g++) {
    if (!((P.vary_groups >> g) & 1u)) {
#pragma unroll
      for (int k = 0; k < 4; k++) fr.v[4 * g + k] = 0.f;
      continue;
    }
    /* a stored group is read whole, with one four-float load per vertex */
    const int o = varyGroupOffset(P, g);
    const float4 qa = *(const float4 *)(va + o), qb = *(const float4 *)(vb + o), qc = *(const float4 *)(vc + o);
    if (g == 0) g0a = qa, g0b = qb, g0c = qc;
    const float ea[4] = {qa.x, qa.y, qa.z, qa.w}, eb[4] = {qb.x, qb.y, qb.z, qb.w}, ec[4] = {qc.x, qc.y, qc.z, qc.w};
#pragma unroll
    for (int k = 0; k < 4; k++)
      fr.v[4 * g + k] = varyUsed<SHADING>(P, 4 * g + k) ? interp(b, ea[k], eb[k], ec[k]) : 0.f;
  }
  fr.tc[0] = fr.tc[1] = fr.tc[2] = {0.f, 0.f};
  if (need_deriv && (P.vary_groups & 1u)) { /* texture coordinates are slots 0-1 (group 0, offset 0) */
    /* coarse derivatives from quad pixels 0, 1, 2 (RendererSoft.cpp:329-334) */
    int qx0, qx1, qy0, qy1;
    axisQuads(ps.min_x, ps.max_x, ps.bcx, px, qx0, qx1);
    axisQuads(ps.min_y, ps.max_y, ps.bcy, py, qy0, qy1);
    const int qx = (inst & 1) ? qx1 : qx0, qy = (inst >> 1) ? qy1 : qy0;
    const int k = (px - qx) + 2 * (py - qy);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j == k) {
        fr.tc[j] = {fr.v[0], fr.v[1]};
      } else {
        const float4 pb = shadingBc<S>(T, (float)(qx + (j & 1)), (float)(qy + (j >> 1)));
        fr.tc[j] = {interp(pb, g0a.x, g0b.x, g0c.x), interp(pb, g0a.y, g0b.y, g0c.y)};
      }
    }
  }
  return fr;
}

template <int S, int SHADING, bool U8OUT = false>
__device__ __forceinline__ RGBAf shadeAt(const DrawParams &P, const PrimSetup &ps, const Tri &T, int px, int py,
                                         float4 b, int inst, int need_deriv) {
  return runFS<SHADING, U8OUT>(P, makeFrag<S, SHADING>(P, ps, T, px, py, b, inst, need_deriv));
}

__device__ __forceinline__ uint32_t resolve4(const uint32_t *c) {
  /* multiSampleResolve (RendererSoft.cpp:880-912): u8(sum / 4) */
  int r[4];
#pragma unroll
  for (int ch = 0; ch < 4; ch++) r[ch] = (chan(c[0], ch) + chan(c[1], ch) + chan(c[2], ch) + chan(c[3], ch)) >> 2;
  return pack4(r[0], r[1], r[2], r[3]);
}

/* --------------------------------------------------------------------------- two-phase raster
 *
 * For draws without blending whose depth test writes depth with LESS, LEQUAL,
 * GREATER or GEQUAL, the colour a sample ends with is the fragment shader
 * output of the last fragment that passed its depth test, and which fragment
 * that is follows from depths and API order alone: the minimum (or maximum)
 * depth, ties going to the earliest fragment for the strict tests and the
 * latest for the others. Phase 1 finds that winner per sample with 64-bit LDS
 * atomics on key = (monotonic depth bits << 32) | order word: big triangles one
 * at a time with a lane per pixel, the others with a lane per triangle, all
 * after the hierarchical-Z cull. Phase 2 lists each pixel's distinct winners
 * and shades the list with every lane busy. The fragment shader's inputs depend
 * only on the triangle, the pixel and the 2x2 quad it is shaded in, so the
 * result equals the sequential one. */
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

/* depthKey's inverse (a zero key returns +0) */
__device__ __forceinline__ float keyDepth(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
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

/* --------------------------------------------------------------------------- hierarchical Z
 *
 * Per tile, one depth bound in depthKey order. For LESS / LEQUAL it is the
 * deepest depth any sample of the tile can hold at the end of the draw: the
 * deepest non-NaN depth already there, lowered to the far depth of every big
 * triangle that covers the whole tile (such a triangle puts a fragment at every
 * sample no deeper than that). A triangle whose nearest possible depth lies
 * beyond the bound cannot win or pass at any sample and is skipped; results are
 * unchanged. GREATER / GEQUAL mirror it. */

/* True if the triangle covers every sample of pixels [x0, x1] x [y0, y1] -- all
 * of which lie in [x0 + 1/8, x1 + 7/8] x [y0 + 1/8, y1 + 7/8] -- with 1/16 of a
 * pixel to spare (the single-precision edge functions err by about 1e-3 pixel
 * at 4K coordinates), inside its quads and, multisampled, inside the depth
 * range; [zmin, zmax] then bounds its depth over them. Evaluated in double. */
template <int S>
__device__ bool coversRect(const PrimSetup &ps, const float4 &a, const float4 &b, const float4 &c, int x0, int y0,
                           int x1, int y1, float &zmin, float &zmax) {
  if (ps.qx0 > x0 || ps.qx1 < x1 || ps.qy0 > y0 || ps.qy1 < y1) return false;
  const double ax = a.x, ay = a.y, bx = b.x, by = b.y, cx = c.x, cy = c.y;
  const double area = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
  if (!(fabs(area) > 1.0)) return false;
  const double sg = area > 0 ? 1.0 : -1.0;
  const double dzdx = ((double)(b.z - a.z) * (cy - ay) - (double)(c.z - a.z) * (by - ay)) / area;
  const double dzdy = ((double)(c.z - a.z) * (bx - ax) - (double)(b.z - a.z) * (cx - ax)) / area;
  double lo = INFINITY, hi = -INFINITY;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double X = (k & 1) ? x1 + 0.9375 : x0 + 0.0625, Y = (k & 2) ? y1 + 0.9375 : y0 + 0.0625;
    const double e0 = ((bx - ax) * (Y - ay) - (by - ay) * (X - ax)) * sg;
    const double e1 = ((cx - bx) * (Y - by) - (cy - by) * (X - bx)) * sg;
    const double e2 = ((ax - cx) * (Y - cy) - (ay - cy) * (X - cx)) * sg;
    if (!(e0 > 0 && e1 > 0 && e2 > 0)) return false;
    const double z = (double)a.z + dzdx * (X - ax) + dzdy * (Y - ay);
    lo = fmin(lo, z);
    hi = fmax(hi, z);
  }
  const double sl = 1e-5 + 16.0 * 1.19209290e-07 * (fabs((double)a.z) + fabs((double)b.z) + fabs((double)c.z));
  lo -= sl;
  hi += sl;
  if (!(lo == lo) || !(hi == hi)) return false;
  if (S == 4 && (lo < 0.0 || hi > 1.0)) return false;
  zmin = (float)lo;
  zmax = (float)hi;
  return true;
}

/* The same test for the convex polygon clipped triangle t became (RendererSoft.cpp:489-559);
 * the depth range is taken over the planes of all its fan triangles. It only chooses which
 * triangles go first: the bound that culls comes from the depths their fans actually leave. */
template <int S>
__device__ bool polyCoversRect(const DrawParams &P, int t, int x0, int y0, int x1, int y1, float &zmin,
                               float &zmax) {
  const int *poly = P.clip_poly + (size_t)t * (kMaxPoly + 1);
  const int n = poly[kMaxPoly];
  if (n < 3) return false;
  double vx[kMaxPoly], vy[kMaxPoly], vz[kMaxPoly];
  double area = 0, zs = 0;
  for (int k = 0; k < n; k++) {
    const float4 f = P.frag[poly[k]];
    vx[k] = f.x, vy[k] = f.y, vz[k] = f.z;
    zs += fabs((double)f.z);
  }
  for (int k = 0; k < n; k++) {
    const int k1 = k + 1 == n ? 0 : k + 1;
    area += vx[k] * vy[k1] - vx[k1] * vy[k];
  }
  if (!(fabs(area) > 2.0)) return false;
  const double sg = area > 0 ? 1.0 : -1.0;
  double lo = INFINITY, hi = -INFINITY;
  for (int c = 0; c < 4; c++) {
    const double X = (c & 1) ? x1 + 0.9375 : x0 + 0.0625, Y = (c & 2) ? y1 + 0.9375 : y0 + 0.0625;
    for (int k = 0; k < n; k++) {
      const int k1 = k + 1 == n ? 0 : k + 1;
      const double ex = vx[k1] - vx[k], ey = vy[k1] - vy[k];
      if (ex * ex + ey * ey < 1e-12) continue; /* clipping can repeat a vertex */
      const double e = (ex * (Y - vy[k]) - ey * (X - vx[k])) * sg;
      if (!(e > 0)) return false;
    }
    for (int k = 1; k + 1 < n; k++) {
      const double ar = (vx[k] - vx[0]) * (vy[k + 1] - vy[0]) - (vy[k] - vy[0]) * (vx[k + 1] - vx[0]);
      if (!(fabs(ar) > 0)) continue;
      const double dzdx = ((vz[k] - vz[0]) * (vy[k + 1] - vy[0]) - (vz[k + 1] - vz[0]) * (vy[k] - vy[0])) / ar;
      const double dzdy = ((vz[k + 1] - vz[0]) * (vx[k] - vx[0]) - (vz[k] - vz[0]) * (vx[k + 1] - vx[0])) / ar;
      const double z = vz[0] + dzdx * (X - vx[0]) + dzdy * (Y - vy[0]);
      lo = fmin(lo, z);
      hi = fmax(hi, z);
    }
  }
  const double sl = 1e-5 + 16.0 * 1.19209290e-07 * zs;
  lo -= sl;
  hi += sl;
  if (!(lo == lo) || !(hi == hi)) return false;
  if (S == 4 && (lo < 0.0 || hi > 1.0)) return false;
  zmin = (float)lo;
  zmax = (float)hi;
  return true;
}

__device__ __forceinline__ bool zCulled(const PrimSetup &ps, uint32_t bound, bool mx) {
  return mx ? depthKey(gclamp(ps.zhi, 0.f, 1.f)) < bound : depthKey(gclamp(ps.zlo, 0.f, 1.f)) > bound;
}

/* Bound from the depths a lane holds; combined over the block with atomicMax (LESS) / atomicMin. */
template <int S>
__device__ __forceinline__ void hizFromDepth(const float *dep, bool mx, uint32_t *bound) {
  uint32_t kb = mx ? 0xffffffffu : 0u;
  bool any = false;
#pragma unroll
  for (int s = 0; s < S; s++) {
    if (dep[s] != dep[s]) continue;
    const uint32_t k = depthKey(dep[s]);
    kb = mx ? min(kb, k) : max(kb, k);
    any = true;
  }
  if (!any) return;
  if (mx)
    atomicMin(bound, kb);
  else
    atomicMax(bound, kb);
}

/* Phase-1 visibility of triangle `id` at pixel (x, y): every quad of the triangle
 * holding the pixel is one fragment; of several, the test keeps the first or the last. */
template <int S, int MODE>
__device__ __forceinline__ void visPixel(const PrimSetup &ps, const Tri &T, int id, int x, int y,
                                         unsigned long long *kp) {
  const float lx[4] = {0.375f, 0.875f, 0.125f, 0.625f};
  const float ly[4] = {0.875f, 0.625f, 0.375f, 0.125f};
  int q0, q1;
  const int nqy = axisQuads(ps.min_y, ps.max_y, ps.bcy, y, q0, q1);
  if (!nqy) return;
  const int nqx = axisQuads(ps.min_x, ps.max_x, ps.bcx, x, q0, q1);
  if (!nqx) return;
  const uint32_t inst = visFirst<MODE>() ? 0u : (uint32_t)((nqy - 1) * 2 + (nqx - 1));
  const uint32_t low = visLow<MODE>((((uint32_t)id << 2) | inst) + 1u);
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
    if (!bary(T, fx + 0.5f, fy + 0.5f, bc)) return;
    const float z = gclamp(dpps3(bc, T.z), 0.f, 1.f);
    if (z != z) return;
    const unsigned long long key = ((unsigned long long)depthKey(z) << 32) | low;
    if (visMax<MODE>())
      atomicMax(kp, key);
    else
      atomicMin(kp, key);
  }
}


/* Big triangles (coarse and global bins) per 128x128 coarse bin, before the
 * tiles: the depth bound that covering ones give the whole bin, and the list
 * of those that bound does not rule out (or -1 when longer than kHizCap: the
 * tiles then read the full bins). One block per coarse bin. */
constexpr int kHizCap = 1024;

template <int S, bool MX>
__device__ __forceinline__ void hizCoarseBody(const DrawParams &P, int c) {
  __shared__ uint32_t s_bound, s_pz;
  __shared__ int s_cnt, s_pt;
  const int cx = c % P.coarse_x, cy = c / P.coarse_x;
  const int cpx = P.coarse_tiles * P.tile;
  const int x0 = cx * cpx, y0 = cy * cpx;
  const int x1 = min(x0 + cpx, P.fb_w) - 1, y1 = min(y0 + cpx, P.fb_h) - 1;
  const int gb = P.nbins - 1, cb = P.ntiles + c;
  const int cs = P.bin_start[cb], cn = P.bin_start[cb + 1] - cs;
  const int gs = P.bin_start[gb], gn = P.bin_start[gb + 1] - gs;
  const int n = cn + gn;
  const int lp = threadIdx.x;
  if (P.hiz_off) { /* no bound: every tile reads the whole coarse and global bins */
    if (lp == 0) {
      P.hiz_bound[c] = MX ? 0u : 0xffffffffu;
      P.hiz_count[c] = -1;
      P.hiz_seed[c] = -1;
      P.hiz_poly[c] = 0;
    }
    return;
  }
  if (lp == 0) {
    s_bound = MX ? 0u : 0xffffffffu;
    s_cnt = 0;
  }
  __syncthreads();
  for (int i = lp; i < n; i += blockDim.x) {
    const int id = i < cn ? P.bin_items[cs + i] : P.bin_items[gs + i - cn];
    const PrimSetup ps = P.prims[id];
    float zmin, zmax;
    if (coversRect<S>(ps, P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]], x0, y0, x1, y1, zmin, zmax)) {
      if (MX)
        atomicMax(&s_bound, depthKey(gclamp(zmin, 0.f, 1.f)));
      else
        atomicMin(&s_bound, depthKey(gclamp(zmax, 0.f, 1.f)));
    }
  }
  __syncthreads();
  /* the nearest clipped polygon covering the whole bin: its depth range over
   * the bin bounds the bin too when its live fans leave every sample of the bin
   * covered. Fans rasterize with float barycentrics and may be culled as
   * back-facing, so that is not certain: a tile checks it against the depths
   * the big triangles actually leave and, where it fails, rasterizes what this
   * bound ruled out (visTile, hiz_poly) */
  uint32_t my_pz = MX ? 0u : 0xffffffffu;
  int my_pt = -1;
  for (int i = lp; i < n; i += blockDim.x) {
    const int id = i < cn ? P.bin_items[cs + i] : P.bin_items[gs + i - cn];
    if (P.prims[id].src != id || !(P.prims[id].flags & kPrimLive)) continue;
    float zmin, zmax;
    if (polyCoversRect<S>(P, id, x0, y0, x1, y1, zmin, zmax)) {
      const uint32_t k = MX ? depthKey(gclamp(zmin, 0.f, 1.f)) : depthKey(gclamp(zmax, 0.f, 1.f));
      if (MX ? k > my_pz : k < my_pz) {
        my_pz = k;
        my_pt = id;
      }
    }
  }
  if (lp == 0) {
    s_pz = MX ? 0u : 0xffffffffu;
    s_pt = -1;
  }
  __syncthreads();
  if (my_pt >= 0) {
    if (MX)
      atomicMax(&s_pz, my_pz);
    else
      atomicMin(&s_pz, my_pz);
  }
  __syncthreads();
  if (my_pt >= 0 && my_pz == s_pz) s_pt = my_pt;
  __syncthreads();
  const uint32_t bound = s_pt < 0 ? s_bound : (MX ? max(s_bound, s_pz) : min(s_bound, s_pz));
  for (int i = lp; i < n; i += blockDim.x) {
    const int id = i < cn ? P.bin_items[cs + i] : P.bin_items[gs + i - cn];
    const PrimSetup &ps = P.prims[id];
    if (ps.qx0 > x1 || ps.qx1 < x0 || ps.qy0 > y1 || ps.qy1 < y0 || zCulled(ps, bound, MX)) continue;
    const int k = atomicAdd(&s_cnt, 1);
    if (k < kHizCap) P.hiz_surv[c * kHizCap + k] = id;
  }
  __syncthreads();
  if (lp == 0) {
    P.hiz_bound[c] = bound;
    P.hiz_count[c] = s_cnt <= kHizCap ? s_cnt : -1;
    P.hiz_seed[c] = s_pt;
    P.hiz_poly[c] = bound != s_bound ? 1 : 0;
  }
}

template <int S, bool MX>
__global__ __launch_bounds__(256) void k_hiz_coarse(DrawParams P) {
  hizCoarseBody<S, MX>(P, blockIdx.x);
}

/* One tile of the two-phase raster: TILE x TILE pixels, 64 * WAVES lanes,
 * PPL pixels per lane (lane l owns pixels l, l + lanes, ...). */
template <int S, int MODE, int TILE, int WAVES>
__device__ void visTile(const DrawParams &P, int tile) {
  constexpr int NPIX = TILE * TILE, NT = 64 * WAVES, PPL = NPIX / NT;
  static_assert(PPL >= 1 && PPL * NT == NPIX, "tile and wavefronts do not divide");
  constexpr bool MX = visMax<MODE>();
  __shared__ unsigned long long keys[NPIX * S];
  __shared__ int big[NT];
  __shared__ uint32_t big_z[NT];
  __shared__ uint32_t s_bound, s_bound2;
  __shared__ int s_n, s_base;
  const int tx0 = (tile % P.tiles_x) * TILE, ty0 = (tile / P.tiles_x) * TILE;
  const int X1 = min(tx0 + TILE, P.fb_w) - 1, Y1 = min(ty0 + TILE, P.fb_h) - 1;
  const int lp = threadIdx.x;
  const bool hasC = P.color != nullptr;
  const TileBins tb = tileBins(P, tile);
  const int cpx = P.coarse_tiles * TILE;
  const int coarse = (ty0 / cpx) * P.coarse_x + tx0 / cpx;
  const float lx[4] = {0.375f, 0.875f, 0.125f, 0.625f};
  const float ly[4] = {0.875f, 0.625f, 0.375f, 0.125f};

  float dep[PPL][S]; /* loaded while the bin counts are still on their way */
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    const int pl = lp + q * NT, px = tx0 + pl % TILE, py = ty0 + pl / TILE;
#pragma unroll
    for (int s = 0; s < S; s++)
      dep[q][s] = px < P.fb_w && py < P.fb_h ? P.depth[((size_t)py * P.fb_w + px) * S + s] : 0.f;
  }
  if (tb.total() == 0) return;
  __syncthreads(); /* LDS of a previous tile of this block is no longer read */
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    const int pl = lp + q * NT, px = tx0 + pl % TILE, py = ty0 + pl / TILE;
#pragma unroll
    for (int s = 0; s < S; s++) keys[pl * S + s] = visInit<MODE>(dep[q][s], px < P.fb_w && py < P.fb_h);
  }
  /* big triangles: the coarse bin's survivors, or all of its and the global bin's entries */
  const int hc = P.hiz_count[coarse];
  const int nbig = hc >= 0 ? hc : tb.n[1] + tb.n[2];
  auto bigAt = [&](int i) { return hc >= 0 ? P.hiz_surv[coarse * kHizCap + i] : tb.at(P, tb.n[0] + i); };
  if (lp == 0) {
    s_bound = P.hiz_bound[coarse];
    s_bound2 = MX ? 0xffffffffu : 0u;
    s_n = 0;
  }
  __syncthreads();
  const bool hiz = !P.hiz_off;
  const uint32_t coarse_bound = P.hiz_bound[coarse];
  /* the tile's own depths bound it too */
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    const int pl = lp + q * NT;
    if (hiz && tx0 + pl % TILE < P.fb_w && ty0 + pl / TILE < P.fb_h) hizFromDepth<S>(dep[q], MX, &s_bound2);
  }
  __syncthreads();
  if (lp == 0 && hiz) s_bound = MX ? max(s_bound, s_bound2) : min(s_bound, s_bound2);
  __syncthreads();

  /* One staged triangle per block thread: a big triangle of phase 1a's current
   * chunk, later one of phase 1b's batches. Phase 1a walks its triangles one at
   * a time; reading them here instead of from the setup record and then its
   * vertices saves each step two dependent global loads. */
  __shared__ float4 s_pr[WAVES][64][6];
  float4 *const slot = s_pr[lp >> 6][lp & 63];
  auto stageBig = [&](const PrimSetup &ps, const float4 &f0, const float4 &f1, const float4 &f2, int id) {
    const Tri T = makeTri(f0, f1, f2);
    slot[0] = make_float4(ps.min_x, ps.min_y, ps.max_x, ps.max_y);
    slot[1] = make_float4(__int_as_float(ps.bcx), __int_as_float(ps.bcy), __int_as_float(id), ps.zlo);
    slot[2] = make_float4(ps.zhi, 0.f, T.v0x, T.v0y);
    slot[3] = make_float4(T.a0, T.a1, T.b0, T.b1);
    slot[4] = make_float4(T.uz, T.z[0], T.z[1], T.z[2]);
    slot[5] = make_float4(__int_as_float(ps.qx0), __int_as_float(ps.qx1), __int_as_float(ps.qy0),
                          __int_as_float(ps.qy1));
  };
  auto bigKey = [&](float zlo, float zhi) { return depthKey(gclamp(MX ? zhi : zlo, 0.f, 1.f)); };
  auto touches = [&](const PrimSetup &ps) { return ps.qx0 <= X1 && ps.qx1 >= tx0 && ps.qy0 <= Y1 && ps.qy1 >= ty0; };

  /* big triangles that cover the tile tighten the bound; the first chunk's
   * triangles that touch the tile are staged (id -1 marks the others) */
  for (int i = lp; i < (hiz ? nbig : 0); i += NT) {
    const int id = bigAt(i);
    const PrimSetup ps = P.prims[id];
    if (!touches(ps)) { /* nor can it cover the tile */
      if (i < NT) slot[1].z = __int_as_float(-1);
      continue;
    }
    const float4 f0 = P.frag[ps.v[0]], f1 = P.frag[ps.v[1]], f2 = P.frag[ps.v[2]];
    float zmin, zmax;
    if (coversRect<S>(ps, f0, f1, f2, tx0, ty0, X1, Y1, zmin, zmax)) {
      if (MX)
        atomicMax(&s_bound, depthKey(gclamp(zmin, 0.f, 1.f)));
      else
        atomicMin(&s_bound, depthKey(gclamp(zmax, 0.f, 1.f)));
    }
    if (i < NT) stageBig(ps, f0, f1, f2, id);
  }
  __syncthreads();
  /* the nearest clipped polygon that covers the tile (no single fan of it may):
   * its fans go first, and the depths they leave bound what follows */
  __shared__ uint32_t s_pz;
  __shared__ int s_pt;
  /* A polygon covering the whole coarse bin already bounds its tiles through
   * hiz_bound; rasterizing it first in every tile costs more than the tighter
   * per-tile bound saves (4K stress frame: raster 1.11 -> 0.84 ms). Tiles of a
   * bin without one, whose big list overflowed, look for their own. */
  const int cseed = P.hiz_seed[coarse];
  int seed = -1; /* block-uniform */
  if (hiz && cseed < 0 && hc < 0 && nbig > 0) {
    if (lp == 0) {
      s_pz = MX ? 0u : 0xffffffffu;
      s_pt = -1;
    }
    __syncthreads();
    uint32_t my_pz = MX ? 0u : 0xffffffffu;
    int my_pt = -1;
    for (int i = lp; i < nbig; i += NT) {
      const int id = bigAt(i);
      const PrimSetup &ps = P.prims[id];
      if (ps.src != id) continue; /* each polygon once, through its first fan */
      float zmin, zmax;
      if (polyCoversRect<S>(P, id, tx0, ty0, X1, Y1, zmin, zmax)) {
        const uint32_t k = MX ? depthKey(gclamp(zmin, 0.f, 1.f)) : depthKey(gclamp(zmax, 0.f, 1.f));
        if (MX ? k > my_pz : k < my_pz) {
          my_pz = k;
          my_pt = id;
        }
      }
    }
    if (my_pt >= 0) {
      if (MX)
        atomicMax(&s_pz, my_pz);
      else
        atomicMin(&s_pz, my_pz);
    }
    __syncthreads();
    if (my_pt >= 0 && my_pz == s_pz) s_pt = my_pt; /* any one of equals */
    __syncthreads();
    seed = s_pt;
  }
  if (seed >= 0) {
    const int nf = P.clip_poly[(size_t)seed * (kMaxPoly + 1) + kMaxPoly] - 2;
    for (int k = 0; k < nf; k++) {
      const int id = k == 0 ? seed : P.fan_first[seed] + k - 1;
      const PrimSetup ps = P.prims[id];
      if (!(ps.flags & kPrimLive)) continue;
      const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
#pragma unroll
      for (int q = 0; q < PPL; q++) {
        const int pl = lp + q * NT, px = tx0 + pl % TILE, py = ty0 + pl / TILE;
        if (px < P.fb_w && py < P.fb_h && px >= ps.qx0 && px <= ps.qx1 && py >= ps.qy0 && py <= ps.qy1)
          visPixel<S, MODE>(ps, T, id, px, py, &keys[pl * S]);
      }
    }
    /* the deepest (LESS) depth any sample holds now bounds the tile */
    uint32_t lim = MX ? 0xffffffffu : 0u;
#pragma unroll
    for (int q = 0; q < PPL; q++)
#pragma unroll
      for (int s = 0; s < S; s++) {
        const uint32_t k = (uint32_t)(keys[(lp + q * NT) * S + s] >> 32);
        lim = MX ? min(lim, k) : max(lim, k);
      }
    if (lp == 0) s_pz = MX ? 0xffffffffu : 0u;
    __syncthreads();
    if (MX)
      atomicMin(&s_pz, lim); /* GREATER: the tile's shallowest sample */
    else
      atomicMax(&s_pz, lim); /* LESS: the tile's deepest sample */
    __syncthreads();
    if (lp == 0) s_bound = MX ? max(s_bound, s_pz) : min(s_bound, s_pz);
    __syncthreads();
  }
  const uint32_t bound = s_bound;
  unsigned long long nculled = 0;

  /* phase 1a: surviving big triangles one at a time, lanes over pixels; a
   * wavefront skips a triangle no pixel of it can take at its current depths */
  auto laneLimit = [&]() {
    uint32_t mine = MX ? 0xffffffffu : 0u;
#pragma unroll
    for (int q = 0; q < PPL; q++)
#pragma unroll
      for (int s = 0; s < S; s++) {
        const uint32_t k = (uint32_t)(keys[(lp + q * NT) * S + s] >> 32);
        mine = MX ? min(mine, k) : max(mine, k);
      }
    return mine;
  };
  uint32_t mine = laneLimit();
  uint32_t wave_lim = MX ? waveMinU(mine) : waveMaxU(mine);
  for (int c0 = 0; c0 < nbig; c0 += NT) {
    if (c0 + lp < nbig) {
      bool touch, take;
      uint32_t zk;
      if (hiz && c0 == 0) { /* staged above */
        const float4 d1 = slot[1];
        touch = __float_as_int(d1.z) >= 0;
        zk = touch ? bigKey(d1.w, slot[2].x) : 0u;
      } else {
        const int id = bigAt(c0 + lp);
        const PrimSetup ps = P.prims[id];
        touch = touches(ps);
        zk = bigKey(ps.zlo, ps.zhi);
        if (touch && (MX ? zk >= bound : zk <= bound)) stageBig(ps, P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]], id);
      }
      take = touch && (MX ? zk >= bound : zk <= bound); /* zCulled's test */
      if (take) {
        const int k = atomicAdd(&s_n, 1);
        big[k] = lp;
        big_z[k] = zk;
      } else if (touch) {
        nculled++;
      }
    }
    __syncthreads();
    const int nb = s_n;
    for (int j = 0; j < nb; j++) {
      const uint32_t zk = big_z[j];
      if (hiz && (MX ? zk < wave_lim : zk > wave_lim)) continue; /* wave-uniform */
      const float4 *d = s_pr[big[j] >> 6][big[j] & 63];
      const float4 d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4], d5 = d[5];
      PrimSetup ps;
      ps.min_x = d0.x, ps.min_y = d0.y, ps.max_x = d0.z, ps.max_y = d0.w;
      ps.bcx = __float_as_int(d1.x), ps.bcy = __float_as_int(d1.y);
      Tri T;
      T.v0x = d2.z, T.v0y = d2.w, T.a0 = d3.x, T.a1 = d3.y, T.b0 = d3.z, T.b1 = d3.w;
      T.uz = d4.x, T.z[0] = d4.y, T.z[1] = d4.z, T.z[2] = d4.w;
      const int id = __float_as_int(d1.z);
      const int qx0 = __float_as_int(d5.x), qx1 = __float_as_int(d5.y);
      const int qy0 = __float_as_int(d5.z), qy1 = __float_as_int(d5.w);
#pragma unroll
      for (int q = 0; q < PPL; q++) {
        const int pl = lp + q * NT, px = tx0 + pl % TILE, py = ty0 + pl / TILE;
        if (px < P.fb_w && py < P.fb_h && px >= qx0 && px <= qx1 && py >= qy0 && py <= qy1)
          visPixel<S, MODE>(ps, T, id, px, py, &keys[pl * S]);
      }
      mine = laneLimit();
      wave_lim = MX ? waveMinU(mine) : waveMaxU(mine);
    }
    __syncthreads();
    if (lp == 0) s_n = 0;
    __syncthreads();
  }
  /* the depths after the big triangles bound the rest */
  auto tileLimit = [&]() {
    __syncthreads();
    if (lp == 0) s_bound2 = MX ? 0xffffffffu : 0u;
    __syncthreads();
    if (MX)
      atomicMin(&s_bound2, mine);
    else
      atomicMax(&s_bound2, mine);
    __syncthreads();
  };
  tileLimit();
  uint32_t bound2 = MX ? max(bound, s_bound2) : min(bound, s_bound2);
  /* The coarse bound may rest on the bin's nearest covering polygon
   * (hizCoarseBody). It holds for this tile when every sample now holds a depth
   * at least as near as that bound: a triangle it ruled out is then farther
   * than every sample. Where a sample is farther (a gap between the polygon's
   * fans, a sliver fan culled as back-facing, rounding), every big triangle of
   * the bins that the tile's actual depths do not rule out is rasterized too;
   * those already rasterized leave their keys unchanged. */
  if (hiz && P.hiz_poly[coarse] && (MX ? s_bound2 < coarse_bound : s_bound2 > coarse_bound)) {
    const uint32_t actual = s_bound2;
    const int nfull = tb.n[1] + tb.n[2];
    for (int i = 0; i < nfull; i++) {
      const int id = tb.at(P, tb.n[0] + i);
      const PrimSetup ps = P.prims[id];
      if (ps.qx0 > X1 || ps.qx1 < tx0 || ps.qy0 > Y1 || ps.qy1 < ty0 || zCulled(ps, actual, MX)) continue;
      const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
#pragma unroll
      for (int q = 0; q < PPL; q++) {
        const int pl = lp + q * NT, px = tx0 + pl % TILE, py = ty0 + pl / TILE;
        if (px < P.fb_w && py < P.fb_h && px >= ps.qx0 && px <= ps.qx1 && py >= ps.qy0 && py <= ps.qy1)
          visPixel<S, MODE>(ps, T, id, px, py, &keys[pl * S]);
      }
    }
    mine = laneLimit();
    tileLimit();
    bound2 = s_bound2;
  }
  if (!hiz) bound2 = MX ? 0u : 0xffffffffu;

  /* phase 1b: the tile's own triangles. A wavefront takes 64 at a time, stages
   * what visPixel needs of each in LDS and spreads their (triangle, pixel) pairs
   * evenly over its lanes, 64 pairs per round, instead of each lane walking its
   * own triangle's pixels (the keys' atomics make the order irrelevant). Pairs
   * that the division-free test does not rule out are queued and tested exactly
   * 64 at a time: most pixels of a small triangle's quad range are uncovered. */
  {
    __shared__ int s_start[WAVES][64];
    __shared__ unsigned long long s_marks[WAVES];
    __shared__ uint32_t s_q[WAVES][128]; /* queued pairs: triangle | x << 8 | y << 16, tile-relative */
    const int wv = lp >> 6, lane = lp & 63;
    /* the exact test of queued pairs [0, cnt) */
    auto exactPairs = [&](int cnt) {
      if (lane < cnt) {
        const uint32_t e = s_q[wv][lane];
        const int x = tx0 + (int)((e >> 8) & 255u), y = ty0 + (int)(e >> 16);
        const float4 *d = s_pr[wv][e & 255u];
        const float4 d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4];
        PrimSetup ps;
        ps.min_x = d0.x, ps.min_y = d0.y, ps.max_x = d0.z, ps.max_y = d0.w;
        ps.bcx = __float_as_int(d1.x), ps.bcy = __float_as_int(d1.y);
        Tri T;
        T.v0x = d2.z, T.v0y = d2.w, T.a0 = d3.x, T.a1 = d3.y, T.b0 = d3.z, T.b1 = d3.w;
        T.uz = d4.x, T.z[0] = d4.y, T.z[1] = d4.z, T.z[2] = d4.w;
        visPixel<S, MODE>(ps, T, __float_as_int(d1.z), x, y, &keys[((y - ty0) * TILE + (x - tx0)) * S]);
      }
    };
    for (int b0 = wv * 64; b0 < tb.n[0]; b0 += NT) {
      int n = 0;
      if (b0 + lane < tb.n[0]) {
        const int id = P.bin_items[tb.s[0] + b0 + lane];
        const PrimSetup ps = P.prims[id];
        if (zCulled(ps, bound2, MX)) {
          nculled++;
        } else {
          const int x0 = max(ps.qx0, tx0), x1 = min(ps.qx1, X1);
          const int y0 = max(ps.qy0, ty0), y1 = min(ps.qy1, Y1);
          if (x0 <= x1 && y0 <= y1) {
            n = (x1 - x0 + 1) * (y1 - y0 + 1);
            const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
            float4 *d = s_pr[wv][lane];
            d[0] = make_float4(ps.min_x, ps.min_y, ps.max_x, ps.max_y);
            d[1] = make_float4(__int_as_float(ps.bcx), __int_as_float(ps.bcy), __int_as_float(id),
                               __int_as_float(x0));
            d[2] = make_float4(__int_as_float(y0), __int_as_float(x1 - x0 + 1), T.v0x, T.v0y);
            d[3] = make_float4(T.a0, T.a1, T.b0, T.b1);
            d[4] = make_float4(T.uz, T.z[0], T.z[1], T.z[2]);
          }
        }
      }
      const int inc = waveInclusiveScan(n);
      const int off = inc - n, total = __builtin_amdgcn_readlane(inc, 63);
      s_pr[wv][lane][5].x = __int_as_float(off);
      int carry = 0, qn = 0;
      for (int r0 = 0; r0 < total; r0 += 64) {
        if (lane == 0) s_marks[wv] = 0ull;
        waveSync();
        if (n > 0 && off >= r0 && off < r0 + 64) {
          s_start[wv][off - r0] = lane;
          atomicOr(&s_marks[wv], 1ull << (off - r0));
        }
        waveSync();
        /* the owner of pair r0 + lane: the last triangle starting at or before it */
        const unsigned long long m = s_marks[wv] & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const int owner = m ? s_start[wv][63 - __clzll(m)] : carry;
        carry = __builtin_amdgcn_readlane(owner, 63);
        const int t = r0 + lane;
        bool keep = false;
        uint32_t e = 0;
        if (t < total) {
          const float4 *d = s_pr[wv][owner];
          const float4 d1 = d[1], d2 = d[2], d3 = d[3];
          Tri T;
          T.v0x = d2.z, T.v0y = d2.w, T.a0 = d3.x, T.a1 = d3.y, T.b0 = d3.z, T.b1 = d3.w;
          T.uz = d[4].x;
          const int w = __float_as_int(d2.y), r = t - __float_as_int(d[5].x);
          const int ry = r / w, rx = r - ry * w;
          const int x = __float_as_int(d1.w) + rx, y = __float_as_int(d2.x) + ry;
          keep = !pixelSurelyOut<S>(T, x, y);
          e = (uint32_t)owner | ((uint32_t)(x - tx0) << 8) | ((uint32_t)(y - ty0) << 16);
        }
        const unsigned long long kb = __ballot(keep);
        if (keep) s_q[wv][qn + __popcll(kb & ((1ull << lane) - 1ull))] = e;
        qn += __popcll(kb);
        waveSync();
        if (qn >= 64) {
          exactPairs(64);
          waveSync();
          if (lane < qn - 64) s_q[wv][lane] = s_q[wv][64 + lane];
          qn -= 64;
        }
        waveSync(); /* marks, starts and the queue are read before the next round changes them */
      }
      if (qn > 0) exactPairs(qn);
      waveSync(); /* staged triangles are read before the next batch overwrites them */
    }
  }
  __syncthreads();

  /* phase 2: each pixel's distinct winners write their depth here; colour is
   * queued, one job per (pixel, winner), for k_shade */
  uint32_t win[PPL][S];
  int nj = 0;
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    const int pl = lp + q * NT, px = tx0 + pl % TILE, py = ty0 + pl / TILE;
    const bool in = px < P.fb_w && py < P.fb_h;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t low = (uint32_t)keys[pl * S + s];
      const uint32_t ord = (MODE == VIS_LESS || MODE == VIS_GEQUAL) ? low : ~low;
      win[q][s] = (in && ord) ? ord - 1u : 0xffffffffu;
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
      bool seen = win[q][s] == 0xffffffffu;
#pragma unroll
      for (int t = 0; t < s; t++) seen = seen || win[q][t] == win[q][s];
      nj += seen ? 0 : 1;
    }
  }
  int base = 0;
  if (hasC) {
    if (lp == 0) s_n = 0;
    __syncthreads();
    const int off = nj ? atomicAdd(&s_n, nj) : 0;
    __syncthreads();
    if (lp == 0 && s_n) s_base = atomicAdd(&P.dcount[2], s_n);
    __syncthreads();
    base = s_base + off;
  }
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    const int pl = lp + q * NT, px = tx0 + pl % TILE, py = ty0 + pl / TILE;
    if (px >= P.fb_w || py >= P.fb_h) continue;
    const size_t pix = (size_t)py * P.fb_w + px;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t w = win[q][s];
      if (w == 0xffffffffu) continue;
      bool seen = false;
#pragma unroll
      for (int t = 0; t < s; t++) seen = seen || win[q][t] == w;
      if (seen) continue;
      const int id = (int)(w >> 2);
      /* A winning key holds the depth the triangle's sample took; it returns
       * that depth exactly unless it was a zero (depthKey folds -0 into +0) or
       * a NaN, which take the triangle's interpolation again. */
      uint32_t m = 0;
      bool again = false;
#pragma unroll
      for (int t = s; t < S; t++) {
        if (win[q][t] != w) continue;
        m |= 1u << t;
        const float z = keyDepth((uint32_t)(keys[pl * S + t] >> 32));
        if (z > 0.f && z <= 1.f)
          dep[q][t] = z;
        else
          again = true;
      }
      if (again) {
        const PrimSetup ps = P.prims[id];
        const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
#pragma unroll
        for (int t = s; t < S; t++) {
          if (win[q][t] != w) continue;
          float4 bc;
          if (S == 4)
            bary(T, lx[t] + (float)px, ly[t] + (float)py, bc);
          else
            bary(T, (float)px + 0.5f, (float)py + 0.5f, bc);
          dep[q][t] = gclamp(dpps3(bc, T.z), 0.f, 1.f);
        }
      }
      if (hasC) P.jobs[base++] = make_uint2((uint32_t)pix | (m << 28), w);
    }
#pragma unroll
    for (int s = 0; s < S; s++) P.depth[pix * S + s] = dep[q][s];
  }
  waveAddCounter(P, 6, nculled);
}

/* Four wavefronts per SIMD (at most 128 registers, a few spilled) hide the
 * latency of the tile's loads and LDS atomics better than three (measured: C2
 * rasters -9 to -17 %, C3 -17 %). */
template <int S, int MODE, int TILE, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_raster_vis(DrawParams P) {
  visTile<S, MODE, TILE, WAVES>(P, blockIdx.x); /* tiles stay dealt round-robin: their costs are uneven */
}

/* Colour of the fragments the two-phase raster queued: one lane per (pixel,
 * winning triangle), every sample the triangle won gets its shader output. */
template <int S, int SHADING>
__device__ __forceinline__ void shadeJob(const DrawParams &P, int j, int need_deriv, unsigned long long &nfs,
                                         unsigned long long &nwr) {
  const float lx[4] = {0.375f, 0.875f, 0.125f, 0.625f};
  const float ly[4] = {0.875f, 0.625f, 0.375f, 0.125f};
  {
    const uint2 jb = P.jobs[j];
    const uint32_t pix = jb.x & 0x0fffffffu, m = jb.x >> 28, w = jb.y;
    const int px = (int)(pix % (uint32_t)P.fb_w), py = (int)(pix / (uint32_t)P.fb_w);
    const int id = (int)(w >> 2), inst = (int)(w & 3u);
    const PrimSetup ps = P.prims[id];
    const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
    const float fx = (float)px, fy = (float)py;
    /* the shading sample: the centre if inside, else the first inside sample (RendererSoft.cpp:776-802) */
    float4 bs;
    if (S == 4) {
      if (!bary(T, fx + 0.5f, fy + 0.5f, bs)) {
#pragma unroll
        for (int t = 3; t >= 0; t--) {
          float4 b;
          if (bary(T, lx[t] + fx, ly[t] + fy, b)) bs = b;
        }
      }
    } else {
      bary(T, fx + 0.5f, fy + 0.5f, bs);
    }
    RGBAf c = shadeAt<S, SHADING, true>(P, ps, T, px, py, bs, inst, need_deriv); /* written as bytes, no blending */
    c = {gclamp(c.r, 0.f, 1.f), gclamp(c.g, 0.f, 1.f), gclamp(c.b, 0.f, 1.f), gclamp(c.a, 0.f, 1.f)};
    const uint32_t packed = pack4(u8t(c.r * 255.f), u8t(c.g * 255.f), u8t(c.b * 255.f), u8t(c.a * 255.f));
    nfs++;
#pragma unroll
    for (int t = 0; t < S; t++)
      if ((m >> t) & 1u) {
        P.color[(size_t)pix * S + t] = packed;
        nwr++;
      }
  }
}

/* Adds the lanes' fragment-shader and sample-write counts to the draw counters
 * (one atomic each per wavefront; every lane of the wavefront calls it). */
__device__ __forceinline__ void addShadeCounts(const DrawParams &P, unsigned long long nfs, unsigned long long nwr) {
  waveAddCounter(P, 0, nfs);
  waveAddCounter(P, 1, nwr);
}

template <int S, int SHADING>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_shade(DrawParams P, int need_deriv) {
  const int n = P.dcount[2];
  unsigned long long nfs = 0, nwr = 0;
  const XcdShare x = xcdShare(n);
  for (int j = x.lo + x.idx * blockDim.x + threadIdx.x; j < x.hi; j += x.cnt * blockDim.x)
    shadeJob<S, SHADING>(P, j, need_deriv, nfs, nwr);
  addShadeCounts(P, nfs, nwr);
}

/* multiSampleResolve (RendererSoft.cpp:880-912) of a whole 4x colour target */
__device__ __forceinline__ void resolveBody(const DrawParams &P, int blk, int nblocks) {
  const size_t n = (size_t)P.fb_w * P.fb_h;
  for (size_t p = (size_t)blk * blockDim.x + threadIdx.x; p < n; p += (size_t)nblocks * blockDim.x) {
    const uint4 c = ((const uint4 *)P.color)[p];
    const uint32_t cs[4] = {c.x, c.y, c.z, c.w};
    P.color_resolved[p] = resolve4(cs);
  }
}

static __global__ __launch_bounds__(256) void k_resolve(DrawParams P) { resolveBody(P, blockIdx.x, gridDim.x); }

/* --------------------------------------------------------------------------- ordered raster
 *
 * One workgroup per 16x16 tile, one wavefront per 8x8 quarter, one lane per
 * pixel. The tile's triangles are sorted by primitive id in LDS (in windows of
 * kOrdWin ids when there are more). Walking them in order, each lane finds the
 * fragments of its pixel (one per quad of the triangle holding it), tests
 * coverage and early-Z, and appends them to its wavefront's queue; when the
 * queue fills, the wavefront shades the queued fragments with all lanes busy,
 * then every lane applies its own fragments' per-sample operations in order
 * (late depth test, depth write, blending). Early-Z against a depth that
 * fragments still in the queue may lower is conservative for the monotonic
 * tests: such a fragment also fails the late test. For NOTEQUAL with depth
 * writes the queue is drained after every triangle. */
constexpr int kOrdWin = 2048;
constexpr int kOrdJobsW = 256; /* queue entries per wavefront (64 lanes x 4 quads at most per triangle) */
constexpr int kOrdStage = 256; /* triangles of a window whose walk data is staged in LDS */

template <int S, int SHADING>
__device__ __forceinline__ void ordTile(const DrawParams &P, int tile, int need_deriv, int flush_each) {
  __shared__ uint32_t win[kOrdWin];
  __shared__ uint32_t job_id[4 * kOrdJobsW];
  __shared__ uint32_t job_meta[4 * kOrdJobsW];
  __shared__ short job_next[4 * kOrdJobsW];
  __shared__ float4 job_col[4 * kOrdJobsW];
  __shared__ float job_z[4 * kOrdJobsW * S];
  __shared__ float4 s_stage[kOrdStage][5]; /* quad range, box, block counts and edge/depth setup per triangle */
  __shared__ uint32_t s_bound;
  __shared__ int s_n, s_cnt;
  const int tx0 = (tile % P.tiles_x) * kTile, ty0 = (tile / P.tiles_x) * kTile;
  const int X1 = min(tx0 + kTile, P.fb_w) - 1, Y1 = min(ty0 + kTile, P.fb_h) - 1;
  const int lp = threadIdx.x, wave = lp >> 6, lane = lp & 63;
  const int wx0 = tx0 + (wave & 1) * kSub, wy0 = ty0 + (wave >> 1) * kSub;
  const int px = wx0 + (lane & 7), py = wy0 + (lane >> 3);
  const int pl = (py - ty0) * kTile + (px - tx0);
  const bool inFb = px < P.fb_w && py < P.fb_h;
  const size_t pix = (size_t)py * P.fb_w + px;
  const bool hasC = P.color != nullptr, hasD = P.depth != nullptr;
  const TileBins tb = tileBins(P, tile);
  const int n = tb.total();
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
    if (hasD)
      for (int s = 0; s < S; s++) dep[s] = P.depth[pix * S + s];
    if (hasC)
      for (int s = 0; s < S; s++) col[s] = P.color[pix * S + s];
  }
  if (n == 0) {
    if (inFb && hasC && S == 4) P.color_resolved[pix] = resolve4(col);
    return;
  }
  const int dfunc = P.st.depthFunc;
  const bool dtest = P.st.depthTest && hasD;
  const bool hiz = dtest && (dfunc == 1 || dfunc == 3 || dfunc == 4 || dfunc == 6);
  const bool mx = dfunc == 4 || dfunc == 6;
  if (lp == 0) {
    s_bound = mx ? 0xffffffffu : 0u;
    s_n = 0;
  }
  __syncthreads();
  if (hiz && inFb) hizFromDepth<S>(dep, mx, &s_bound);
  __syncthreads();
  const uint32_t bound = s_bound;

  const int maxid = P.dcount[1] - 1;
  const int jb = wave * kOrdJobsW;
  unsigned long long nculled = 0;
  int head = -1, tail = -1, wcnt = 0;
  unsigned long long nfs = 0, nwr = 0;
  int lo = 0;
  while (lo <= maxid) {
    int hi = maxid;
    if (n > kOrdWin) {
      /* the largest hi with at most kOrdWin entries in [lo, hi] */
      int L = lo, R = maxid;
      while (L < R) {
        const int mid = L + (R - L + 1) / 2;
        if (lp == 0) s_cnt = 0;
        __syncthreads();
        int cnt = 0;
        for (int i = lp; i < n; i += 256) {
          const int id = tb.at(P, i);
          cnt += (id >= lo && id <= mid) ? 1 : 0;
        }
        if (cnt) atomicAdd(&s_cnt, cnt);
        __syncthreads();
        const int c = s_cnt;
        __syncthreads();
        if (c <= kOrdWin)
          L = mid;
        else
          R = mid - 1;
      }
      hi = L;
    }
    for (int i = lp; i < n; i += 256) {
      const int id = tb.at(P, i);
      if (id < lo || id > hi) continue;
      const PrimSetup &ps = P.prims[id];
      if (i >= tb.n[0] && (ps.qx0 > X1 || ps.qx1 < tx0 || ps.qy0 > Y1 || ps.qy1 < ty0)) continue;
      if (hiz && zCulled(ps, bound, mx)) {
        nculled++;
        continue;
      }
      win[atomicAdd(&s_n, 1)] = (uint32_t)id;
    }
    __syncthreads();
    const int m = s_n;
    sortIds(win, m);
    /* what the walk reads of the window's first triangles, staged once per window
     * instead of two dependent global loads per triangle and wavefront */
    const int ms = min(m, kOrdStage);
    for (int i = lp; i < ms; i += 256) {
      const PrimSetup ps = P.prims[win[i]];
      const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
      s_stage[i][0] = make_float4(__int_as_float(ps.qx0), __int_as_float(ps.qx1), __int_as_float(ps.qy0),
                                  __int_as_float(ps.qy1));
      s_stage[i][1] = make_float4(ps.min_x, ps.max_x, ps.min_y, ps.max_y);
      s_stage[i][2] = make_float4(__int_as_float(ps.bcx), __int_as_float(ps.bcy), T.v0x, T.v0y);
      s_stage[i][3] = make_float4(T.a0, T.a1, T.b0, T.b1);
      s_stage[i][4] = make_float4(T.uz, T.z[0], T.z[1], T.z[2]);
    }
    __syncthreads();

    /* walk the window in order; the queue is drained at one place */
    int k = 0, tot = 0, nI = 0, pend_id = 0;
    uint32_t meta = 0;
    int qn = 0;
    bool have = false;
    float zs[S];
#pragma unroll
    for (int s = 0; s < S; s++) zs[s] = 0.f;
    for (;;) {
      const bool drain = wcnt > 0 && (have ? (wcnt + tot > kOrdJobsW) : (k == m || flush_each));
      if (drain) {
        waveSync();
        for (int j = lane; j < wcnt; j += 64) {
          const int slot = jb + j;
          if (!hasC) break;
          const int id = (int)job_id[slot];
          const uint32_t mt = job_meta[slot];
          const int jp = (int)(mt & 255u), inst = (int)((mt >> 8) & 3u), shade = (int)((mt >> 14) & 7u);
          const int jx = tx0 + (jp & (kTile - 1)), jy = ty0 + jp / kTile;
          const PrimSetup ps = P.prims[id];
          const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
          const float fx = (float)jx, fy = (float)jy;
          float4 b;
          if (S == 4 && shade < 4)
            bary(T, sampleX(shade) + fx, sampleY(shade) + fy, b);
          else
            bary(T, fx + 0.5f, fy + 0.5f, b);
          const RGBAf c = shadeAt<S, SHADING>(P, ps, T, jx, jy, b, inst, need_deriv);
          job_col[slot] = make_float4(gclamp(c.r, 0.f, 1.f), gclamp(c.g, 0.f, 1.f), gclamp(c.b, 0.f, 1.f),
                                      gclamp(c.a, 0.f, 1.f));
          nfs++;
        }
        waveSync();
        /* per-sample operations of this lane's fragments, in order (RendererSoft.cpp:358-407) */
        for (int j = head; j >= 0; j = job_next[j]) {
          const int smask = (int)((job_meta[j] >> 10) & 15u);
          const float4 c = job_col[j];
#pragma unroll
          for (int s = 0; s < S; s++) {
            if (!((smask >> s) & 1)) continue;
            if (dtest) {
              const float z = job_z[j * S + s];
              if (!depthPass(z, dep[s], dfunc)) continue;
              if (P.st.depthMask) dep[s] = z;
            }
            if (!hasC) continue;
            RGBAf o = {c.x, c.y, c.z, c.w};
            if (P.st.blend) o = blendColor(o, unpack01(col[s]), P.st);
            col[s] = pack4(u8t(o.r * 255.f), u8t(o.g * 255.f), u8t(o.b * 255.f), u8t(o.a * 255.f));
            nwr++;
          }
        }
        head = tail = -1;
        wcnt = 0;
        waveSync();
      }
      if (have) {
        /* append this lane's fragments of triangle pend_id */
        const int base = jb + wcnt + qn - nI;
        int q = 0;
        for (int iy = 0; iy < 2; iy++)
          for (int ix = 0; ix < 2; ix++) {
            if (q >= nI || !((meta >> (20 + iy * 2 + ix)) & 1u)) continue;
            const int slot = base + q++;
            job_id[slot] = (uint32_t)pend_id;
            job_meta[slot] = (meta & 0x1ffffu) | ((uint32_t)(iy * 2 + ix) << 8);
            job_next[slot] = -1;
#pragma unroll
            for (int s = 0; s < S; s++) job_z[slot * S + s] = zs[s];
            if (tail >= 0)
              job_next[tail] = (short)slot;
            else
              head = slot;
            tail = slot;
          }
        wcnt += tot;
        have = false;
        k++;
        continue;
      }
      if (k == m) break;
      const int id = (int)win[k];
      PrimSetup ps;
      Tri T;
      if (k < ms) { /* wave-uniform k: broadcast LDS reads */
        const float4 d0 = s_stage[k][0], d1 = s_stage[k][1], d2 = s_stage[k][2], d3 = s_stage[k][3],
                     d4 = s_stage[k][4];
        ps.qx0 = __float_as_int(d0.x), ps.qx1 = __float_as_int(d0.y);
        ps.qy0 = __float_as_int(d0.z), ps.qy1 = __float_as_int(d0.w);
        ps.min_x = d1.x, ps.max_x = d1.y, ps.min_y = d1.z, ps.max_y = d1.w;
        ps.bcx = __float_as_int(d2.x), ps.bcy = __float_as_int(d2.y);
        T.v0x = d2.z, T.v0y = d2.w, T.a0 = d3.x, T.a1 = d3.y, T.b0 = d3.z, T.b1 = d3.w;
        T.uz = d4.x, T.z[0] = d4.y, T.z[1] = d4.z, T.z[2] = d4.w;
      } else {
        ps = P.prims[id];
        T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
      }
      if (ps.qx1 < wx0 || ps.qx0 > wx0 + kSub - 1 || ps.qy1 < wy0 || ps.qy0 > wy0 + kSub - 1) {
        k++;
        continue;
      }
      nI = 0;
      meta = 0;
      if (inFb) {
        int qa, qb;
        const int nqx = axisQuads(ps.min_x, ps.max_x, ps.bcx, px, qa, qb);
        const int nqy = nqx ? axisQuads(ps.min_y, ps.max_y, ps.bcy, py, qa, qb) : 0;
        if (nqx && nqy) {
          const float fx = (float)px, fy = (float)py;
          /* coverage, z and the depth clip (RendererSoft.cpp:776-799) */
          bool sin[5];
          float4 sbc[5];
          float sz[5];
          int shade = 0;
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
#pragma unroll
            for (int s = 0; s < 4; s++) {
              sz[s] = 0.f;
              if (!sin[s]) continue;
              sz[s] = dpps3(sbc[s], T.z);
              if (sz[s] < 0.f || sz[s] > 1.f) sin[s] = false;
            }
          } else {
            sin[0] = bary(T, fx + 0.5f, fy + 0.5f, sbc[0]);
            pin = sin[0];
            sz[0] = pin ? dpps3(sbc[0], T.z) : 0.f;
          }
          /* early-Z (RendererSoft.cpp:853-878): test without writing */
          if (pin && P.st.depthTest) {
            if (S == 4) {
              bool any = false;
#pragma unroll
              for (int s = 0; s < 4; s++) {
                if (!sin[s]) continue;
                if (hasD) sin[s] = depthPass(gclamp(sz[s], 0.f, 1.f), dep[s], dfunc);
                any = any || sin[s];
              }
              pin = any;
            } else {
              pin = hasD ? depthPass(gclamp(sz[0], 0.f, 1.f), dep[0], dfunc) : true;
            }
          }
          if (pin) {
            uint32_t smask = 0;
            if (S == 4) {
#pragma unroll
              for (int s = 0; s < 4; s++) smask |= (sin[s] ? 1u : 0u) << s;
            } else {
              smask = 1u;
            }
#pragma unroll
            for (int s = 0; s < S; s++) zs[s] = gclamp(sz[s], 0.f, 1.f);
            uint32_t qm = 0;
            for (int iy = 0; iy < nqy; iy++)
              for (int ix = 0; ix < nqx; ix++) qm |= 1u << (20 + iy * 2 + ix);
            nI = nqx * nqy;
            meta = (uint32_t)pl | (smask << 10) | ((uint32_t)shade << 14) | qm;
          }
        }
      }
      qn = waveInclusiveScan(nI);
      tot = __shfl(qn, 63, 64);
      if (tot == 0) {
        k++;
        continue;
      }
      pend_id = id;
      have = true;
    }
    __syncthreads();
    if (lp == 0) s_n = 0;
    __syncthreads();
    if (hi >= maxid) break;
    lo = hi + 1;
  }

  if (inFb) {
    if (hasD)
      for (int s = 0; s < S; s++) P.depth[pix * S + s] = dep[s];
    if (hasC) {
      for (int s = 0; s < S; s++) P.color[pix * S + s] = col[s];
      if (S == 4) P.color_resolved[pix] = resolve4(col);
    }
  }
  waveAddCounter(P, 0, nfs);
  waveAddCounter(P, 1, nwr);
  waveAddCounter(P, 6, nculled);
}

/* The ordered raster of every tile, one workgroup per tile. */
template <int S, int SHADING>
__global__ __launch_bounds__(256) void k_raster_ord(DrawParams P, int need_deriv, int flush_each) {
  ordTile<S, SHADING>(P, blockIdx.x, need_deriv, flush_each);
}

/* --------------------------------------------------------------------------- listed ordered raster
 *
 * For a draw that writes no depth, every fragment's depth test sees the depths
 * the draw started with, so which samples a fragment reaches does not depend on
 * the others; only the per-sample writes (RendererSoft.cpp:358-407) depend on
 * their order. k_ord_list walks a tile's triangles in primitive-id order and
 * lists each fragment that covers a sample and passes the depth clip and the
 * depth test, pixel by pixel in API order; k_ord_shade runs the fragment shader
 * of every listed fragment at once over the whole GPU; k_ord_apply makes each
 * pixel's writes in order. A tile whose fragments do not fit is marked, and
 * k_ord_rest gives it the ordered raster instead. */
constexpr int kOrdEmpty = 0, kOrdListed = 1, kOrdOverflow = 2;
constexpr int kListCap = 512;              /* fragments one wavefront's 8x8 quarter can list */
constexpr uint32_t kListVoid = 0xffffffffu; /* a jobs slot reserved by a tile that overflowed */

constexpr int kListStage = 96; /* triangles of a window whose walk data is staged in LDS */

/* One tile by one workgroup, one wavefront per 8x8 quarter. A wavefront takes
 * the window's triangles 64 at a time and spreads their (triangle, pixel) pairs
 * over its lanes, 64 pairs per round; fragments are listed in round and lane
 * order, which is primitive-id order for every pixel. */
template <int S>
__device__ __forceinline__ void listTile(const DrawParams &P, int tile) {
  __shared__ uint32_t win[kOrdWin];
  __shared__ float4 s_stage[kListStage][5];
  __shared__ uint2 s_frag[4][kListCap];
  __shared__ uint32_t s_rank[4][kListCap]; /* rank among its pixel's fragments << 8 | pixel in the tile */
  __shared__ uint32_t s_lq[4][128];        /* queued pairs: triangle | x << 8 | y << 16, tile-relative */
  __shared__ float s_dep[kTile * kTile * S];
  __shared__ int s_cnt[kTile * kTile];
  __shared__ int4 s_run[4][64]; /* per triangle of a wavefront's batch: first pair, x0, y0, width */
  __shared__ int s_start[4][64];
  __shared__ unsigned long long s_marks[4];
  __shared__ uint32_t s_bound;
  __shared__ int s_n, s_c, s_over, s_base, s_wn[4];
  const int tx0 = (tile % P.tiles_x) * kTile, ty0 = (tile / P.tiles_x) * kTile;
  const int X1 = min(tx0 + kTile, P.fb_w) - 1, Y1 = min(ty0 + kTile, P.fb_h) - 1;
  const int lp = threadIdx.x, wave = lp >> 6, lane = lp & 63;
  const int wx0 = tx0 + (wave & 1) * kSub, wy0 = ty0 + (wave >> 1) * kSub;
  const int wx1 = min(wx0 + kSub - 1, X1), wy1 = min(wy0 + kSub - 1, Y1);
  const int px = wx0 + (lane & 7), py = wy0 + (lane >> 3);
  const int pl = (py - ty0) * kTile + (px - tx0);
  const bool inFb = px < P.fb_w && py < P.fb_h;
  const size_t pix = (size_t)py * P.fb_w + px;
  const bool hasD = P.depth != nullptr;
  const TileBins tb = tileBins(P, tile);
  const int n = tb.total();
  if (n == 0) {
    if (lp == 0) P.ord_state[tile] = kOrdEmpty;
    return;
  }
  const float lx[4] = {0.375f, 0.875f, 0.125f, 0.625f};
  const float ly[4] = {0.875f, 0.625f, 0.375f, 0.125f};
  float dep[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    dep[s] = (inFb && hasD) ? P.depth[pix * S + s] : 0.f;
    s_dep[pl * S + s] = dep[s];
  }
  s_cnt[pl] = 0;
  const int dfunc = P.st.depthFunc;
  const bool dtest = P.st.depthTest && hasD;
  const bool hiz = dtest && (dfunc == 1 || dfunc == 3 || dfunc == 4 || dfunc == 6);
  const bool mx = dfunc == 4 || dfunc == 6;
  if (lp == 0) {
    s_bound = mx ? 0xffffffffu : 0u;
    s_n = 0;
    s_over = 0;
  }
  __syncthreads();
  if (hiz && inFb) hizFromDepth<S>(dep, mx, &s_bound);
  __syncthreads();
  const uint32_t bound = s_bound;

  const int maxid = P.dcount[1] - 1;
  unsigned long long nculled = 0;
  int wn = 0;        /* fragments this wavefront has listed */
  bool full = false; /* its list overflowed (wave-uniform) */
  int lo = 0;
  while (lo <= maxid) {
    int hi = maxid;
    if (n > kOrdWin) {
      /* the largest hi with at most kOrdWin entries in [lo, hi] */
      int L = lo, R = maxid;
      while (L < R) {
        const int mid = L + (R - L + 1) / 2;
        if (lp == 0) s_c = 0;
        __syncthreads();
        int cnt = 0;
        for (int i = lp; i < n; i += 256) {
          const int id = tb.at(P, i);
          cnt += (id >= lo && id <= mid) ? 1 : 0;
        }
        if (cnt) atomicAdd(&s_c, cnt);
        __syncthreads();
        const int c = s_c;
        __syncthreads();
        if (c <= kOrdWin)
          L = mid;
        else
          R = mid - 1;
      }
      hi = L;
    }
    for (int i = lp; i < n; i += 256) {
      const int id = tb.at(P, i);
      if (id < lo || id > hi) continue;
      const PrimSetup &ps = P.prims[id];
      if (i >= tb.n[0] && (ps.qx0 > X1 || ps.qx1 < tx0 || ps.qy0 > Y1 || ps.qy1 < ty0)) continue;
      if (hiz && zCulled(ps, bound, mx)) {
        nculled++;
        continue;
      }
      win[atomicAdd(&s_n, 1)] = (uint32_t)id;
    }
    __syncthreads();
    const int m = s_n;
    sortIds(win, m);
    const int ms = min(m, kListStage);
    for (int i = lp; i < ms; i += 256) {
      const PrimSetup ps = P.prims[win[i]];
      const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
      s_stage[i][0] = make_float4(__int_as_float(ps.qx0), __int_as_float(ps.qx1), __int_as_float(ps.qy0),
                                  __int_as_float(ps.qy1));
      s_stage[i][1] = make_float4(ps.min_x, ps.max_x, ps.min_y, ps.max_y);
      s_stage[i][2] = make_float4(__int_as_float(ps.bcx), __int_as_float(ps.bcy), T.v0x, T.v0y);
      s_stage[i][3] = make_float4(T.a0, T.a1, T.b0, T.b1);
      s_stage[i][4] = make_float4(T.uz, T.z[0], T.z[1], T.z[2]);
    }
    __syncthreads();

    for (int c0 = 0; c0 < m && !full; c0 += 64) {
      /* this lane's triangle of the batch: its pixels in the quarter */
      const int k = c0 + lane;
      int np = 0, x0 = 0, y0 = 0, w = 1;
      if (k < m) {
        int qx0, qx1, qy0, qy1;
        if (k < ms) {
          const float4 d0 = s_stage[k][0];
          qx0 = __float_as_int(d0.x), qx1 = __float_as_int(d0.y), qy0 = __float_as_int(d0.z), qy1 = __float_as_int(d0.w);
        } else {
          const PrimSetup &ps = P.prims[win[k]];
          qx0 = ps.qx0, qx1 = ps.qx1, qy0 = ps.qy0, qy1 = ps.qy1;
        }
        x0 = max(qx0, wx0);
        y0 = max(qy0, wy0);
        const int x1 = min(qx1, wx1), y1 = min(qy1, wy1);
        if (x0 <= x1 && y0 <= y1) {
          w = x1 - x0 + 1;
          np = w * (y1 - y0 + 1);
        }
      }
      const int inc = waveInclusiveScan(np);
      const int off = inc - np, total = __builtin_amdgcn_readlane(inc, 63);
      s_run[wave][lane] = make_int4(off, x0, y0, w);
      int carry = 0, qn = 0;
      /* the exact test of queued pairs [0, cnt); the queue keeps primitive order */
      auto exactPairs = [&](int cnt) {
        int nI = 0, id = 0, x = 0, y = 0, owner = 0;
        uint32_t word = 0, qm = 0;
        if (lane < cnt) {
          const uint32_t e = s_lq[wave][lane];
          owner = (int)(e & 255u);
          x = tx0 + (int)((e >> 8) & 255u);
          y = ty0 + (int)(e >> 16);
          const int kk = c0 + owner;
          id = (int)win[kk];
          float mnx, mxx, mny, mxy;
          int bcx, bcy;
          Tri T;
          if (kk < ms) {
            const float4 d1 = s_stage[kk][1], d2 = s_stage[kk][2], d3 = s_stage[kk][3], d4 = s_stage[kk][4];
            mnx = d1.x, mxx = d1.y, mny = d1.z, mxy = d1.w;
            bcx = __float_as_int(d2.x), bcy = __float_as_int(d2.y);
            T.v0x = d2.z, T.v0y = d2.w, T.a0 = d3.x, T.a1 = d3.y, T.b0 = d3.z, T.b1 = d3.w;
            T.uz = d4.x, T.z[0] = d4.y, T.z[1] = d4.z, T.z[2] = d4.w;
          } else {
            const PrimSetup ps = P.prims[id];
            mnx = ps.min_x, mxx = ps.max_x, mny = ps.min_y, mxy = ps.max_y;
            bcx = ps.bcx, bcy = ps.bcy;
            T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
          }
          int qa, qb;
          const int nqx = axisQuads(mnx, mxx, bcx, x, qa, qb);
          const int nqy = nqx ? axisQuads(mny, mxy, bcy, y, qa, qb) : 0;
          if (nqx && nqy) {
            const float fx = (float)x, fy = (float)y;
            const float *dp = &s_dep[((y - ty0) * kTile + (x - tx0)) * S];
            /* coverage, z and the depth clip (RendererSoft.cpp:776-799), then
             * the depth test, which nothing in this draw can change */
            bool sin[4];
            float sz[4];
            bool pin;
            if (S == 4) {
              pin = false;
#pragma unroll
              for (int s = 0; s < 4; s++) {
                float4 bc;
                sin[s] = bary(T, lx[s] + fx, ly[s] + fy, bc);
                sz[s] = 0.f;
                pin = pin || sin[s];
                if (!sin[s]) continue;
                sz[s] = dpps3(bc, T.z);
                if (sz[s] < 0.f || sz[s] > 1.f) sin[s] = false;
              }
            } else {
              float4 bc;
              sin[0] = bary(T, fx + 0.5f, fy + 0.5f, bc);
              pin = sin[0];
              sz[0] = pin ? dpps3(bc, T.z) : 0.f;
            }
            if (pin && P.st.depthTest) {
              if (S == 4) {
                bool any = false;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                  if (!sin[s]) continue;
                  if (hasD) sin[s] = depthPass(gclamp(sz[s], 0.f, 1.f), dp[s], dfunc);
                  any = any || sin[s];
                }
                pin = any;
              } else {
                pin = hasD ? depthPass(gclamp(sz[0], 0.f, 1.f), dp[0], dfunc) : true;
              }
            }
            if (pin) {
              uint32_t smask = 1u;
              if (S == 4) {
                smask = 0;
#pragma unroll
                for (int s = 0; s < 4; s++) smask |= (sin[s] ? 1u : 0u) << s;
              }
              for (int iy = 0; iy < nqy; iy++)
                for (int ix = 0; ix < nqx; ix++) qm |= 1u << (iy * 2 + ix);
              nI = nqx * nqy;
              word = (uint32_t)((size_t)y * P.fb_w + x) | (smask << 28);
            }
          }
        }
        const int pos = waveInclusiveScan(nI);
        const int tot = __builtin_amdgcn_readlane(pos, 63);
        if (wn + tot > kListCap) {
          full = true;
          return;
        }
        /* ranks within each pixel, triangle by triangle: a triangle's pairs
         * are distinct pixels, and triangles come in primitive order */
        const int lpx = (y - ty0) * kTile + (x - tx0);
        int rank = 0;
        const int o0 = __builtin_amdgcn_readlane(owner, 0);
        const int o1 = __builtin_amdgcn_readlane(owner, cnt - 1);
        for (int o = o0; o <= o1; o++) {
          if (owner == o && nI) {
            rank = s_cnt[lpx];
            s_cnt[lpx] = rank + nI;
          }
          waveSync();
        }
        if (nI) {
          int q = wn + pos - nI;
          for (int f = 0; f < 4; f++)
            if ((qm >> f) & 1u) {
              s_frag[wave][q] = make_uint2(word, ((uint32_t)id << 2) | (uint32_t)f);
              s_rank[wave][q++] = ((uint32_t)rank++ << 8) | (uint32_t)lpx;
            }
        }
        wn += tot;
      };
      for (int r0 = 0; r0 < total && !full; r0 += 64) {
        if (lane == 0) s_marks[wave] = 0ull;
        waveSync();
        if (np > 0 && off >= r0 && off < r0 + 64) {
          s_start[wave][off - r0] = lane;
          atomicOr(&s_marks[wave], 1ull << (off - r0));
        }
        waveSync();
        /* the owner of pair r0 + lane: the last triangle starting at or before it */
        const unsigned long long mk = s_marks[wave] & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const int owner = mk ? s_start[wave][63 - __clzll(mk)] : carry;
        carry = __builtin_amdgcn_readlane(owner, 63);
        const int t = r0 + lane;
        bool keep = false;
        uint32_t e = 0;
        if (t < total) {
          const int kk = c0 + owner;
          const int4 run = s_run[wave][owner];
          const int r = t - run.x, ry = r / run.w, rx = r - ry * run.w;
          const int x = run.y + rx, y = run.z + ry;
          keep = true;
          if (kk < ms) { /* pairs the division-free test rules out never reach the exact test */
            const float4 d2 = s_stage[kk][2], d3 = s_stage[kk][3];
            Tri T;
            T.v0x = d2.z, T.v0y = d2.w, T.a0 = d3.x, T.a1 = d3.y, T.b0 = d3.z, T.b1 = d3.w;
            T.uz = s_stage[kk][4].x;
            keep = !pixelSurelyOut<S>(T, x, y);
          }
          e = (uint32_t)owner | ((uint32_t)(x - tx0) << 8) | ((uint32_t)(y - ty0) << 16);
        }
        const unsigned long long kb = __ballot(keep);
        if (keep) s_lq[wave][qn + __popcll(kb & ((1ull << lane) - 1ull))] = e;
        qn += __popcll(kb);
        waveSync();
        if (qn >= 64) {
          exactPairs(64);
          waveSync();
          if (lane < qn - 64) s_lq[wave][lane] = s_lq[wave][64 + lane];
          qn -= 64;
        }
        waveSync(); /* marks, starts and the queue are read before the next round changes them */
      }
      if (qn > 0 && !full) exactPairs(qn);
      waveSync(); /* the batch's runs are read before the next batch overwrites them */
    }
    if (full && lane == 0) s_over = 1;
    __syncthreads();
    const bool stop = s_over != 0;
    if (lp == 0) s_n = 0;
    __syncthreads();
    if (stop || hi >= maxid) break;
    lo = hi + 1;
  }

  /* each pixel's fragments, contiguous and in order, at one range of the draw's list */
  const int c = s_cnt[pl];
  const int inc = waveInclusiveScan(c);
  if (lane == 63) s_wn[wave] = inc;
  __syncthreads();
  const int all = s_wn[0] + s_wn[1] + s_wn[2] + s_wn[3];
  if (lp == 0) s_base = (s_over || all == 0) ? 0 : atomicAdd(&P.dcount[3], all);
  __syncthreads();
  const int gb = s_base;
  if (s_over || gb + all > P.ord_cap) {
    if (!s_over) /* the range reserved past the end is never written: mark what lies before it */
      for (int i = gb + lp; i < min(gb + all, P.ord_cap); i += 256) P.jobs[i] = make_uint2(kListVoid, 0u);
    if (lp == 0) {
      P.ord_state[tile] = kOrdOverflow;
      P.ord_list[atomicAdd(&P.dcount[7], 1)] = tile;
    }
    return;
  }
  int base = gb + inc - c;
  for (int v = 0; v < wave; v++) base += s_wn[v];
  s_cnt[pl] = base; /* every pixel of this quarter: where its fragments start */
  waveSync();
  for (int e = lane; e < wn; e += 64) {
    const uint32_t rk = s_rank[wave][e];
    P.jobs[s_cnt[rk & 255u] + (int)(rk >> 8)] = s_frag[wave][e];
  }
  if (inFb) P.ord_span[pix] = make_int2(base, c);
  if (lp == 0) {
    P.ord_state[tile] = kOrdListed;
    if (all) P.ord_list[P.ntiles - 1 - atomicAdd(&P.dcount[8], 1)] = tile;
  }
  waveAddCounter(P, 6, nculled);
}

/* Tiles dealt round-robin to a grid that fits the GPU at once: most tiles of a
 * draw like this are empty, and the busy ones sit together. A workgroup finds
 * which of its next 64 tiles have triangles with one load per tile, in
 * parallel, and walks only those. */
template <int S>
__global__ __launch_bounds__(256) void k_ord_list(DrawParams P) {
  __shared__ int s_tiles[64];
  __shared__ int s_nt;
  for (int t0 = blockIdx.x; t0 < P.ntiles; t0 += 64 * gridDim.x) {
    if (threadIdx.x == 0) s_nt = 0;
    __syncthreads();
    if (threadIdx.x < 64) {
      const int tile = t0 + (int)threadIdx.x * (int)gridDim.x;
      if (tile < P.ntiles) {
        if (tileBins(P, tile).total() > 0)
          s_tiles[atomicAdd(&s_nt, 1)] = tile;
        else
          P.ord_state[tile] = kOrdEmpty;
      }
    }
    __syncthreads();
    const int nt = s_nt;
    for (int i = 0; i < nt; i++) {
      listTile<S>(P, s_tiles[i]);
      __syncthreads();
    }
  }
}

/* The fragment shader of every listed fragment: the centre sample if it is
 * inside, else the first inside sample (RendererSoft.cpp:776-802). */
template <int S, int SHADING>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_ord_shade(DrawParams P, int need_deriv) {
  const int n = min(P.dcount[3], P.ord_cap);
  const XcdShare xs = xcdShare(n);
  for (int j = xs.lo + xs.idx * blockDim.x + threadIdx.x; j < xs.hi; j += xs.cnt * blockDim.x) {
    const uint2 jb = P.jobs[j];
    if (jb.x == kListVoid) continue;
    const uint32_t pix = jb.x & 0x0fffffffu;
    const int x = (int)(pix % (uint32_t)P.fb_w), y = (int)(pix / (uint32_t)P.fb_w);
    const int id = (int)(jb.y >> 2), inst = (int)(jb.y & 3u);
    const PrimSetup ps = P.prims[id];
    const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
    const float fx = (float)x, fy = (float)y;
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
    const RGBAf c = shadeAt<S, SHADING>(P, ps, T, x, y, bs, inst, need_deriv);
    P.ord_col[j] = make_float4(gclamp(c.r, 0.f, 1.f), gclamp(c.g, 0.f, 1.f), gclamp(c.b, 0.f, 1.f),
                               gclamp(c.a, 0.f, 1.f));
  }
}

/* The IBL precompute shaders (irradiance: 252 x 63 terms per fragment;
 * prefilter: 1024 samples) on listed fragments, one wavefront per fragment:
 * the lanes compute 64 terms at a time and every lane then adds them in the
 * shader's order from LDS, so the sum is the sequential one, bit for bit. A
 * face of 32^2 fragments fills 1024 wavefronts instead of 16. */
template <int S>
__global__ __launch_bounds__(256) void k_ord_shade_ibl(DrawParams P, int need_deriv) {
  __shared__ float4 terms[4][64];
  const int n = min(P.dcount[3], P.ord_cap);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  float4 *L = terms[w];
  for (int j = blockIdx.x * (blockDim.x >> 6) + w; j < n; j += nwaves) {
    const uint2 jb = P.jobs[j];
    if (jb.x == kListVoid) continue;
    const uint32_t pix = jb.x & 0x0fffffffu;
    const int x = (int)(pix % (uint32_t)P.fb_w), y = (int)(pix / (uint32_t)P.fb_w);
    const int id = (int)(jb.y >> 2), inst = (int)(jb.y & 3u);
    const PrimSetup ps = P.prims[id];
    const Tri T = makeTri(P.frag[ps.v[0]], P.frag[ps.v[1]], P.frag[ps.v[2]]);
    const float fx = (float)x, fy = (float)y;
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
    const Frag f = makeFrag<S, kShadingEnv>(P, ps, T, x, y, bs, inst, need_deriv);
    RGBAf c;
    if (P.shading == RHS_SHADING_IBL_IRRADIANCE) {
      const IrrSetup st = irrSetup(f);
      const int total = P.ibl->nphi * P.ibl->ntheta;
      f3 irr{0.f, 0.f, 0.f};
      float nr = 0.f;
      for (int base = 0; base < total; base += 64) {
        const int k = base + lane;
        L[lane] = k < total ? irrTerm(P, st, k) : make_float4(0.f, 0.f, 0.f, 0.f);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int m = min(64, total - base);
        for (int q = 0; q < m; q++) {
          irr = irrAdd(irr, L[q]);
          nr += 1.f;
        }
        __builtin_amdgcn_wave_barrier();
      }
      c = irrFinish(irr, nr);
    } else {
      const PrefSetup st = prefSetup(P, f);
      f3 acc{0.f, 0.f, 0.f};
      float tw = 0.f;
      for (uint32_t base = 0; base < 1024u; base += 64) {
        L[lane] = prefSample(P, st, base + (uint32_t)lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int q = 0; q < 64; q++) {
          const float4 t = L[q];
          if (t.w > 0.0f) {
            acc = {fmaf(t.x, t.w, acc.x), fmaf(t.y, t.w, acc.y), fmaf(t.z, t.w, acc.z)};
            tw += t.w;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      c = {acc.x / tw, acc.y / tw, acc.z / tw, 1.f};
    }
    if (lane == 0)
      P.ord_col[j] = make_float4(gclamp(c.r, 0.f, 1.f), gclamp(c.g, 0.f, 1.f), gclamp(c.b, 0.f, 1.f),
                                 gclamp(c.a, 0.f, 1.f));
  }
}

/* Each pixel's listed fragments written to its samples in order, blended when
 * blending is on: one workgroup per tile with fragments, one lane per pixel. */
template <int S>
__global__ __launch_bounds__(256) void k_ord_apply(DrawParams P) {
  const int n = P.dcount[8];
  unsigned long long nfs = 0, nwr = 0;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int tile = P.ord_list[P.ntiles - 1 - i];
    const int x = (tile % P.tiles_x) * kTile + (int)(threadIdx.x & (kTile - 1));
    const int y = (tile / P.tiles_x) * kTile + (int)(threadIdx.x / kTile);
    if (x >= P.fb_w || y >= P.fb_h) continue;
    const size_t p = (size_t)y * P.fb_w + x;
    const int2 sp = P.ord_span[p];
    if (sp.y == 0) continue;
    uint32_t col[S];
    if (S == 4) {
      const uint4 c4 = ((const uint4 *)P.color)[p];
      col[0] = c4.x, col[1] = c4.y, col[2] = c4.z, col[3] = c4.w;
    } else {
      col[0] = P.color[p];
    }
    for (int j = sp.x; j < sp.x + sp.y; j++) {
      const uint32_t m = P.jobs[j].x >> 28;
      const float4 c = P.ord_col[j];
      nfs++;
#pragma unroll
      for (int s = 0; s < S; s++) {
        if (!((m >> s) & 1u)) continue;
        RGBAf o = {c.x, c.y, c.z, c.w};
        if (P.st.blend) o = blendColor(o, unpack01(col[s]), P.st);
        col[s] = pack4(u8t(o.r * 255.f), u8t(o.g * 255.f), u8t(o.b * 255.f), u8t(o.a * 255.f));
        nwr++;
      }
    }
    if (S == 4)
      ((uint4 *)P.color)[p] = make_uint4(col[0], col[1], col[2], col[3]);
    else
      P.color[p] = col[0];
  }
  addShadeCounts(P, nfs, nwr);
}

/* The ordered raster of the tiles whose fragments did not fit the list. */
template <int S, int SHADING>
__global__ __launch_bounds__(256) void k_ord_rest(DrawParams P, int need_deriv) {
  const int n = P.dcount[7];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    ordTile<S, SHADING>(P, P.ord_list[i], need_deriv, 0);
    __syncthreads();
  }
}

/* --------------------------------------------------------------------------- textures */

/* generateMipmaps / sampleBufferBilinear (SamplerSoft.h:89-110, 240-252): level i is a
 * clamp-to-edge bilinear resample of level i-1 at (x*r + r/2, y*r + r/2). */
template <typename T>
static __global__ void k_mipgen(const TexDesc *d, int layer, int level) {
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
