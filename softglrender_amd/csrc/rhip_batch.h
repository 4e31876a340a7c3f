/*
 * rhip_batch.h -- kernels of the batched multi-view path.
 *
 * Every stage of a batch is one launch over all of its views: a block finds
 * its view and its block within that view in a per-stage table of
 * (view << kViewShift | local block) entries written by the host, and runs the
 * same stage body a single draw runs. The raster is persistent: workgroups take
 * (view, tile) items from a queue until it is empty. Shading walks 64-fragment
 * chunks of every view's fragment queue; the wavefront that finishes a view's
 * last chunk stamps the view's completion time.
 */
#pragma once
#include "rhip_kernels.h"

namespace rhip {

constexpr int kViewShift = 20;
constexpr uint32_t kLocalMask = (1u << kViewShift) - 1;
constexpr int kMaxBatchViews = 1024;  /* 3 fill ranges per view must fit the 12-bit range index */
constexpr uint32_t kFillBlock = 4096; /* words one fill block writes */

/* `n` words at `p` set to `v` */
struct FillRange {
  uint32_t *p;
  uint32_t n, v;
};

struct ViewBlock {
  const DrawParams *P;
  int local;
};

/* Table entries are read at a block-uniform index, so they and the view's
 * parameters come in through scalar loads. */
__device__ __forceinline__ ViewBlock viewBlock(const DrawParams *__restrict__ Pv, const uint32_t *__restrict__ tab) {
  const uint32_t e = tab[blockIdx.x];
  return {Pv + (e >> kViewShift), (int)(e & kLocalMask)};
}

/* Clears (bins, attachments, counters) of every view; block 0 stamps the start time. */
static __global__ __launch_bounds__(256) void kb_fill(const FillRange *__restrict__ r, const uint32_t *__restrict__ tab,
                                              unsigned long long *t0) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *t0 = wall_clock64();
  const uint32_t e = tab[blockIdx.x];
  const FillRange R = r[e >> kViewShift];
  const uint32_t lo = (e & kLocalMask) * kFillBlock, hi = min(lo + kFillBlock, R.n);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) R.p[i] = R.v;
}

static __global__ __launch_bounds__(256) void kb_vertex(const DrawParams *__restrict__ Pv, const uint32_t *__restrict__ tab) {
  const ViewBlock vb = viewBlock(Pv, tab);
  vertexBody(*vb.P, vb.local);
}

static __global__ __launch_bounds__(256) void kb_setup(const DrawParams *__restrict__ Pv, const uint32_t *__restrict__ tab) {
  const ViewBlock vb = viewBlock(Pv, tab);
  setupBody(*vb.P, vb.local);
}

static __global__ __launch_bounds__(256) void kb_clip(const DrawParams *__restrict__ Pv, const uint32_t *__restrict__ tab) {
  const ViewBlock vb = viewBlock(Pv, tab);
  clipBody(*vb.P, vb.local);
}

/* one block per view */
static __global__ __launch_bounds__(1024) void kb_scan_fans(const DrawParams *__restrict__ Pv) { scanFansBody(Pv[blockIdx.x]); }

static __global__ __launch_bounds__(256) void kb_clip_fans(const DrawParams *__restrict__ Pv,
                                                    const uint32_t *__restrict__ tab) {
  const ViewBlock vb = viewBlock(Pv, tab);
  clipFansBody<false>(*vb.P, vb.local);
}

/* one block per view */
static __global__ __launch_bounds__(1024) void kb_bin_scan(const DrawParams *__restrict__ Pv) { binScanBody(Pv[blockIdx.x]); }

/* the view's setup blocks also walk the fan triangles clipping appended */
static __global__ __launch_bounds__(256) void kb_bin_emit(const DrawParams *__restrict__ Pv,
                                                   const uint32_t *__restrict__ tab) {
  const ViewBlock vb = viewBlock(Pv, tab);
  binEmitBody(*vb.P, vb.local, vb.P->nblk);
}

template <int S, bool MX>
__global__ __launch_bounds__(256) void kb_hiz_coarse(const DrawParams *__restrict__ Pv,
                                                     const uint32_t *__restrict__ tab) {
  const ViewBlock vb = viewBlock(Pv, tab);
  hizCoarseBody<S, MX>(*vb.P, vb.local);
}

/* Persistent two-phase raster: each workgroup takes items (view << kViewShift
 * | tile) in queue order until the queue is empty, so every wave reaches the
 * exit once `*queue` passes `nitems`. */
template <int S, int MODE, int TILE, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_raster_vis(const DrawParams *__restrict__ Pv,
                                                            const uint32_t *__restrict__ items, int nitems,
                                                            int *queue) {
  __shared__ int s_item;
  for (;;) {
    __syncthreads(); /* the previous item's index has been read by every thread */
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(s_item);
    if (i >= nitems) return;
    const uint32_t e = items[i];
    visTile<S, MODE, TILE, WAVES>(Pv[e >> kViewShift], (int)(e & kLocalMask));
  }
}

/* One block of 1024: chunk_first[v] = first 64-fragment chunk of view v in the
 * shading walk, chunk_first[nv] = all chunks. Views without fragments to shade
 * are complete once the raster is. */
static __global__ __launch_bounds__(1024) void kb_job_chunks(const DrawParams *__restrict__ Pv, int nv, int *chunk_first,
                                                      unsigned long long *view_time) {
  __shared__ int part[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int v0 = 0; v0 < nv; v0 += 1024) {
    const int v = v0 + threadIdx.x;
    int c = 0;
    if (v < nv) {
      const DrawParams &P = Pv[v];
      c = P.color ? (P.dcount[2] + 63) / 64 : 0;
      if (c == 0) view_time[v] = wall_clock64();
    }
    const int inc = blockInclusiveScan<1024>(c, part);
    if (v < nv) chunk_first[v] = carry + inc - c;
    __syncthreads();
    if (threadIdx.x == 1023) carry += inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) chunk_first[nv] = carry;
}

/* Shading of every view's queued fragments, one 64-fragment chunk per
 * wavefront step, chunks in view order. */
template <int S, int SHADING>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_shade(const DrawParams *__restrict__ Pv, const int *__restrict__ chunk_first,
                                                int nv, int *chunks_done, unsigned long long *view_time) {
  const int lane = threadIdx.x & 63;
  const int total = chunk_first[nv];
  const int nw = gridDim.x * (blockDim.x / 64);
  unsigned long long nfs = 0, nwr = 0;
  for (int c = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)); c < total;
       c += nw) {
    int lo = 0, hi = nv - 1; /* the view of chunk c: the last v with chunk_first[v] <= c */
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (chunk_first[mid] <= c)
        lo = mid;
      else
        hi = mid - 1;
    }
    const DrawParams &P = Pv[lo];
    const int first = chunk_first[lo];
    const int j = (c - first) * 64 + lane;
    if (j < P.dcount[2]) shadeJob<S, SHADING>(P, j, P.need_deriv, nfs, nwr);
    if (lane == 0 && atomicAdd(&chunks_done[lo], 1) == chunk_first[lo + 1] - first - 1) view_time[lo] = wall_clock64();
  }
  addShadeCounts(Pv[0], nfs, nwr);
}

/* multiSampleResolve of every multisampled view: one entry per 256 pixels */
static __global__ __launch_bounds__(256) void kb_resolve(const DrawParams *__restrict__ Pv, const uint32_t *__restrict__ tab) {
  const ViewBlock vb = viewBlock(Pv, tab);
  resolveBody(*vb.P, vb.local, (int)((vb.P->fb_w * vb.P->fb_h + 255) / 256));
}

/* One instance of the persistent raster, for the host's registry. */
struct VisKernel {
  int samples, mode, tile, waves;
  void (*launch)(const DrawParams *Pv, const uint32_t *items, int nitems, int *queue, int grid, hipStream_t st);
  const void *fn;
};

/* Instantiates one persistent raster for the registry. */
template <int S, int MODE, int TILE, int WAVES>
void visLaunch(const DrawParams *Pv, const uint32_t *items, int nitems, int *queue, int grid, hipStream_t st) {
  kb_raster_vis<S, MODE, TILE, WAVES><<<grid, 64 * WAVES, 0, st>>>(Pv, items, nitems, queue);
}
#define RHIP_VIS_KERNEL(S, M, T, W) {S, M, T, W, visLaunch<S, M, T, W>, (const void *)kb_raster_vis<S, M, T, W>}
#define RHIP_VIS_MODES(S, T, W)                                                                    \
  RHIP_VIS_KERNEL(S, VIS_LESS, T, W), RHIP_VIS_KERNEL(S, VIS_LEQUAL, T, W),                      \
      RHIP_VIS_KERNEL(S, VIS_GREATER, T, W), RHIP_VIS_KERNEL(S, VIS_GEQUAL, T, W)

/* The instances compiled for one and for four samples (one translation unit
 * each, so they build in parallel); *n gets their count. */
const VisKernel *visKernels1(int *n);
const VisKernel *visKernels4(int *n);

}  // namespace rhip
