/*
 * rhip_internal.h -- POD records shared by librhip's host code and its gfx950
 * kernels. Nothing here is part of the public C-ABI.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rhip_libm.h"
#include "rhip_stream.h"

namespace rhip {

constexpr int kMaxLevels = 16;
constexpr int kMaxLayers = 6;
constexpr int kTile = 16;         /* default screen tile edge in pixels (one workgroup) */
constexpr int kSub = 8;           /* 8x8 sub-tile per wavefront */
/* statistics counters: kCounterSlots copies of the 8 counters, kCounterStride
 * words (128 bytes) apart; a workgroup adds to copy blockIdx % kCounterSlots
 * and the host sums the copies */
constexpr int kCounterSlots = 64, kCounterStride = 16;
constexpr int kVaryStride = 32;   /* varying slots of a vertex shader output (the reference layout) */
constexpr int kMaxClipVerts = 12; /* new vertices one clipped triangle can create */
constexpr int kMaxPoly = 9;       /* polygon vertices after six planes */
/* 64-block chunks of a draw's 256-triangle blocks, whose fan counts are summed together */
__host__ __device__ inline int fanChunks(int nblk) { return (nblk + 63) >> 6; }
constexpr int kMaxFans = kMaxPoly - 3; /* extra triangles one clipped triangle can append */
/* Binning: a primitive whose quads touch at most kBinFine tiles is listed in
 * each of those tiles; a larger one in each kCoarse x kCoarse-tile coarse bin
 * it touches if those are at most kBinCoarse, else in the one global bin. So
 * no primitive is listed more than kMaxBinItems times. */
constexpr int kBinFine = 16;
constexpr int kCoarse = 8;
constexpr int kBinCoarse = 16;
constexpr int kMaxBinItems = 16;

enum Layout { LAYOUT_LINEAR = 0, LAYOUT_TILED = 1, LAYOUT_MORTON = 2 };

/* One mip level of one layer. `inner_w` is the padded row length in texels
 * (whole 4x4 or 32x32 tiles for the tiled layouts), `tiles_x` the tile count per row. */
struct TexLevel {
  uint64_t offset; /* bytes from TexDesc::base */
  int w, h;
  int inner_w, inner_h;
  int tiles_x;
  int pad;
};

/* Device-resident texture descriptor (read with scalar loads by the kernels). */
struct TexDesc {
  uint8_t *base;
  uint8_t *resolved; /* single-sample image of a multisampled colour target */
  int type, format, width, height;
  int layers, levels, layout, msaa;
  int filter_min, filter_mag, wrap, border;
  TexLevel lv[kMaxLayers][kMaxLevels];
};

/* Uniform blocks decoded to floats (Material.h:63-95 layouts). */
struct Uniforms {
  int reverse_z;
  float model[16], mvp[16], inv_t[12], shadow_mvp[16];
  float ambient[3], cam_pos[3], light_pos[3], light_color[3];
  int enable_light, enable_ibl, enable_shadow;
  float point_size, k_specular;
  float base_color[4];
  float aux[4]; /* the material block's first 16 bytes as floats: UniformsIBLPrefilter {srcResolution,
                   roughness} or UniformsQuadFilter {screenSize} for those shaders */
};

/* Shading models drawn by one kernel specialisation that switches on
 * DrawParams::shading (skybox, IBL irradiance and prefilter, FXAA). */
constexpr int kShadingEnv = 100;
__host__ __device__ inline bool isEnvShading(int s) {
  return s == RHS_SHADING_SKYBOX || s == RHS_SHADING_IBL_IRRADIANCE || s == RHS_SHADING_IBL_PREFILTER ||
         s == RHS_SHADING_FXAA;
}

/* Loop tables of the image-based-lighting shaders, from the host libm
 * (rhip_libm.hip): IBLIrradianceSoft.h:84-85 steps phi over [0, 2 PI) and
 * theta over [0, PI / 2) by 0.025 in float; IBLPrefilterSoft.h:106-112 takes
 * cos and sin of 2 PI i / 1024. */
struct IblTables {
  int nphi, ntheta;
  float sphi[256], cphi[256];
  float stheta[64], ctheta[64];
  float hcos[1024], hsin[1024];
};

/* Per-primitive setup record (after clipping / culling). */
struct PrimSetup {
  int v[3];            /* vertex ids */
  int flags;           /* bit0 live, bit1 front facing */
  float min_x, min_y;  /* triangleBoundingBox min - 1 (RendererSoft.cpp:724-725) */
  float max_x, max_y;  /* triangleBoundingBox max */
  int bcx, bcy;        /* 32-px block counts */
  int qx0, qy0, qx1, qy1; /* pixel range touched by the triangle's quads (inclusive), empty if qx1 < qx0 */
  float zlo, zhi;         /* bounds of every depth the triangle's samples can take */
  int src;                /* triangle a clipped polygon came from (its fans share it), -1 if not clipped */
};

/* Varying slots a fragment shader reads, in the reference's per-shader slot layout
 * (BlinnPhongSoft.h:58-68, PbrSoft.h:63-72): 0-1 uv, 4-6 normal, 12-14 camera
 * direction, 16-18 light direction, 20-23 shadow position (Blinn-Phong), then the
 * TBN normal and tangent when a normal map is bound. World position (8-10) only
 * feeds the vertex shader's own outputs. */
__host__ __device__ inline bool varySlotUsed(int shading, uint32_t defines, int enable_shadow, int j) {
  if (shading == RHS_SHADING_VERTEX_COLOR) return j < 4;
  if (shading == RHS_SHADING_FXAA) return j < 2;  /* v_texCoord */
  if (isEnvShading(shading)) return j < 3;        /* v_worldPos */
  if (shading == RHS_SHADING_BASIC) return false;
  const bool nm = (defines & RHS_DEF_NORMAL_MAP) != 0;
  if (j < 2) return true;
  if (j >= 4 && j <= 6) return !nm;
  if (j >= 12 && j <= 14) return true;
  if (j >= 16 && j <= 18) return shading != RHS_SHADING_NEURAL_PBR; /* the learned shader has no point light */
  if (shading == RHS_SHADING_PBR || shading == RHS_SHADING_NEURAL_PBR)
    return nm && ((j >= 20 && j <= 22) || (j >= 24 && j <= 26));
  if (j >= 20 && j <= 23) return enable_shadow != 0;
  return nm && ((j >= 24 && j <= 26) || (j >= 28 && j <= 30));
}

/* Varyings are stored per vertex as the groups of 4 slots that hold a read slot,
 * packed in slot order: bit g is set when slots 4g..4g+3 are stored. */
__host__ __device__ inline uint32_t varyGroups(int shading, uint32_t defines, int enable_shadow) {
  uint32_t g = 0;
  for (int j = 0; j < kVaryStride; j++)
    if (varySlotUsed(shading, defines, enable_shadow, j)) g |= 1u << (j >> 2);
  return g;
}

/* Device-side error words (DrawParams.errors): a kernel that has to drop work
 * sets its word to 1; rhip_wait_idle and rhip_texture_read report and clear them. */
constexpr int kErrClip = 0; /* a clipped polygon exceeded kMaxPoly vertices: the triangle was dropped */
constexpr int kErrSlots = 4;

/* PrimSetup.flags */
constexpr int kPrimLive = 1, kPrimFront = 2, kPrimCoarse = 4, kPrimGlobal = 8;

struct DrawParams {
  Uniforms u;
  const TexDesc *tex[RHS_SLOT_COUNT];
  RhsStates st;
  int shading;
  uint32_t defines;
  int need_deriv; /* a sampler of the draw needs dFdx/dFdy (mipmapped minification) */
  const float *gamma_lut; /* [0, 256): powf(k / 255, 2.2); [256, 512): the PBR output-byte thresholds
                             (host libm, rhip_libm.hip) */
  const LibmTables *libm; /* the host libm's log2f / exp2f / powf on the device (rhip_libm.h) */
  const IblTables *ibl;   /* sin / cos tables of the IBL shaders */
  float base_lin[3];      /* powf(base colour, 2.2) by the host libm (PbrSoft.h:235) */
  uint32_t vary_groups; /* stored groups of 4 varying slots (varyGroups) */
  int vary_stride;      /* floats per vertex in `vary`: 4 per stored group */
  /* viewport (RendererSoft.cpp:92-113) */
  float vp_w, vp_h, o_x, o_y, p_x, p_y;
  float pl_point_size; /* gl->PointSize after the vertex shader: u_pointSize for the base colour shader, else 1 */
  /* framebuffer */
  uint32_t *color;          /* RGBA8, `samples` per pixel */
  uint32_t *color_resolved; /* RGBA8 single sample, when samples == 4 */
  float *depth;             /* f32, `samples` per pixel */
  int fb_w, fb_h;           /* attachment size */
  int samples;              /* 1 or 4 */
  int tiles_x, tiles_y;
  /* geometry */
  const float *verts;
  const float *clip_raw;
  int stride_f; /* floats per vertex */
  int nverts;
  const int *idx;
  int ntris;
  /* vertex stage outputs */
  float4 *clip;
  float4 *frag;
  int *mask;
  float *vary;
  /* clipping: triangle t owns new-vertex slots nverts + t * kMaxClipVerts ... */
  int *clip_poly; /* kMaxPoly vertex ids + count per triangle of a block that clips */
  int *fan_first; /* per clipped triangle: id of its second fan triangle (the first keeps its id) */
  int *blk_clip;  /* per 256-triangle block: triangles that need clipping */
  int *blk_fans;  /* per block: fan triangles its clipped triangles append */
  int *blk_fan_off; /* per block: its fans' offset (batched views) */
  int *chunk_fans;  /* per 64 blocks: their fan triangles (a single draw) */
  int nblk;
  int *dcount; /* device scalars: [0] appended fan triangles, [1] primitive count, [2] queued fragments,
                [3] listed fragments, [7] tiles left to the ordered raster, [8] tiles with listed fragments */
  /* primitives */
  PrimSetup *prims;
  /* bins: ntiles fine bins of tile x tile pixels, then coarse_x * coarse_y coarse bins of
   * coarse_tiles x coarse_tiles tiles, then the global bin */
  int tile, coarse_tiles;
  int ntiles, coarse_x, coarse_y, nbins;
  int *bin_count;  /* [nbins] */
  int *bin_start;  /* [nbins + 1] */
  int *bin_cursor; /* [nbins] */
  int *bin_items;  /* primitive ids, unordered within a bin */
  /* two-phase raster */
  uint2 *jobs;          /* (pixel | sample mask << 28, winner) per visible fragment; count in dcount[2]
                           (the listed ordered raster: dcount[3]) */
  uint32_t *hiz_bound; /* per coarse bin: depth bound from big triangles covering it */
  int *hiz_count;      /* per coarse bin: big triangles that bound leaves, -1 if over kHizCap */
  int *hiz_surv;       /* per coarse bin: those triangles (kHizCap slots) */
  int *hiz_seed;       /* per coarse bin: nearest clipped polygon covering it, or -1 */
  int *hiz_poly;       /* per coarse bin: 1 when that polygon's depth range tightened hiz_bound */
  int hiz_off;         /* 1: no depth-bound culling at all (every triangle is rasterized; for tests) */
  /* listed ordered raster (draws that write no depth): fragments in jobs[0, dcount[3]) */
  float4 *ord_col;  /* shaded colour per listed fragment */
  int2 *ord_span;   /* per pixel of a listed tile: its fragments jobs[x, x + y), in API order */
  int *ord_state;   /* per tile: kOrdEmpty, kOrdListed or kOrdOverflow */
  int *ord_list;    /* tiles whose fragments did not fit (dcount[7] from the front) and tiles with
                       listed fragments (dcount[8] from the back) */
  int ord_cap;      /* jobs slots */
  uint32_t *errors; /* kErrSlots words, see kErrClip */
  unsigned long long *counters; /* [0] fs invocations, [1] samples written, [2] clip overflows,
                                   [3] bin items, [4] primitives binned, [5] clipped triangles,
                                   [6] (tile, triangle) pairs skipped by the hierarchical-Z test */
};

}  // namespace rhip
