/*
 * rhip_api.hip -- librhip: the C-ABI declared in include/rhip.h.
 *
 * Host orchestration of RendererHIP: resources live in device memory, every
 * call queues work on the device's HIP stream, and one draw runs the kernel
 * pipeline of rhip_kernels.h.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "rhip.h"
#include "rhip_host.h"
#include "rhip_neural.h"
#include "rhip_lines.h"

#include <hipcub/hipcub.hpp>

using namespace rhip;

/* Every kernel launch goes through this macro (one place to change how kernels are started). */
/* a comma inside a macro argument (template argument lists) */
#define RHIP_COMMA ,

#ifndef RHIP_LAUNCH
#define RHIP_LAUNCH(kern, grid, block, stream, ...) kern<<<(grid), (block), 0, (stream)>>>(__VA_ARGS__)
#endif

namespace rhip {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace rhip

namespace {

/* Texel index in a level's storage order (host side of texelIndex). */
size_t hostTexelIndex(const TexLevel &l, int layout, int x, int y) {
  if (layout == LAYOUT_TILED)
    return ((size_t)((uint32_t)(y >> 2) * l.tiles_x + (uint32_t)(x >> 2)) << 4) + ((y & 3) << 2) + (x & 3);
  if (layout == LAYOUT_MORTON) {
    uint32_t r = (uint32_t)(x & 31) | ((uint32_t)(y & 31) << 16);
    r = (r | (r << 4)) & 0x0f0f0f0fu;
    r = (r | (r << 2)) & 0x33333333u;
    r = (r | (r << 1)) & 0x55555555u;
    r = (r | (r >> 15)) & 0xffffu;
    return ((size_t)((uint32_t)(y >> 5) * l.tiles_x + (uint32_t)(x >> 5)) << 10) + r;
  }
  return (size_t)y * l.w + x;
}

int levelCount(int w, int h) { return (int)std::floor(std::log2((double)std::max(w, h))) + 1; }

/* Allocates storage for every level of every layer in the texture's layout. */
int allocTexture(rhip_device d, Texture &t) {
  const rhip_texture_desc &ds = t.desc;
  TexDesc &h = t.hd;
  memset(&h, 0, sizeof h);
  h.type = ds.type;
  h.format = ds.format;
  h.width = ds.width;
  h.height = ds.height;
  h.layers = ds.type == 1 ? 6 : 1;
  h.msaa = ds.multi_sample ? 1 : 0;
  /* attachments and multisampled images are always stored linearly */
  h.layout = (ds.multi_sample || (ds.usage & 12)) ? LAYOUT_LINEAR : ds.layout;
  h.levels = ds.use_mipmaps ? levelCount(ds.width, ds.height) : 1;
  if (h.levels > kMaxLevels) return fail(RHIP_E_UNSUPPORTED, "texture too large for %d mip levels", kMaxLevels);
  size_t texel = (size_t)4 * (h.msaa ? 4 : 1);
  size_t off = 0;
  for (int l = 0; l < h.layers; l++)
    for (int v = 0; v < h.levels; v++) {
      TexLevel &lv = h.lv[l][v];
      lv.w = std::max(1, ds.width >> v);
      lv.h = std::max(1, ds.height >> v);
      int ts = h.layout == LAYOUT_TILED ? 4 : (h.layout == LAYOUT_MORTON ? 32 : 1);
      lv.tiles_x = (lv.w + ts - 1) / ts;
      lv.inner_w = lv.tiles_x * ts;
      lv.inner_h = ((lv.h + ts - 1) / ts) * ts;
      lv.offset = off;
      off += (size_t)lv.inner_w * lv.inner_h * texel;
      off = (off + 255) & ~(size_t)255;
    }
  t.bytes = off;
  if (hipMalloc(&t.data, off) != hipSuccess) return fail(RHIP_E_OOM, "texture allocation of %zu bytes failed", off);
  HIPCHK(hipMemsetAsync(t.data, 0, off, d->stream));
  if (h.msaa && ds.format == 0) {
    size_t rb = (size_t)ds.width * ds.height * 4;
    if (hipMalloc(&t.resolved, rb) != hipSuccess) return fail(RHIP_E_OOM, "resolve target allocation failed");
    HIPCHK(hipMemsetAsync(t.resolved, 0, rb, d->stream));
  }
  h.base = t.data;
  h.resolved = t.resolved;
  return RHIP_OK;
}

int syncDesc(rhip_device d, Texture &t) {
  t.hd.filter_min = t.samp.filter_min;
  t.hd.filter_mag = t.samp.filter_mag;
  t.hd.wrap = t.samp.wrap_s; /* the reference samples with wrapS on both axes (SamplerSoft.h:392) */
  t.hd.border = t.samp.border_color;
  HIPCHK(hipMemcpyAsync(t.dd, &t.hd, sizeof(TexDesc), hipMemcpyHostToDevice, d->stream));
  return RHIP_OK;
}

int fillLevel(rhip_device d, Texture &t, int layer, int level, const uint8_t *src) {
  const TexLevel &lv = t.hd.lv[layer][level];
  size_t n = (size_t)lv.inner_w * lv.inner_h;
  std::vector<uint32_t> tmp(n, 0);
  const uint32_t *s = (const uint32_t *)src;
  for (int y = 0; y < lv.h; y++)
    for (int x = 0; x < lv.w; x++) tmp[hostTexelIndex(lv, t.hd.layout, x, y)] = s[(size_t)y * lv.w + x];
  HIPCHK(hipMemcpyAsync(t.data + lv.offset, tmp.data(), n * 4, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return RHIP_OK;
}

}  // namespace

/* decode the 256 / 64 / 48-byte uniform blocks (Material.h:63-95) */
void rhip::decodeUniforms(const rhip_draw_desc &dd, Uniforms &u) {
  const uint8_t *m = (const uint8_t *)dd.uniform_model, *s = (const uint8_t *)dd.uniform_scene,
                *t = (const uint8_t *)dd.uniform_material;
  memset(&u, 0, sizeof u);
  if (m) {
    memcpy(&u.reverse_z, m, 4);
    memcpy(u.model, m + 16, 64);
    memcpy(u.mvp, m + 80, 64);
    memcpy(u.inv_t, m + 144, 48);
    memcpy(u.shadow_mvp, m + 192, 64);
  }
  if (s) {
    memcpy(u.ambient, s, 12);
    memcpy(u.cam_pos, s + 16, 12);
    memcpy(u.light_pos, s + 32, 12);
    memcpy(u.light_color, s + 48, 12);
  }
  if (t) {
    memcpy(&u.enable_light, t, 4);
    memcpy(&u.enable_ibl, t + 4, 4);
    memcpy(&u.enable_shadow, t + 8, 4);
    memcpy(&u.point_size, t + 12, 4);
    memcpy(&u.k_specular, t + 16, 4);
    memcpy(u.base_color, t + 32, 16);
    memcpy(u.aux, t, 16);
  }
}


namespace {

/* Scratch of rhip_draw_noalloc: the arena of a draw that may not allocate
 * (vertex slots, setup records, bins, fragment jobs), one copy per GPU, held
 * until the process ends. Host code never uses the name itself for device
 * memory: each device's address comes from hipGetSymbolAddress at
 * rhip_device_create (rhip_device_s::scratch). */
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

/* After a synchronisation: reports (RHIP_E_DEVICE, naming the stage) and
 * clears what the kernels' error words recorded since the last check. */
int takeDeviceErrors(rhip_device d) {
  uint32_t e[kErrSlots];
  HIPCHK(hipMemcpy(e, d->errors, sizeof e, hipMemcpyDeviceToHost));
  if (!e[kErrClip]) return RHIP_OK;
  HIPCHK(hipMemsetAsync(d->errors, 0, sizeof e, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return fail(RHIP_E_DEVICE,
              "clipping: a clipped polygon exceeded %d vertices and its triangle was dropped; the output differs "
              "from the reference's",
              kMaxPoly);
}

__global__ void k_set_error(uint32_t *errors, int slot) { errors[slot] = 1u; }

hipEvent_t takeEvent(rhip_device d) {
  if (!d->event_pool.empty()) {
    hipEvent_t e = d->event_pool.back();
    d->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

/* Brackets one kernel launch with events when profiling is on. */
struct Timer {
  rhip_device d;
  const char *name;
  hipEvent_t a = nullptr;
  Timer(rhip_device dev, const char *n) : d(dev), name(n) {
    if (d->profiling && !d->no_alloc_call) {
      a = takeEvent(d);
      hipEventRecord(a, d->stream);
    }
  }
  ~Timer() {
    if (a) {
      hipEvent_t b = takeEvent(d);
      hipEventRecord(b, d->stream);
      d->timed.push_back({name, a, b});
    }
  }
};

/* Wavefronts per tile of a single draw's two-phase raster: four for 4x targets,
 * two without MSAA, whose tiles carry a quarter of the samples (measured on the
 * C2 frame: 2048^2 shadow pass 126 -> 111 us with two; the 4x pass 114 -> 152 us). */
template <int S>
constexpr int visWaves() {
  return S == 4 ? 4 : 2;
}

/* Two-phase raster: coarse hierarchical Z, then visibility and depth per tile. */
template <int S>
void launchVis(rhip_device d, const DrawParams &P, int ntiles, int ncoarse, hipStream_t st) {
  const int f = P.st.depthFunc;
  {
    Timer tm(d, "k_hiz_coarse");
    if (f == 4 || f == 6) {
      RHIP_LAUNCH(k_hiz_coarse<S RHIP_COMMA true>, ncoarse, 256, st, P);
    } else {
      RHIP_LAUNCH(k_hiz_coarse<S RHIP_COMMA false>, ncoarse, 256, st, P);
    }
  }
  Timer tm(d, S == 4 ? "k_raster_vis_msaa4" : "k_raster_vis");
  switch (f) {
    case 1: RHIP_LAUNCH(k_raster_vis<S RHIP_COMMA VIS_LESS RHIP_COMMA kTile RHIP_COMMA visWaves<S>()>, ntiles, 64 * visWaves<S>(), st, P); break;
    case 3: RHIP_LAUNCH(k_raster_vis<S RHIP_COMMA VIS_LEQUAL RHIP_COMMA kTile RHIP_COMMA visWaves<S>()>, ntiles, 64 * visWaves<S>(), st, P); break;
    case 4: RHIP_LAUNCH(k_raster_vis<S RHIP_COMMA VIS_GREATER RHIP_COMMA kTile RHIP_COMMA visWaves<S>()>, ntiles, 64 * visWaves<S>(), st, P); break;
    default: RHIP_LAUNCH(k_raster_vis<S RHIP_COMMA VIS_GEQUAL RHIP_COMMA kTile RHIP_COMMA visWaves<S>()>, ntiles, 64 * visWaves<S>(), st, P); break;
  }
}

/* Shading of the fragments the two-phase raster queued. */
template <int S>
void launchShade(const DrawParams &P, int grid, int need_deriv, hipStream_t st) {
  switch (P.shading) {
    case RHS_SHADING_BASIC: RHIP_LAUNCH(k_shade<S RHIP_COMMA RHS_SHADING_BASIC>, grid, 256, st, P, need_deriv); break;
    case RHS_SHADING_BLINN_PHONG:
      RHIP_LAUNCH(k_shade<S RHIP_COMMA RHS_SHADING_BLINN_PHONG>, grid, 256, st, P, need_deriv);
      break;
    case RHS_SHADING_BLINN_PHONG_DIR:
      RHIP_LAUNCH(k_shade<S RHIP_COMMA RHS_SHADING_BLINN_PHONG_DIR>, grid, 256, st, P, need_deriv);
      break;
    case RHS_SHADING_PBR: RHIP_LAUNCH(k_shade<S RHIP_COMMA RHS_SHADING_PBR>, grid, 256, st, P, need_deriv); break;
    default: RHIP_LAUNCH(k_shade<S RHIP_COMMA RHS_SHADING_VERTEX_COLOR>, grid, 256, st, P, need_deriv); break;
  }
}

/* Shading of the listed ordered raster's fragments, and its fallback for tiles that did not fit. */
template <int S>
void launchOrdShade(const DrawParams &P, int grid, int need_deriv, hipStream_t st) {
  switch (P.shading) {
    case RHS_SHADING_BASIC: RHIP_LAUNCH(k_ord_shade<S RHIP_COMMA RHS_SHADING_BASIC>, grid, 256, st, P, need_deriv); break;
    case RHS_SHADING_BLINN_PHONG:
      RHIP_LAUNCH(k_ord_shade<S RHIP_COMMA RHS_SHADING_BLINN_PHONG>, grid, 256, st, P, need_deriv);
      break;
    case RHS_SHADING_BLINN_PHONG_DIR:
      RHIP_LAUNCH(k_ord_shade<S RHIP_COMMA RHS_SHADING_BLINN_PHONG_DIR>, grid, 256, st, P, need_deriv);
      break;
    case RHS_SHADING_PBR: RHIP_LAUNCH(k_ord_shade<S RHIP_COMMA RHS_SHADING_PBR>, grid, 256, st, P, need_deriv); break;
    case RHS_SHADING_IBL_IRRADIANCE:
    case RHS_SHADING_IBL_PREFILTER:
      /* one wavefront per fragment: `grid` blocks of 256 lanes cover the fragments one lane each */
      RHIP_LAUNCH(k_ord_shade_ibl<S>, std::min(grid * 64, 16384), 256, st, P, need_deriv);
      break;
    case RHS_SHADING_SKYBOX:
    case RHS_SHADING_FXAA: RHIP_LAUNCH(k_ord_shade<S RHIP_COMMA kShadingEnv>, grid, 256, st, P, need_deriv); break;
    default: RHIP_LAUNCH(k_ord_shade<S RHIP_COMMA RHS_SHADING_VERTEX_COLOR>, grid, 256, st, P, need_deriv); break;
  }
}

template <int S>
void launchOrdRest(const DrawParams &P, int grid, int need_deriv, hipStream_t st) {
  switch (P.shading) {
    case RHS_SHADING_BASIC: RHIP_LAUNCH(k_ord_rest<S RHIP_COMMA RHS_SHADING_BASIC>, grid, 256, st, P, need_deriv); break;
    case RHS_SHADING_BLINN_PHONG:
      RHIP_LAUNCH(k_ord_rest<S RHIP_COMMA RHS_SHADING_BLINN_PHONG>, grid, 256, st, P, need_deriv);
      break;
    case RHS_SHADING_BLINN_PHONG_DIR:
      RHIP_LAUNCH(k_ord_rest<S RHIP_COMMA RHS_SHADING_BLINN_PHONG_DIR>, grid, 256, st, P, need_deriv);
      break;
    case RHS_SHADING_PBR: RHIP_LAUNCH(k_ord_rest<S RHIP_COMMA RHS_SHADING_PBR>, grid, 256, st, P, need_deriv); break;
    case RHS_SHADING_SKYBOX:
    case RHS_SHADING_IBL_IRRADIANCE:
    case RHS_SHADING_IBL_PREFILTER:
    case RHS_SHADING_FXAA: RHIP_LAUNCH(k_ord_rest<S RHIP_COMMA kShadingEnv>, grid, 256, st, P, need_deriv); break;
    default: RHIP_LAUNCH(k_ord_rest<S RHIP_COMMA RHS_SHADING_VERTEX_COLOR>, grid, 256, st, P, need_deriv); break;
  }
}

/* Listed ordered raster of a draw that writes no depth. */

/* The multisample resolve a draw ends with (RendererSoft.cpp:161-163): launched now, or, for rhip_draw,
 * left pending until something may read the resolved image (checkDev, drawOne). */
}  // namespace

namespace rhip {
int flushResolve(rhip_device d) {
  if (!d->resolve.on) return RHIP_OK;
  d->resolve.on = false;
  Timer tm(d, "k_resolve");
  RHIP_LAUNCH(k_resolve, d->resolve.grid, 256, d->stream, d->resolve.P);
  HIPCHK(hipGetLastError());
  return RHIP_OK;
}
}  // namespace rhip

namespace {

static void resolveAfterDraw(rhip_device d, const DrawParams &P, unsigned grid, hipStream_t s) {
  if (d->defer_resolve && s == d->stream) {
    d->resolve.on = true;
    d->resolve.tex = d->fbo[d->cur_fbo].color;
    d->resolve.grid = grid;
    d->resolve.P = P;
    return;
  }
  Timer tm(d, "k_resolve");
  RHIP_LAUNCH(k_resolve, grid, 256, s, P);
}

template <int S>
void launchListed(rhip_device d, const DrawParams &P, int need_deriv, hipStream_t st) {
  const size_t npx = (size_t)P.fb_w * P.fb_h;
  {
    Timer tm(d, S == 4 ? "k_ord_list_msaa4" : "k_ord_list");
    RHIP_LAUNCH(k_ord_list<S>, std::min(P.ntiles, 768), 256, st, P);
  }
  if (P.color) {
    {
      Timer tm(d, S == 4 ? "k_ord_shade_msaa4" : "k_ord_shade");
      launchOrdShade<S>(P, (int)std::min<size_t>(blocks(npx, 256), 4096), need_deriv, st);
    }
    Timer tm(d, S == 4 ? "k_ord_apply_msaa4" : "k_ord_apply");
    RHIP_LAUNCH(k_ord_apply<S>, std::min(P.ntiles, 1024), 256, st, P);
  }
  {
    Timer tm(d, S == 4 ? "k_ord_rest_msaa4" : "k_ord_rest");
    launchOrdRest<S>(P, std::min(P.ntiles, 512), need_deriv, st);
  }
  if (S == 4 && P.color) resolveAfterDraw(d, P, (unsigned)std::min<size_t>(blocks(npx, 256), 4096), st);
}

/* Ordered raster (every state the two-phase raster does not take). */
template <int S>
void launchOrdered(const DrawParams &P, int ntiles, int need_deriv, int flush_each, hipStream_t st) {
  switch (P.shading) {
    case RHS_SHADING_BASIC: {
#define K_ORD k_raster_ord<S, RHS_SHADING_BASIC>
      RHIP_LAUNCH(K_ORD, ntiles, 256, st, P, need_deriv, flush_each);
#undef K_ORD
      break;
    }
    case RHS_SHADING_BLINN_PHONG: {
#define K_ORD k_raster_ord<S, RHS_SHADING_BLINN_PHONG>
      RHIP_LAUNCH(K_ORD, ntiles, 256, st, P, need_deriv, flush_each);
#undef K_ORD
      break;
    }
    case RHS_SHADING_BLINN_PHONG_DIR: {
#define K_ORD k_raster_ord<S, RHS_SHADING_BLINN_PHONG_DIR>
      RHIP_LAUNCH(K_ORD, ntiles, 256, st, P, need_deriv, flush_each);
#undef K_ORD
      break;
    }
    case RHS_SHADING_PBR: {
#define K_ORD k_raster_ord<S, RHS_SHADING_PBR>
      RHIP_LAUNCH(K_ORD, ntiles, 256, st, P, need_deriv, flush_each);
#undef K_ORD
      break;
    }
    default: {
#define K_ORD k_raster_ord<S, RHS_SHADING_VERTEX_COLOR>
      RHIP_LAUNCH(K_ORD, ntiles, 256, st, P, need_deriv, flush_each);
#undef K_ORD
      break;
    }
  }
}

}  // namespace

namespace rhip {

int drawSetup(rhip_device d, int fbo, const int vp[4], const rhip_draw_desc &dd, int tile, DrawParams &P,
              DrawShape &sh) {
  sh = DrawShape{};
  memset(&P, 0, sizeof P);
  auto vit = d->vao.find(dd.vao);
  auto pit = d->prog.find(dd.program);
  if (vit == d->vao.end()) return fail(RHIP_E_INVALID, "invalid vao %d", dd.vao);
  if (pit == d->prog.end()) return fail(RHIP_E_INVALID, "invalid program %d", dd.program);
  const Vao &va = *vit->second;
  const Program &pg = pit->second;
  const RhsStates &st = dd.states;
  if (st.primitiveType < 0 || st.primitiveType > 2 || st.polygonMode < 0 || st.polygonMode > 2)
    return fail(RHIP_E_INVALID, "primitive type %d, polygon mode %d", st.primitiveType, st.polygonMode);
  sh.lines = st.primitiveType != 2 || st.polygonMode != 2;
  if (sh.lines) {
    /* points and lines shade without a derivative context (RendererSoft.cpp:637-661); the reference's
     * Viewer draws them with the base colour shader only (Viewer.cpp:264-275, ModelLoader.cpp:60-85) */
    if (pg.shading != RHS_SHADING_BASIC && pg.shading != RHS_SHADING_VERTEX_COLOR)
      return fail(RHIP_E_UNSUPPORTED, "points and lines draw with the base colour or vertex colour shader only");
    if (!(st.lineWidth >= 0.f && st.lineWidth <= 1024.f))
      return fail(RHIP_E_UNSUPPORTED, "line width %g (0 to 1024)", (double)st.lineWidth);
  }
  if (st.depthFunc < 0 || st.depthFunc > 7 || st.blendFuncRgb < 0 || st.blendFuncRgb > 4)
    return fail(RHIP_E_INVALID, "invalid render states");
  auto fit = d->fbo.find(fbo);
  if (fit == d->fbo.end()) return fail(RHIP_E_INVALID, "invalid framebuffer %d", fbo);
  const Fbo &f = fit->second;
  Texture *ct = f.color ? getTex(d, f.color) : nullptr;
  Texture *dt = f.depth ? getTex(d, f.depth) : nullptr;
  if (!ct && !dt) {
    sh.empty = true;
    return RHIP_OK;
  }
  if (vp[2] <= 0 || vp[3] <= 0) return fail(RHIP_E_INVALID, "viewport not set");
  sh.two_phase = !sh.lines && !st.blend && dt && st.depthTest && st.depthMask &&
                 (st.depthFunc == 1 || st.depthFunc == 3 || st.depthFunc == 4 || st.depthFunc == 6);
  /* a draw that writes no depth: every fragment's tests see the depths the draw started with */
  sh.listed = !sh.lines && !sh.two_phase && !(dt && st.depthTest && st.depthMask);

  /* a learned-shader draw of a material that failed its error bound draws with the plain shader */
  int shading = pg.shading;
  if (shading == RHS_SHADING_NEURAL_PBR && d->nn.draw_material >= 0 && d->nn.draw_material < 64 &&
      ((d->nn.fallback >> d->nn.draw_material) & 1ull))
    shading = RHS_SHADING_PBR;
  if (isEnvShading(shading) && !sh.listed && !sh.lines)
    return fail(RHIP_E_UNSUPPORTED,
                "skybox, IBL and FXAA programs draw without depth writes (the reference's states for them); "
                "this draw writes depth");
  if (shading == RHS_SHADING_NEURAL_PBR) {
    uint32_t enable_light = 0;
    if (dd.uniform_material) memcpy(&enable_light, dd.uniform_material, 4);
    if (!d->nn.ready) return fail(RHIP_E_INVALID, "learned shader: no network (rhip_neural_init or set_weights)");
    if (!sh.two_phase || enable_light || (pg.defines & (RHS_DEF_AO_MAP | RHS_DEF_EMISSIVE_MAP)))
      return fail(RHIP_E_UNSUPPORTED,
                  "learned shader: image-based lighting only (no point light, AO or emissive map), opaque "
                  "draws that write depth with LESS/LEQUAL/GREATER/GEQUAL");
  }
  decodeUniforms(dd, P.u);
  P.st = st;
  P.pl_point_size = shading == RHS_SHADING_BASIC ? P.u.point_size : 1.0f; /* BasicSoft.h:68 */
  if (sh.lines && !(P.pl_point_size >= 0.f && P.pl_point_size <= 1024.f))
    return fail(RHIP_E_UNSUPPORTED, "point size %g (0 to 1024)", (double)P.pl_point_size);
  P.shading = shading;
  P.defines = pg.defines;
  P.vary_groups = varyGroups(P.shading, P.defines, P.u.enable_shadow);
  P.vary_stride = 4 * __builtin_popcount(P.vary_groups);
  for (int s = 0; s < RHS_SLOT_COUNT; s++) {
    int id = dd.sampler_tex[s];
    if (!id) continue;
    Texture *t = getTex(d, id);
    if (!t) return fail(RHIP_E_INVALID, "sampler slot %d: invalid texture %d", s, id);
    if (!t->initialised) return fail(RHIP_E_INVALID, "sampler slot %d: texture %d has no image", s, id);
    P.tex[s] = t->dd;
    /* only the Blinn-Phong and PBR shaders set a derivative LOD on their samplers (PbrSoft.h:138-151) */
    if (t->samp.filter_min > 1 && s != RHS_SLOT_METAL_ROUGHNESS && s <= RHS_SLOT_AO && !isEnvShading(shading))
      P.need_deriv = 1;
  }
  P.vp_w = (float)vp[2];
  P.vp_h = (float)vp[3];
  P.o_x = (float)vp[0] + P.vp_w / 2.f;
  P.o_y = (float)vp[1] + P.vp_h / 2.f;
  P.p_x = P.vp_w / 2.f;
  P.p_y = P.vp_h / 2.f;
  int fbw = 0, fbh = 0, samples = 1;
  if (ct) {
    const TexLevel &lv = ct->hd.lv[f.color_layer][f.color_level];
    P.color = (uint32_t *)(ct->data + lv.offset);
    P.color_resolved = (uint32_t *)ct->resolved;
    fbw = lv.w;
    fbh = lv.h;
    samples = ct->hd.msaa ? 4 : 1;
  }
  if (dt) {
    const TexLevel &lv = dt->hd.lv[0][0];
    P.depth = (float *)(dt->data + lv.offset);
    if (!ct) {
      fbw = lv.w;
      fbh = lv.h;
      samples = dt->hd.msaa ? 4 : 1;
    } else if (lv.w != fbw || lv.h != fbh || (dt->hd.msaa ? 4 : 1) != samples) {
      return fail(RHIP_E_UNSUPPORTED, "colour and depth attachments differ in size or sample count");
    }
  }
  P.fb_w = fbw;
  P.fb_h = fbh;
  P.samples = samples;
  P.tile = tile;
  P.coarse_tiles = kCoarse;
  P.tiles_x = (fbw + tile - 1) / tile;
  P.tiles_y = (fbh + tile - 1) / tile;
  P.ntiles = P.tiles_x * P.tiles_y;
  P.coarse_x = (P.tiles_x + kCoarse - 1) / kCoarse;
  P.coarse_y = (P.tiles_y + kCoarse - 1) / kCoarse;
  P.nbins = P.ntiles + P.coarse_x * P.coarse_y + 1;
  P.verts = va.verts;
  P.stride_f = va.stride / 4;
  P.nverts = va.nverts;
  P.idx = va.idx;
  P.ntris = va.nidx / 3;
  P.counters = d->counters;
  P.errors = d->errors;
  P.hiz_off = d->hiz_off ? 1 : 0;
  P.gamma_lut = d->gamma_lut;
  P.libm = d->libm;
  P.ibl = d->ibl;
  for (int c = 0; c < 3; c++) P.base_lin[c] = hostPowf(P.u.base_color[c], 2.2f); /* PbrSoft.h:235 */
  sh.V = (size_t)va.nverts;
  sh.T = (size_t)P.ntris;
  if (sh.lines) {
    sh.units = plUnitCount(st.primitiveType, va.nidx);
    sh.T = 0; /* no triangle pipeline */
    P.nblk = 0;
    P.nbins = 0;
    return RHIP_OK;
  }
  sh.stride = va.stride;
  P.nblk = (int)blocks(sh.T, 256);
  const size_t vcap = sh.V + sh.T * kMaxClipVerts, pcap = sh.T * (1 + kMaxFans);
  if (pcap * kMaxBinItems >= (size_t)INT32_MAX || vcap >= (size_t)INT32_MAX)
    return fail(RHIP_E_UNSUPPORTED, "draw too large (%zu triangles)", sh.T);
  return RHIP_OK;
}

/* Arenas sized for the worst case, so that no count has to come back to the
 * host: every triangle may clip (kMaxClipVerts new vertices, kMaxFans extra
 * triangles) and every primitive is listed in at most kMaxBinItems bins. */
size_t carveArenas(DrawParams &P, const DrawShape &sh, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t V = sh.V, T = sh.T, vcap = V + T * kMaxClipVerts, pcap = T * (1 + kMaxFans);
  const int nblk = P.nblk;
  P.clip = (float4 *)take(vcap * 16);
  P.frag = (float4 *)take(vcap * 16);
  P.mask = (int *)take(vcap * 4);
  P.vary = (float *)take(vcap * P.vary_stride * 4);
  P.clip_raw = (float *)take(T * kMaxClipVerts * sh.stride);
  P.clip_poly = (int *)take(T * (kMaxPoly + 1) * 4);
  P.fan_first = (int *)take(T * 4);
  P.blk_clip = (int *)take(((size_t)nblk * 3 + fanChunks(nblk)) * 4);
  P.blk_fans = P.blk_clip ? P.blk_clip + nblk : nullptr;
  P.blk_fan_off = P.blk_fans ? P.blk_fans + nblk : nullptr;
  P.chunk_fans = P.blk_fan_off ? P.blk_fan_off + nblk : nullptr;
  P.dcount = (int *)take(64);
  P.prims = (PrimSetup *)take(pcap * sizeof(PrimSetup));
  P.bin_count = (int *)take(((size_t)P.nbins * 3 + 1) * 4);
  P.bin_start = P.bin_count ? P.bin_count + P.nbins : nullptr;
  P.bin_cursor = P.bin_start ? P.bin_start + P.nbins + 1 : nullptr;
  P.bin_items = (int *)take(pcap * kMaxBinItems * 4);
  if (sh.two_phase) {
    const int ncoarse = P.coarse_x * P.coarse_y;
    P.jobs = (uint2 *)take((size_t)P.fb_w * P.fb_h * P.samples * 8);
    P.hiz_bound = (uint32_t *)take((size_t)ncoarse * (4 + kHizCap) * 4);
    P.hiz_count = P.hiz_bound ? (int *)P.hiz_bound + ncoarse : nullptr;
    P.hiz_seed = P.hiz_count ? P.hiz_count + ncoarse : nullptr;
    P.hiz_poly = P.hiz_seed ? P.hiz_seed + ncoarse : nullptr;
    P.hiz_surv = P.hiz_poly ? P.hiz_poly + ncoarse : nullptr;
  }
  if (sh.listed) {
    const size_t px = (size_t)P.fb_w * P.fb_h;
    const size_t cap = std::min<size_t>(2 * px + 65536, (size_t)1 << 30);
    P.ord_cap = (int)cap;
    P.jobs = (uint2 *)take(cap * 8);
    P.ord_col = (float4 *)take(cap * 16);
    P.ord_span = (int2 *)take(px * 8);
    P.ord_state = (int *)take((size_t)P.ntiles * 4);
    P.ord_list = (int *)take((size_t)P.ntiles * 4);
  }
  return off;
}

}  // namespace rhip

extern "C" {

const char *rhip_last_error(void) { return g_err.c_str(); }

int rhip_device_create(int hip_device, rhip_device *out) {
  if (!out) return fail(RHIP_E_INVALID, "null output");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (hip_device < 0 || hip_device >= n) return fail(RHIP_E_INVALID, "no HIP device %d (%d present)", hip_device, n);
  HIPCHK(hipSetDevice(hip_device));
  std::unique_ptr<rhip_device_s> d(new rhip_device_s());
  d->hip_device = hip_device;
  HIPCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  HIPCHK(hipGetSymbolAddress((void **)&d->scratch, HIP_SYMBOL(scratch)));
  HIPCHK(hipMalloc(&d->counters, kCounterSlots * kCounterStride * sizeof(unsigned long long)));
  HIPCHK(hipMemset(d->counters, 0, kCounterSlots * kCounterStride * sizeof(unsigned long long)));
  HIPCHK(hipMalloc(&d->gamma_lut, 512 * sizeof(float)));
  HIPCHK(hipMalloc(&d->errors, kErrSlots * sizeof(uint32_t)));
  HIPCHK(hipMemset(d->errors, 0, kErrSlots * sizeof(uint32_t)));
  {
    float gt[512];
    gammaTablesHost(gt);
    HIPCHK(hipMemcpy(d->gamma_lut, gt, sizeof gt, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMalloc(&d->ibl, sizeof(IblTables)));
  {
    std::unique_ptr<IblTables> it(new IblTables());
    iblTablesHost(it.get());
    HIPCHK(hipMemcpy(d->ibl, it.get(), sizeof(IblTables), hipMemcpyHostToDevice));
  }
  /* the host libm on the device: every table, built once per process */
  {
    /* log2f and exp2f need no table (glibc's own algorithms): only the powf tables are built here */
    int rc = libmEnsure(d.get(), (1u << kLmPow5) | (1u << kLmPow128) | (1u << kLmPowGamma));
    if (rc) {
      if (d->libm) hipFree(d->libm);
      for (void *p : d->libm_bufs) hipFree(p);
      return rc;
    }
  }
  *out = d.release();
  return RHIP_OK;
}

int rhip_device_destroy(rhip_device d) {
  if (!d) return RHIP_OK;
  hipSetDevice(d->hip_device);
  hipStreamSynchronize(d->stream);
  d->batches.clear();
  d->tex.clear();
  d->vao.clear();
  if (d->counters) hipFree(d->counters);
  if (d->gamma_lut) hipFree(d->gamma_lut);
  if (d->ibl) hipFree(d->ibl);
  if (d->libm) hipFree(d->libm);
  for (void *p : d->libm_bufs) hipFree(p);
  if (d->errors) hipFree(d->errors);
  for (void *p : {(void *)d->nn.params, (void *)d->nn.grad, (void *)d->nn.m1, (void *)d->nn.m2, (void *)d->nn.packed,
                  (void *)d->nn.scratch, (void *)d->nn.hist, (void *)d->nn.rec_count, (void *)d->nn.partial,
                  (void *)d->nn.lossp})
    if (p) hipFree(p);
  hipStreamDestroy(d->stream);
  delete d;
  return RHIP_OK;
}

void *rhip_device_stream(rhip_device d) { return d ? (void *)d->stream : nullptr; }

/* ------------------------------------------------------------------ textures */

int rhip_texture_create(rhip_device d, const rhip_texture_desc *desc, int *out) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!desc || !out) return fail(RHIP_E_INVALID, "null argument");
  if (desc->width <= 0 || desc->height <= 0 || desc->width > 16384 || desc->height > 16384)
    return fail(RHIP_E_INVALID, "texture size %dx%d out of range", desc->width, desc->height);
  if (desc->type != 0 && desc->type != 1) return fail(RHIP_E_INVALID, "texture type %d", desc->type);
  if (desc->format != 0 && desc->format != 1) return fail(RHIP_E_INVALID, "texture format %d", desc->format);
  if (desc->layout < 0 || desc->layout > 2) return fail(RHIP_E_INVALID, "texture layout %d", desc->layout);
  std::unique_ptr<Texture> t(new Texture());
  t->desc = *desc;
  t->samp.filter_min = t->samp.filter_mag = 0;
  t->samp.wrap_s = t->samp.wrap_t = t->samp.wrap_r = 2;
  if ((rc = allocTexture(d, *t))) return rc;
  HIPCHK(hipMalloc(&t->dd, sizeof(TexDesc)));
  if ((rc = syncDesc(d, *t))) return rc;
  int id = d->next_id++;
  d->tex[id] = std::move(t);
  *out = id;
  return RHIP_OK;
}

int rhip_texture_set_sampler(rhip_device d, int id, const rhip_sampler_desc *s) {
  int rc = checkDev(d);
  if (rc) return rc;
  Texture *t = getTex(d, id);
  if (!t || !s) return fail(RHIP_E_INVALID, "invalid texture %d", id);
  if (s->filter_min < 0 || s->filter_min > 5 || s->wrap_s < 0 || s->wrap_s > 3)
    return fail(RHIP_E_INVALID, "invalid sampler description");
  t->samp = *s;
  return syncDesc(d, *t);
}

int rhip_texture_init(rhip_device d, int id) {
  int rc = checkDev(d);
  if (rc) return rc;
  Texture *t = getTex(d, id);
  if (!t) return fail(RHIP_E_INVALID, "invalid texture %d", id);
  t->initialised = true; /* storage is allocated and zeroed at creation */
  return RHIP_OK;
}

int rhip_texture_upload(rhip_device d, int id, int layers, const void *data, size_t bytes) {
  int rc = checkDev(d);
  if (rc) return rc;
  Texture *t = getTex(d, id);
  if (!t || !data) return fail(RHIP_E_INVALID, "invalid texture %d", id);
  if (t->desc.multi_sample) return fail(RHIP_E_INVALID, "setImageData not supported on a multisampled texture");
  if (layers != t->hd.layers) return fail(RHIP_E_INVALID, "texture has %d layers, got %d", t->hd.layers, layers);
  size_t per = (size_t)t->desc.width * t->desc.height * 4;
  if (bytes < per * layers) return fail(RHIP_E_INVALID, "image data too small: %zu < %zu", bytes, per * layers);
  for (int l = 0; l < layers; l++)
    if ((rc = fillLevel(d, *t, l, 0, (const uint8_t *)data + per * l))) return rc;
  for (int l = 0; l < layers; l++)
    for (int v = 1; v < t->hd.levels; v++) {
      const TexLevel &lv = t->hd.lv[l][v];
      dim3 b(16, 16), g(blocks(lv.w, 16), blocks(lv.h, 16));
      if (t->desc.format == 0)
        {
#define K_MIP k_mipgen<uint32_t>
        RHIP_LAUNCH(K_MIP, g, b, d->stream, t->dd, l, v);
#undef K_MIP
      }
      else
        {
#define K_MIP k_mipgen<float>
        RHIP_LAUNCH(K_MIP, g, b, d->stream, t->dd, l, v);
#undef K_MIP
      }
      HIPCHK(hipGetLastError());
    }
  t->initialised = true;
  return RHIP_OK;
}

int rhip_texture_upload_level(rhip_device d, int id, int layer, int level, const void *data, size_t bytes) {
  int rc = checkDev(d);
  if (rc) return rc;
  Texture *t = getTex(d, id);
  if (!t || !data) return fail(RHIP_E_INVALID, "invalid texture %d", id);
  if (t->desc.multi_sample) return fail(RHIP_E_INVALID, "multisampled texture");
  if (layer < 0 || layer >= t->hd.layers || level < 0 || level >= t->hd.levels)
    return fail(RHIP_E_INVALID, "layer %d / level %d out of range", layer, level);
  const TexLevel &lv = t->hd.lv[layer][level];
  if (bytes < (size_t)lv.w * lv.h * 4) return fail(RHIP_E_INVALID, "level data too small");
  return fillLevel(d, *t, layer, level, (const uint8_t *)data);
}

int rhip_texture_read(rhip_device d, int id, int layer, int level, int which, void *out, size_t bytes) {
  int rc = checkDev(d);
  if (rc) return rc;
  Texture *t = getTex(d, id);
  if (!t || !out) return fail(RHIP_E_INVALID, "invalid texture %d", id);
  if (layer < 0 || layer >= t->hd.layers || level < 0 || level >= t->hd.levels)
    return fail(RHIP_E_INVALID, "layer %d / level %d out of range", layer, level);
  const TexLevel &lv = t->hd.lv[layer][level];
  if (which == 1) {
    if (!t->desc.multi_sample) return fail(RHIP_E_INVALID, "texture is not multisampled");
    size_t n = (size_t)lv.w * lv.h * 16;
    if (bytes < n) return fail(RHIP_E_INVALID, "output too small");
    HIPCHK(hipMemcpyAsync(out, t->data + lv.offset, n, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    return takeDeviceErrors(d);
  }
  size_t n = (size_t)lv.w * lv.h * 4;
  if (bytes < n) return fail(RHIP_E_INVALID, "output too small");
  if (t->desc.multi_sample) {
    if (!t->resolved) return fail(RHIP_E_INVALID, "a multisampled depth texture has no single-sample image");
    HIPCHK(hipMemcpyAsync(out, t->resolved, n, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    return takeDeviceErrors(d);
  }
  std::vector<uint32_t> tmp((size_t)lv.inner_w * lv.inner_h);
  HIPCHK(hipMemcpyAsync(tmp.data(), t->data + lv.offset, tmp.size() * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  uint32_t *o = (uint32_t *)out;
  for (int y = 0; y < lv.h; y++)
    for (int x = 0; x < lv.w; x++) o[(size_t)y * lv.w + x] = tmp[hostTexelIndex(lv, t->hd.layout, x, y)];
  return takeDeviceErrors(d);
}

int rhip_texture_destroy(rhip_device d, int id) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!d->tex.count(id)) return fail(RHIP_E_INVALID, "invalid texture %d", id);
  HIPCHK(hipStreamSynchronize(d->stream));
  d->tex.erase(id);
  return RHIP_OK;
}

/* ------------------------------------------------------------------ vertex arrays */

int rhip_vao_create(rhip_device d, const void *vertices, int stride, int nverts, const int32_t *indices,
                    int nindices, int *out) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!out || !vertices || !indices) return fail(RHIP_E_INVALID, "null argument");
  if (stride <= 0 || stride % 4 || stride > 64) return fail(RHIP_E_UNSUPPORTED, "vertex stride %d (need 4..64, x4)", stride);
  if (nverts <= 0 || nindices < 0) return fail(RHIP_E_INVALID, "bad vertex/index count");
  for (int i = 0; i < nindices; i++)
    if (indices[i] < 0 || indices[i] >= nverts) return fail(RHIP_E_INVALID, "index %d out of range at %d", indices[i], i);
  std::unique_ptr<Vao> v(new Vao());
  v->stride = stride;
  v->nverts = nverts;
  v->nidx = nindices;
  if (hipMalloc(&v->verts, (size_t)stride * nverts) != hipSuccess) return fail(RHIP_E_OOM, "vertex buffer");
  if (hipMalloc(&v->idx, (size_t)std::max(nindices, 1) * 4) != hipSuccess) return fail(RHIP_E_OOM, "index buffer");
  HIPCHK(hipMemcpyAsync(v->verts, vertices, (size_t)stride * nverts, hipMemcpyHostToDevice, d->stream));
  if (nindices) HIPCHK(hipMemcpyAsync(v->idx, indices, (size_t)nindices * 4, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  int id = d->next_id++;
  d->vao[id] = std::move(v);
  *out = id;
  return RHIP_OK;
}

int rhip_vao_update(rhip_device d, int id, const void *vertices, size_t bytes) {
  int rc = checkDev(d);
  if (rc) return rc;
  auto it = d->vao.find(id);
  if (it == d->vao.end() || !vertices) return fail(RHIP_E_INVALID, "invalid vao %d", id);
  size_t n = std::min(bytes, (size_t)it->second->stride * it->second->nverts);
  HIPCHK(hipMemcpyAsync(it->second->verts, vertices, n, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return RHIP_OK;
}

int rhip_vao_destroy(rhip_device d, int id) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!d->vao.count(id)) return fail(RHIP_E_INVALID, "invalid vao %d", id);
  HIPCHK(hipStreamSynchronize(d->stream));
  d->vao.erase(id);
  return RHIP_OK;
}

/* ------------------------------------------------------------------ framebuffers / programs */

int rhip_fbo_create(rhip_device d, int *out) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!out) return fail(RHIP_E_INVALID, "null output");
  int id = d->next_id++;
  d->fbo[id] = Fbo();
  *out = id;
  return RHIP_OK;
}

int rhip_fbo_attach_color(rhip_device d, int id, int tex, int layer, int level) {
  auto it = d ? d->fbo.find(id) : decltype(d->fbo.end())();
  if (!d || it == d->fbo.end()) return fail(RHIP_E_INVALID, "invalid framebuffer %d", id);
  Texture *t = getTex(d, tex);
  if (!t || t->desc.format != 0) return fail(RHIP_E_INVALID, "colour attachment must be an RGBA8 texture");
  if (layer < 0 || layer >= t->hd.layers || level < 0 || level >= t->hd.levels)
    return fail(RHIP_E_INVALID, "attachment layer/level out of range");
  it->second.color = tex;
  it->second.color_layer = layer;
  it->second.color_level = level;
  return RHIP_OK;
}

int rhip_fbo_attach_depth(rhip_device d, int id, int tex) {
  auto it = d ? d->fbo.find(id) : decltype(d->fbo.end())();
  if (!d || it == d->fbo.end()) return fail(RHIP_E_INVALID, "invalid framebuffer %d", id);
  Texture *t = getTex(d, tex);
  if (!t || t->desc.format != 1) return fail(RHIP_E_INVALID, "depth attachment must be a FLOAT32 texture");
  it->second.depth = tex;
  return RHIP_OK;
}

int rhip_fbo_destroy(rhip_device d, int id) {
  if (!d || !d->fbo.erase(id)) return fail(RHIP_E_INVALID, "invalid framebuffer %d", id);
  if (d->cur_fbo == id) d->cur_fbo = 0;
  return RHIP_OK;
}

int rhip_program_create(rhip_device d, int shading, uint32_t defines, int *out) {
  if (!d || !out) return fail(RHIP_E_INVALID, "null argument");
  if (shading < RHS_SHADING_BASIC || shading > RHS_SHADING_FXAA)
    return fail(RHIP_E_UNSUPPORTED, "shading model %d", shading);
  if (defines & ~(uint32_t)(2 * RHS_DEF_EQUIRECTANGULAR_MAP - 1))
    return fail(RHIP_E_INVALID, "unknown define bits 0x%x", defines);
  int id = d->next_id++;
  d->prog[id] = Program{shading, defines};
  *out = id;
  return RHIP_OK;
}

/* ------------------------------------------------------------------ pipeline */

int rhip_begin_render_pass(rhip_device d, int fbo, const rhip_clear_states *cs) {
  int rc = checkDev(d);
  if (rc) return rc;
  auto it = d->fbo.find(fbo);
  if (it == d->fbo.end()) {
    d->cur_fbo = 0; /* the reference ignores draws without a framebuffer (RendererSoft.cpp:62-66) */
    return fail(RHIP_E_INVALID, "invalid framebuffer %d", fbo);
  }
  d->cur_fbo = fbo;
  if (!cs) return RHIP_OK;
  const Fbo &f = it->second;
  /* RendererSoft::beginRenderPass (RendererSoft.cpp:61-90): u8 truncation of clear * 255; both clears in one launch */
  uint32_t *cp = nullptr, *dp = nullptr, cv = 0, dv = 0;
  size_t cn = 0, dn = 0;
  auto launch = [&]() {
    const size_t most = std::max(cn, dn) / 4 + 1;
    RHIP_LAUNCH(k_fill2_u32, std::min(blocks(most, 256), 4096u), 256, d->stream, cp, cn, cv, dp, dn, dv);
    return hipGetLastError();
  };
  if (cs->color_flag && f.color) {
    Texture *t = getTex(d, f.color);
    if (!t) return fail(RHIP_E_INVALID, "colour attachment was destroyed");
    auto u8 = [](float v) { return (uint32_t)(uint8_t)(int32_t)(v * 255); };
    cv = u8(cs->clear_color[0]) | (u8(cs->clear_color[1]) << 8) | (u8(cs->clear_color[2]) << 16) |
         (u8(cs->clear_color[3]) << 24);
    const TexLevel &lv = t->hd.lv[f.color_layer][f.color_level];
    cn = (size_t)lv.w * lv.h * (t->hd.msaa ? 4 : 1);
    cp = (uint32_t *)(t->data + lv.offset);
  }
  if (cs->depth_flag && f.depth) {
    Texture *t = getTex(d, f.depth);
    if (!t) {
      if (cp) HIPCHK(launch()); /* the colour clear still happens, as when the clears were separate */
      return fail(RHIP_E_INVALID, "depth attachment was destroyed");
    }
    const TexLevel &lv = t->hd.lv[0][0];
    dn = (size_t)lv.w * lv.h * (t->hd.msaa ? 4 : 1);
    memcpy(&dv, &cs->clear_depth, 4);
    dp = (uint32_t *)(t->data + lv.offset);
  }
  if (cp || dp) HIPCHK(launch());
  return RHIP_OK;
}

int rhip_set_viewport(rhip_device d, int x, int y, int w, int h) {
  if (!d) return fail(RHIP_E_INVALID, "null device");
  d->vp_x = x;
  d->vp_y = y;
  d->vp_w = w;
  d->vp_h = h;
  return RHIP_OK;
}

int rhip_end_render_pass(rhip_device d) { return checkDev(d); /* launches a deferred resolve */ }

int rhip_wait_idle(rhip_device d) {
  int rc = checkDev(d);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(d->stream));
  return takeDeviceErrors(d);
}

int rhip_set_hierarchical_z(rhip_device d, int enabled) {
  if (!d) return fail(RHIP_E_INVALID, "null device");
  d->hiz_off = enabled == 0;
  return RHIP_OK;
}

int rhip_debug_raise_device_error(rhip_device d, int stage) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (stage != kErrClip) return fail(RHIP_E_INVALID, "no device error stage %d", stage);
  RHIP_LAUNCH(k_set_error, 1, 1, d->stream, d->errors, stage);
  HIPCHK(hipGetLastError());
  return RHIP_OK;
}

}  // extern "C"

namespace {

/* A draw of points, lines, or triangles in point or line mode (rhip_lines.h). The fragment list is
 * sized from the count the units report, which is read back once per draw. */
int drawPointsLines(rhip_device d, DrawParams &P, const DrawShape &sh, bool no_alloc) {
  if (no_alloc)
    return fail(RHIP_E_UNSUPPORTED,
                "rhip_draw_noalloc: points and lines size their fragment list at draw time; use rhip_draw");
  hipStream_t s = d->stream;
  const long long U = sh.units, V = (long long)sh.V;
  const bool has_colour = P.color != nullptr;
  const bool points = P.st.primitiveType == 0 || (P.st.primitiveType == 2 && P.st.polygonMode == 0);
  if (U > (long long)INT32_MAX) return fail(RHIP_E_UNSUPPORTED, "draw too large (%lld points or lines)", U);
  /* phase 1: vertices, units, per-unit counts and offsets, the point sort */
  size_t sort_tmp = 0, scan_tmp = 0;
  if (points)
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (uint32_t *)nullptr, (int)U, 0, 32, s));
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, (unsigned long long *)nullptr,
                                          (unsigned long long *)nullptr, (int)U, s));
  size_t off = 0;
  char *base = nullptr;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return q;
  };
  PlUnit *units = nullptr;
  unsigned long long *count = nullptr, *offset = nullptr;
  uint32_t *pkey = nullptr, *pval = nullptr, *skey = nullptr, *sval = nullptr;
  void *tmp = nullptr;
  for (int pass = 0; pass < 2; pass++) {
    off = 0;
    P.clip = (float4 *)take((size_t)V * 16);
    P.frag = (float4 *)take((size_t)V * 16);
    P.mask = (int *)take((size_t)V * 4);
    P.vary = (float *)take((size_t)V * P.vary_stride * 4);
    units = (PlUnit *)take((size_t)U * sizeof(PlUnit));
    count = (unsigned long long *)take((size_t)U * 8);
    offset = (unsigned long long *)take((size_t)U * 8);
    pkey = (uint32_t *)take((size_t)U * 4);
    pval = (uint32_t *)take((size_t)U * 4);
    skey = (uint32_t *)take((size_t)U * 4);
    sval = (uint32_t *)take((size_t)U * 4);
    tmp = take(std::max(sort_tmp, scan_tmp));
    if (pass == 0) {
      int rc = d->arena.ensure(off);
      if (rc) return rc;
      base = d->arena.as<char>();
    }
  }
  { Timer tm(d, "k_vertex"); RHIP_LAUNCH(k_vertex, std::max(blocks((size_t)V, 256), 1u), 256, s, P); }
  unsigned long long total = 0;
  if (U > 0) {
    const unsigned g = blocks((size_t)U, 256);
    { Timer tm(d, "k_pl_units"); RHIP_LAUNCH(k_pl_units, g, 256, s, P, units, U, count, pkey, pval, has_colour ? 1 : 0); }
    if (points) {
      size_t t = sort_tmp;
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, t, pkey, skey, pval, sval, (int)U, 0, 32, s));
      Timer tm(d, "k_pl_points");
      RHIP_LAUNCH(k_pl_points, g, 256, s, P, units, skey, sval, U, count, has_colour ? 1 : 0);
    }
    size_t t = scan_tmp;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, t, count, offset, (int)U, s));
    unsigned long long last[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&last[0], offset + (U - 1), 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&last[1], count + (U - 1), 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    total = last[0] + last[1];
  }
  HIPCHK(hipGetLastError());
  if (total > (unsigned long long)INT32_MAX)
    return fail(RHIP_E_UNSUPPORTED, "draw too large (%llu point and line fragments)", total);
  if (total > 0) {
    /* phase 2: the fragments, sorted by pixel, then applied pixel by pixel */
    const long long F = (long long)total;
    const uint32_t npix = (uint32_t)P.fb_w * (uint32_t)P.fb_h;
    int bits = 1;
    while (bits < 32 && (1ull << bits) <= (unsigned long long)npix) bits++;
    size_t fsort_tmp = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, fsort_tmp, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (uint32_t *)nullptr, (int)F, 0, bits, s));
    char *fb = nullptr;
    size_t foff = 0;
    auto ftake = [&](size_t bytes) {
      char *q = fb ? fb + foff : nullptr;
      foff += (bytes + 255) & ~(size_t)255;
      return q;
    };
    uint32_t *key = nullptr, *val = nullptr, *key2 = nullptr, *val2 = nullptr;
    float *zf = nullptr;
    float4 *cf = nullptr;
    void *ftmp = nullptr;
    for (int pass = 0; pass < 2; pass++) {
      foff = 0;
      key = (uint32_t *)ftake((size_t)F * 4);
      val = (uint32_t *)ftake((size_t)F * 4);
      key2 = (uint32_t *)ftake((size_t)F * 4);
      val2 = (uint32_t *)ftake((size_t)F * 4);
      zf = (float *)ftake((size_t)F * 4);
      cf = (float4 *)ftake((size_t)F * 16);
      ftmp = ftake(fsort_tmp);
      if (pass == 0) {
        int rc = d->lines.ensure(foff);
        if (rc) return rc;
        fb = d->lines.as<char>();
      }
    }
    const unsigned g = blocks((size_t)U, 256), gf = blocks((size_t)F, 256);
    { Timer tm(d, "k_pl_emit"); RHIP_LAUNCH(k_pl_emit, g, 256, s, P, units, U, offset, key, val, zf, cf); }
    size_t t = fsort_tmp;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(ftmp, t, key, key2, val, val2, (int)F, 0, bits, s));
    Timer tm(d, "k_pl_apply");
    if (P.samples == 4)
      RHIP_LAUNCH(k_pl_apply<4>, gf, 256, s, P, key2, val2, zf, cf, F);
    else
      RHIP_LAUNCH(k_pl_apply<1>, gf, 256, s, P, key2, val2, zf, cf, F);
  }
  /* RendererSoft::draw resolves a multisampled colour target after every draw (:161-163) */
  if (has_colour && P.samples == 4)
    resolveAfterDraw(d, P, (unsigned)std::min<size_t>(blocks((size_t)P.fb_w * P.fb_h, 256), 4096), s);
  HIPCHK(hipGetLastError());
  d->host_stats.draws++;
  return RHIP_OK;
}

/* One draw (RendererSoft::draw, RendererSoft.cpp:136-164) with its arena in
 * the device's grow-only buffer, or, for no_alloc, in the library's static
 * scratch of kScratchBytes. */
int drawOne(rhip_device d, const rhip_draw_desc *dd, bool no_alloc) {
  if (!d) return fail(RHIP_E_INVALID, "null device");
  const hipError_t he = hipSetDevice(d->hip_device);
  if (he != hipSuccess) return fail(RHIP_E_HIP, "hipSetDevice: %s", hipGetErrorString(he));
  int rc = RHIP_OK;
  if (!dd) return fail(RHIP_E_INVALID, "null draw");
  if (!d->cur_fbo) return RHIP_OK; /* RendererSoft::draw returns when no framebuffer is bound (:137-139) */
  /* a deferred resolve waits only for a draw of rhip_draw on the same colour target that does not
   * sample that target */
  if (d->resolve.on) {
    bool keep = d->defer_resolve && d->fbo[d->cur_fbo].color == d->resolve.tex;
    for (int s = 0; keep && s < RHS_SLOT_COUNT; s++) keep = dd->sampler_tex[s] != d->resolve.tex;
    if (!keep && (rc = flushResolve(d))) return rc;
  }
  const int vp[4] = {d->vp_x, d->vp_y, d->vp_w, d->vp_h};
  DrawParams P;
  DrawShape sh;
  if ((rc = drawSetup(d, d->cur_fbo, vp, *dd, kTile, P, sh))) return rc;
  if (sh.empty) return RHIP_OK;
  if (sh.lines) return drawPointsLines(d, P, sh, no_alloc);
  const size_t need = carveArenas(P, sh, nullptr);
  if (no_alloc) {
    if (need > kScratchBytes)
      return fail(RHIP_E_UNSUPPORTED,
                  "rhip_draw_noalloc: this draw needs %zu bytes of scratch, more than the %zu the library holds "
                  "(%zu vertices, %zu triangles, %dx%d x%d samples); use rhip_draw",
                  need, kScratchBytes, sh.V, sh.T, P.fb_w, P.fb_h, P.samples);
    if (!d->scratch) return fail(RHIP_E_HIP, "rhip_draw_noalloc: no scratch address for this device");
    carveArenas(P, sh, d->scratch);
  } else {
    if ((rc = d->arena.ensure(need))) return rc;
    carveArenas(P, sh, d->arena.as<char>());
  }
  if (sh.T == 0) {
    /* no triangles: the reference still resolves a multisampled colour target (:161-163) */
    if (P.color && P.samples == 4) {
      resolveAfterDraw(d, P, (unsigned)std::min<size_t>(blocks((size_t)P.fb_w * P.fb_h, 256), 4096), d->stream);
      HIPCHK(hipGetLastError());
    }
    d->host_stats.draws++;
    return RHIP_OK;
  }
  const RhsStates &st = dd->states;
  const int ntiles = P.ntiles, nblk = P.nblk, fbw = P.fb_w, fbh = P.fb_h, samples = P.samples;
  const bool need_deriv = P.need_deriv != 0;
  const size_t V = sh.V;
  hipStream_t s = d->stream;

  { Timer tm(d, "k_vertex"); RHIP_LAUNCH(k_vertex, std::max(blocks(V, 256), 1u), 256, s, P); }
  { Timer tm(d, "k_setup"); RHIP_LAUNCH(k_setup, nblk, 256, s, P); }
  RHIP_LAUNCH(k_clip_fans, nblk, 256, s, P);
  { Timer tm(d, "k_bin_scan"); RHIP_LAUNCH(k_bin_scan, 1, 1024, s, P); }
  { Timer tm(d, "k_bin_emit"); RHIP_LAUNCH(k_bin_emit, std::min(nblk, 8192), 256, s, P); }
  HIPCHK(hipGetLastError());

  /* draws whose result follows from depths and order alone take the two-phase raster */
  if (sh.two_phase) {
    const int ncoarse = P.coarse_x * P.coarse_y;
    const size_t njobs = (size_t)fbw * fbh * samples;
    if (samples == 4)
      launchVis<4>(d, P, ntiles, ncoarse, s);
    else
      launchVis<1>(d, P, ntiles, ncoarse, s);
    if (P.color) {
      const int grid = (int)std::min<size_t>(blocks(njobs, 256), 4096);
      if (P.shading == RHS_SHADING_NEURAL_PBR) {
        Timer tm(d, samples == 4 ? "k_nn_shade_msaa4" : "k_nn_shade");
        const unsigned ig = (unsigned)std::min<size_t>((njobs + 255) / 256, 2048);
        if (samples == 4)
          RHIP_LAUNCH(k_nn_shade<4>, ig, 256, s, P, d->nn.packed, need_deriv ? 1 : 0);
        else
          RHIP_LAUNCH(k_nn_shade<1>, ig, 256, s, P, d->nn.packed, need_deriv ? 1 : 0);
      } else {
        Timer tm(d, samples == 4 ? "k_shade_msaa4" : "k_shade");
        if (samples == 4)
          launchShade<4>(P, grid, need_deriv ? 1 : 0, s);
        else
          launchShade<1>(P, grid, need_deriv ? 1 : 0, s);
      }
      if (d->nn.capturing && P.shading == RHS_SHADING_PBR) {
        Timer tm(d, "k_capture");
        d->nn.salt = d->nn.salt * 1664525u + 1013904223u;
        float *rec = d->nn.records.as<float>();
        if (samples == 4)
          RHIP_LAUNCH(k_capture<4>, grid, 256, s, P, need_deriv ? 1 : 0, rec, d->nn.rec_cap, d->nn.rec_count,
                      d->nn.shift, d->nn.salt, d->nn.material);
        else
          RHIP_LAUNCH(k_capture<1>, grid, 256, s, P, need_deriv ? 1 : 0, rec, d->nn.rec_cap, d->nn.rec_count,
                      d->nn.shift, d->nn.salt, d->nn.material);
      }
      if (samples == 4) resolveAfterDraw(d, P, (unsigned)std::min<size_t>(blocks((size_t)fbw * fbh, 256), 4096), s);
    }
  } else if (sh.listed) {
    if (samples == 4)
      launchListed<4>(d, P, need_deriv ? 1 : 0, s);
    else
      launchListed<1>(d, P, need_deriv ? 1 : 0, s);
  } else {
    /* NOTEQUAL with depth writes: early-Z must see every earlier write */
    const int flush_each = (P.depth && st.depthTest && st.depthMask && st.depthFunc == 5) ? 1 : 0;
    Timer tm(d, samples == 4 ? "k_raster_ord_msaa4" : "k_raster_ord");
    if (samples == 4)
      launchOrdered<4>(P, ntiles, need_deriv ? 1 : 0, flush_each, s);
    else
      launchOrdered<1>(P, ntiles, need_deriv ? 1 : 0, flush_each, s);
  }
  HIPCHK(hipGetLastError());
  d->host_stats.draws++;
  return RHIP_OK;
}

}  // namespace

extern "C" {

int rhip_draw(rhip_device d, const rhip_draw_desc *dd) {
  if (!d) return fail(RHIP_E_INVALID, "null device");
  d->defer_resolve = true;
  const int rc = drawOne(d, dd, false);
  d->defer_resolve = false;
  return rc;
}

int rhip_draw_noalloc(rhip_device d, const rhip_draw_desc *dd) {
  if (!d) return fail(RHIP_E_INVALID, "null device");
  d->no_alloc_call = true;
  const int rc = drawOne(d, dd, true);
  d->no_alloc_call = false;
  return rc;
}

size_t rhip_draw_noalloc_capacity(void) { return kScratchBytes; }

int rhip_profile_enable(rhip_device d, int on) {
  int rc = checkDev(d);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(d->stream));
  for (auto &t : d->timed) {
    d->event_pool.push_back(t.a);
    d->event_pool.push_back(t.b);
  }
  d->timed.clear();
  d->profiling = on != 0;
  return RHIP_OK;
}

int rhip_profile_query(rhip_device d, int index, char *name, int name_cap, double *total_ms, long long *count) {
  int rc = checkDev(d);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(d->stream));
  std::vector<std::pair<std::string, std::pair<double, long long>>> agg;
  for (auto &t : d->timed) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
    auto it = std::find_if(agg.begin(), agg.end(), [&](const auto &e) { return e.first == t.name; });
    if (it == agg.end()) agg.push_back({t.name, {ms, 1}});
    else it->second.first += ms, it->second.second++;
  }
  if (index < 0 || index >= (int)agg.size()) return fail(RHIP_E_INVALID, "profile entry %d of %zu", index, agg.size());
  if (name && name_cap > 0) snprintf(name, name_cap, "%s", agg[index].first.c_str());
  if (total_ms) *total_ms = agg[index].second.first;
  if (count) *count = agg[index].second.second;
  return RHIP_OK;
}

/* ------------------------------------------------------------------ learned shader */

static int nnEnsure(rhip_device d) {
  auto &n = d->nn;
  if (n.params) return RHIP_OK;
  const size_t pb = kNnParams * sizeof(float);
  HIPCHK(hipMalloc(&n.params, pb));
  HIPCHK(hipMalloc(&n.grad, pb));
  HIPCHK(hipMalloc(&n.m1, pb));
  HIPCHK(hipMalloc(&n.m2, pb));
  HIPCHK(hipMalloc(&n.packed, sizeof(NnPacked)));
  HIPCHK(hipMemsetAsync(n.packed, 0, sizeof(NnPacked), d->stream)); /* padding entries stay zero */
  HIPCHK(hipMalloc(&n.partial, (size_t)kNnMaxParts * kNnParams * sizeof(float)));
  HIPCHK(hipMalloc(&n.lossp, kNnMaxParts * sizeof(double)));
  HIPCHK(hipMalloc(&n.scratch, 8 * sizeof(double)));
  HIPCHK(hipMalloc(&n.hist, 2 * kNnHistBins * sizeof(unsigned int)));
  HIPCHK(hipMalloc(&n.rec_count, sizeof(unsigned long long)));
  HIPCHK(hipMemsetAsync(n.rec_count, 0, sizeof(unsigned long long), d->stream));
  return RHIP_OK;
}

/* parameters in, Adam state and gradient cleared, operands repacked */
static int nnLoad(rhip_device d, const float *w) {
  auto &n = d->nn;
  const size_t pb = kNnParams * sizeof(float);
  HIPCHK(hipMemcpyAsync(n.params, w, pb, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipMemsetAsync(n.m1, 0, pb, d->stream));
  HIPCHK(hipMemsetAsync(n.m2, 0, pb, d->stream));
  RHIP_LAUNCH(k_nn_update, kNnUpdateBlocks, 256, d->stream, n.params, n.partial, 0, n.lossp, (double *)nullptr,
              (float *)nullptr, n.m1, n.m2, n.packed, 0.f, 0.9f, 0.999f, 1e-8f, 0.f, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d->stream));
  n.step = 0;
  n.ready = true;
  return RHIP_OK;
}

int rhip_neural_param_count(void) { return kNnParams; }

int rhip_neural_init(rhip_device d, unsigned int seed) {
  int rc = checkDev(d);
  if (rc || (rc = nnEnsure(d))) return rc;
  /* uniform in +-sqrt(6 / fan_in) (fan_in 24, 64, 64), biases 0; a fixed LCG so runs repeat */
  std::vector<float> w(kNnParams, 0.f);
  uint64_t st = 0x9e3779b97f4a7c15ull ^ seed;
  auto uni = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int cnt, int fan) {
    const double a = std::sqrt(6.0 / fan);
    for (int i = 0; i < cnt; i++) w[off + i] = (float)((2 * uni() - 1) * a);
  };
  fill(kOffW1, kNnHid * kNnIn, kNnIn);
  fill(kOffW2, kNnHid * kNnHid, kNnHid);
  fill(kOffW3, kNnOut * kNnHid, kNnHid);
  return nnLoad(d, w.data());
}

int rhip_neural_set_weights(rhip_device d, const float *w, int n) {
  int rc = checkDev(d);
  if (rc || (rc = nnEnsure(d))) return rc;
  if (!w || n != kNnParams) return fail(RHIP_E_INVALID, "learned shader: %d parameters expected", kNnParams);
  return nnLoad(d, w);
}

int rhip_neural_get_weights(rhip_device d, float *w, int n) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!d->nn.ready) return fail(RHIP_E_INVALID, "learned shader: no network");
  if (!w || n != kNnParams) return fail(RHIP_E_INVALID, "learned shader: %d parameters expected", kNnParams);
  HIPCHK(hipMemcpyAsync(w, d->nn.params, kNnParams * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return RHIP_OK;
}

int rhip_neural_records(rhip_device d, long long capacity) {
  int rc = checkDev(d);
  if (rc || (rc = nnEnsure(d))) return rc;
  if (capacity < 0) return fail(RHIP_E_INVALID, "negative capacity");
  HIPCHK(hipStreamSynchronize(d->stream));
  if ((rc = d->nn.records.ensure((size_t)std::max<long long>(capacity, 1) * kRecFloats * 4))) return rc;
  d->nn.rec_cap = capacity;
  HIPCHK(hipMemsetAsync(d->nn.rec_count, 0, sizeof(unsigned long long), d->stream));
  return RHIP_OK;
}

int rhip_neural_capture(rhip_device d, int enable, int shift, int material) {
  int rc = checkDev(d);
  if (rc || (rc = nnEnsure(d))) return rc;
  if (enable && d->nn.rec_cap <= 0) return fail(RHIP_E_INVALID, "learned shader: no record buffer");
  if (shift < 0 || shift > 20) return fail(RHIP_E_INVALID, "sampling shift out of range");
  d->nn.capturing = enable != 0;
  d->nn.shift = shift;
  d->nn.material = material;
  return RHIP_OK;
}

int rhip_neural_material(rhip_device d, int material) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (material < 0 || material >= 64) return fail(RHIP_E_INVALID, "material %d (0 to 63)", material);
  d->nn.draw_material = material;
  return RHIP_OK;
}

int rhip_neural_fallback(rhip_device d, const int *materials, int n) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (n < 0 || (n > 0 && !materials)) return fail(RHIP_E_INVALID, "fallback list");
  uint64_t mask = 0;
  for (int i = 0; i < n; i++) {
    if (materials[i] < 0 || materials[i] >= 64) return fail(RHIP_E_INVALID, "material %d (0 to 63)", materials[i]);
    mask |= 1ull << materials[i];
  }
  d->nn.fallback = mask;
  return RHIP_OK;
}

int rhip_neural_record_count(rhip_device d, long long *count) {
  int rc = checkDev(d);
  if (rc || (rc = nnEnsure(d))) return rc;
  unsigned long long c = 0;
  HIPCHK(hipMemcpyAsync(&c, d->nn.rec_count, sizeof c, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  *count = std::min<long long>((long long)c, d->nn.rec_cap);
  return RHIP_OK;
}

int rhip_neural_read_records(rhip_device d, long long first, long long count, float *out) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (first < 0 || count < 0 || first + count > d->nn.rec_cap || !out)
    return fail(RHIP_E_INVALID, "record range outside the buffer");
  HIPCHK(hipMemcpyAsync(out, d->nn.records.as<float>() + (size_t)first * kRecFloats,
                        (size_t)count * kRecFloats * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return RHIP_OK;
}

int rhip_neural_write_records(rhip_device d, long long first, long long count, const float *in) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (first < 0 || count < 0 || first + count > d->nn.rec_cap || !in)
    return fail(RHIP_E_INVALID, "record range outside the buffer");
  HIPCHK(hipMemcpyAsync(d->nn.records.as<float>() + (size_t)first * kRecFloats, in, (size_t)count * kRecFloats * 4,
                        hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return RHIP_OK;
}

static int nnRange(rhip_device d, long long first, long long count) {
  if (!d->nn.ready) return fail(RHIP_E_INVALID, "learned shader: no network");
  if (first < 0 || count <= 0 || first + count > d->nn.rec_cap) return fail(RHIP_E_INVALID, "record range invalid");
  return RHIP_OK;
}

int rhip_neural_gradient(rhip_device d, long long first, int count, float *grad, double *loss) {
  int rc = checkDev(d);
  if (rc || (rc = nnRange(d, first, count))) return rc;
  auto &n = d->nn;
  const int parts = std::min<int>((count + 63) / 64, kNnMaxParts);
  RHIP_LAUNCH(k_nn_grad, (unsigned)parts, kNnGradThreads, d->stream, n.packed, n.records.as<float>(), first, 0ll, count, 0u,
              n.partial, n.lossp);
  RHIP_LAUNCH(k_nn_update, kNnUpdateBlocks, 256, d->stream, n.params, n.partial, parts, n.lossp, n.scratch, n.grad,
              n.m1, n.m2, n.packed, 0.f, 0.f, 0.f, 0.f, 0.f, 0);
  HIPCHK(hipGetLastError());
  if (grad) HIPCHK(hipMemcpyAsync(grad, n.grad, kNnParams * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  if (loss) HIPCHK(hipMemcpyAsync(loss, n.scratch, sizeof(double), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return RHIP_OK;
}

int rhip_neural_train(rhip_device d, long long first, long long count, int steps, int batch, float lr_start,
                      float lr_end, unsigned int seed, double *losses) {
  int rc = checkDev(d);
  if (rc || (rc = nnRange(d, first, count))) return rc;
  if (steps < 0 || batch < 64 || count >= (1ll << 32)) return fail(RHIP_E_INVALID, "training arguments invalid");
  auto &n = d->nn;
  std::vector<double> lh;
  double *dl = nullptr;
  if (losses && steps > 0) {
    HIPCHK(hipMalloc(&dl, steps * sizeof(double)));
    HIPCHK(hipMemsetAsync(dl, 0, steps * sizeof(double), d->stream));
  }
  const int grid = std::min<int>((batch + 63) / 64, 512);
  for (int i = 0; i < steps; i++) {
    /* cosine schedule from lr_start to lr_end */
    const double t = steps > 1 ? (double)i / (steps - 1) : 0.0;
    const float lr = (float)(lr_end + 0.5 * (lr_start - lr_end) * (1.0 + std::cos(3.14159265358979 * t)));
    const uint32_t sd = seed * 0x9e3779b9u + (uint32_t)(n.step + 1) * 0x85ebca6bu;
    RHIP_LAUNCH(k_nn_grad, (unsigned)grid, kNnGradThreads, d->stream, n.packed, n.records.as<float>(), first, count, batch, sd,
                n.partial, n.lossp);
    n.step++;
    RHIP_LAUNCH(k_nn_update, kNnUpdateBlocks, 256, d->stream, n.params, n.partial, grid, n.lossp,
                dl ? dl + i : (double *)nullptr, (float *)nullptr, n.m1, n.m2, n.packed, lr, 0.9f, 0.999f, 1e-8f,
                1.f / batch, n.step);
  }
  HIPCHK(hipGetLastError());
  if (dl) {
    lh.resize(steps);
    HIPCHK(hipMemcpyAsync(lh.data(), dl, steps * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  }
  HIPCHK(hipStreamSynchronize(d->stream));
  if (dl) {
    hipFree(dl);
    for (int i = 0; i < steps; i++) losses[i] = lh[i] / ((double)batch * kNnOut);
  }
  return RHIP_OK;
}

static void histStats(const std::vector<unsigned int> &h, double sum, double mx, double n, rhip_nn_error *e) {
  e->samples = (long long)n;
  e->mean = n > 0 ? sum / (n * kNnOut) : 0.0;
  e->max = mx;
  /* p99 over channel values: upper edge of the bin holding the 99th percentile */
  const double total = n * kNnOut, want = 0.99 * total;
  double acc = 0;
  e->p99 = 0;
  for (int b = 0; b < kNnHistBins; b++) {
    acc += h[b];
    if (acc >= want) {
      e->p99 = b == kNnHistBins - 1 ? mx : (b + 1) / 16.0;
      break;
    }
  }
}

int rhip_neural_evaluate(rhip_device d, long long first, long long count, int material, rhip_nn_error *nn_err,
                         rhip_nn_error *plain_err, float *outputs) {
  int rc = checkDev(d);
  if (rc || (rc = nnRange(d, first, count))) return rc;
  if (count > INT32_MAX) return fail(RHIP_E_INVALID, "too many records");
  auto &n = d->nn;
  float *dout = nullptr;
  if (outputs) HIPCHK(hipMalloc(&dout, (size_t)count * 3 * sizeof(float)));
  HIPCHK(hipMemsetAsync(n.hist, 0, 2 * kNnHistBins * sizeof(unsigned int), d->stream));
  HIPCHK(hipMemsetAsync(n.scratch, 0, 8 * sizeof(double), d->stream));
  RHIP_LAUNCH(k_nn_eval, (unsigned)std::min<long long>((count + 255) / 256, 2048), 256, d->stream, n.packed,
              n.records.as<float>(), first, (int)count, material, n.hist, n.hist + kNnHistBins, n.scratch, dout);
  HIPCHK(hipGetLastError());
  std::vector<unsigned int> h(2 * kNnHistBins);
  double sums[8];
  HIPCHK(hipMemcpyAsync(h.data(), n.hist, h.size() * 4, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipMemcpyAsync(sums, n.scratch, sizeof sums, hipMemcpyDeviceToHost, d->stream));
  if (dout) HIPCHK(hipMemcpyAsync(outputs, dout, (size_t)count * 3 * sizeof(float), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  if (dout) hipFree(dout);
  std::vector<unsigned int> hn(h.begin(), h.begin() + kNnHistBins), hp(h.begin() + kNnHistBins, h.end());
  if (nn_err) histStats(hn, sums[0], sums[2], sums[4], nn_err);
  if (plain_err) histStats(hp, sums[1], sums[3], sums[4], plain_err);
  return RHIP_OK;
}

int rhip_stats_get(rhip_device d, rhip_stats *out) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!out) return fail(RHIP_E_INVALID, "null output");
  unsigned long long all[kCounterSlots * kCounterStride], c[8] = {};
  HIPCHK(hipMemcpyAsync(all, d->counters, sizeof all, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  for (int k = 0; k < kCounterSlots; k++)
    for (int i = 0; i < 8; i++) c[i] += all[k * kCounterStride + i];
  *out = d->host_stats;
  out->fs_invocations = (long long)c[0];
  out->samples_written = (long long)c[1];
  out->tile_pairs = (long long)c[3];
  out->triangles_binned = (long long)c[4];
  out->clipped_triangles = (long long)c[5];
  out->hiz_culled = (long long)c[6];
  return RHIP_OK;
}

int rhip_stats_reset(rhip_device d) {
  int rc = checkDev(d);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(d->counters, 0, kCounterSlots * kCounterStride * sizeof(unsigned long long), d->stream));
  d->host_stats = rhip_stats();
  return RHIP_OK;
}

}  // extern "C"
