/*
 * rhip_batch.hip -- batched views (rhip_batch_* in include/rhip.h).
 *
 * A batch holds one DrawParams per view, all views' arenas in one allocation,
 * a block table per stage and a captured HIP graph: the fill kernel (bins,
 * clears, counters), one launch per geometry stage over all views, the coarse
 * hierarchical Z, the persistent raster over (view, tile) items, the shading
 * walk and, multisampled, the resolve. Replays take one hipGraphLaunch.
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <set>
#include <string>

#include "rhip_batch.h"
#include "rhip_host.h"

namespace rhip {

/* device scratch of a batch: t0 | view_time[nv] | queue | chunks_done[nv] | chunk_first[nv + 1] */
struct BatchMisc {
  size_t view_time, queue, chunks_done, chunk_first, bytes;
  explicit BatchMisc(int nv) {
    view_time = 8;
    queue = view_time + 8 * (size_t)nv;
    chunks_done = queue + 8;
    chunk_first = chunks_done + 4 * (size_t)nv;
    bytes = chunk_first + 4 * ((size_t)nv + 1);
  }
};

struct Batch {
  int nv = 0;
  std::vector<rhip_batch_view> views; /* uniform pointers into `ubytes` */
  std::vector<std::vector<uint8_t>> ubytes;
  rhip_batch_config cfg{};
  int samples = 1, func = 1, shading = 0;
  std::string signature; /* workload key of the tuning cache */
  DevBuf arena, params, tabs, fills, misc;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double clock_khz = 0;
  /* resources the graph reads or writes, with their storage when it was built */
  std::vector<std::pair<int, const void *>> tex_refs, vao_refs;
  void dropGraph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~Batch() { dropGraph(); }
};

namespace {

std::mutex g_tuned_mu;
std::map<std::string, rhip_batch_config> g_tuned; /* workload signature -> fastest configuration */

int visModeOf(int func) { return func == 1 ? VIS_LESS : func == 3 ? VIS_LEQUAL : func == 4 ? VIS_GREATER : VIS_GEQUAL; }

const VisKernel *findVis(int samples, int func, const rhip_batch_config &c) {
  int n = 0;
  const VisKernel *k = samples == 4 ? visKernels4(&n) : visKernels1(&n);
  const int mode = visModeOf(func);
  for (int i = 0; i < n; i++)
    if (k[i].mode == mode && k[i].tile == c.tile && k[i].waves == c.waves) return &k[i];
  return nullptr;
}

/* every configuration the raster is compiled for at this sample count */
std::vector<rhip_batch_config> candidates(int samples) {
  int n = 0;
  const VisKernel *k = samples == 4 ? visKernels4(&n) : visKernels1(&n);
  std::vector<rhip_batch_config> out;
  for (int i = 0; i < n; i++)
    if (k[i].mode == VIS_LESS) out.push_back({k[i].tile, k[i].waves});
  return out;
}

int cuCount(int dev) {
  int n = 0;
  hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
  return n > 0 ? n : 1;
}

template <int S, int SHADING>
void shadeLaunch(const DrawParams *Pv, const int *chunk_first, int nv, int *done, unsigned long long *vt, int grid,
                 hipStream_t st) {
  kb_shade<S, SHADING><<<grid, 256, 0, st>>>(Pv, chunk_first, nv, done, vt);
}

/* The shading walk for one sample count and shading model. */
struct ShadeKernel {
  void (*launch)(const DrawParams *, const int *, int, int *, unsigned long long *, int, hipStream_t);
  const void *fn;
};

template <int S>
ShadeKernel shadeKernelS(int shading) {
  switch (shading) {
    case RHS_SHADING_BASIC: return {shadeLaunch<S, RHS_SHADING_BASIC>, (const void *)kb_shade<S, RHS_SHADING_BASIC>};
    case RHS_SHADING_BLINN_PHONG:
      return {shadeLaunch<S, RHS_SHADING_BLINN_PHONG>, (const void *)kb_shade<S, RHS_SHADING_BLINN_PHONG>};
    case RHS_SHADING_BLINN_PHONG_DIR:
      return {shadeLaunch<S, RHS_SHADING_BLINN_PHONG_DIR>, (const void *)kb_shade<S, RHS_SHADING_BLINN_PHONG_DIR>};
    case RHS_SHADING_PBR: return {shadeLaunch<S, RHS_SHADING_PBR>, (const void *)kb_shade<S, RHS_SHADING_PBR>};
    default: return {shadeLaunch<S, RHS_SHADING_VERTEX_COLOR>, (const void *)kb_shade<S, RHS_SHADING_VERTEX_COLOR>};
  }
}

ShadeKernel shadeKernel(int samples, int shading) {
  return samples == 4 ? shadeKernelS<4>(shading) : shadeKernelS<1>(shading);
}

uint32_t clearColorWord(const rhip_clear_states &cs) {
  /* RendererSoft::beginRenderPass (RendererSoft.cpp:61-90): u8 truncation of clear * 255 */
  auto u8 = [](float v) { return (uint32_t)(uint8_t)(int32_t)(v * 255); };
  return u8(cs.clear_color[0]) | (u8(cs.clear_color[1]) << 8) | (u8(cs.clear_color[2]) << 16) |
         (u8(cs.clear_color[3]) << 24);
}

/* appends table entries (key << kViewShift | b) for b in [0, count) */
void addEntries(std::vector<uint32_t> &tab, uint32_t key, size_t count) {
  for (size_t b = 0; b < count; b++) tab.push_back((key << kViewShift) | (uint32_t)b);
}

/* Lays out the batch for configuration `c` and captures its graph. */
int buildBatch(rhip_device d, Batch &b, const rhip_batch_config &c) {
  const VisKernel *vk = findVis(b.samples, b.func, c);
  if (!vk)
    return fail(RHIP_E_UNSUPPORTED, "no raster compiled for tile %d with %d wavefronts at %d samples", c.tile,
                c.waves, b.samples);
  HIPCHK(hipStreamSynchronize(d->stream)); /* no replay still reads what is rebuilt */
  b.dropGraph();
  const int nv = b.nv;
  std::vector<DrawParams> P(nv);
  std::vector<DrawShape> sh(nv);
  std::vector<size_t> off(nv);
  size_t total = 0;
  int rc;
  for (int v = 0; v < nv; v++) {
    const rhip_batch_view &bv = b.views[v];
    if ((rc = drawSetup(d, bv.fbo, bv.viewport, bv.draw, c.tile, P[v], sh[v]))) return rc;
    if (sh[v].empty || !sh[v].two_phase)
      return fail(RHIP_E_UNSUPPORTED, "batch view %d: not a draw the two-phase raster takes", v);
    if (P[v].ntiles >= (1 << kViewShift) || P[v].nblk >= (1 << kViewShift) ||
        blocks(sh[v].V, 256) >= (1u << kViewShift))
      return fail(RHIP_E_UNSUPPORTED, "batch view %d: too large for a batch", v);
    off[v] = total;
    total += carveArenas(P[v], sh[v], nullptr);
  }
  if ((rc = b.arena.ensure(total))) return rc;
  for (int v = 0; v < nv; v++) carveArenas(P[v], sh[v], b.arena.as<char>() + off[v]);
  b.tex_refs.clear();
  b.vao_refs.clear();
  for (int v = 0; v < nv; v++) {
    const rhip_draw_desc &dd = b.views[v].draw;
    const Fbo &f = d->fbo[b.views[v].fbo];
    for (int t : {f.color, f.depth})
      if (t) b.tex_refs.push_back({t, getTex(d, t)->data});
    for (int s = 0; s < RHS_SLOT_COUNT; s++)
      if (dd.sampler_tex[s]) b.tex_refs.push_back({dd.sampler_tex[s], getTex(d, dd.sampler_tex[s])->data});
    b.vao_refs.push_back({dd.vao, d->vao[dd.vao]->verts});
  }

  /* fill ranges: per view its bins and the clears, then the batch counters */
  const BatchMisc M(nv);
  if ((rc = b.misc.ensure(M.bytes))) return rc;
  char *misc = b.misc.as<char>();
  std::vector<FillRange> fills;
  for (int v = 0; v < nv; v++) {
    fills.push_back({(uint32_t *)P[v].bin_count, (uint32_t)P[v].nbins, 0u});
    const rhip_clear_states &cs = b.views[v].clear;
    const size_t px = (size_t)P[v].fb_w * P[v].fb_h * P[v].samples;
    if (cs.color_flag && P[v].color) fills.push_back({P[v].color, (uint32_t)px, clearColorWord(cs)});
    if (cs.depth_flag && P[v].depth) {
      uint32_t bits;
      memcpy(&bits, &cs.clear_depth, 4);
      fills.push_back({(uint32_t *)P[v].depth, (uint32_t)px, bits});
    }
  }
  fills.push_back({(uint32_t *)(misc + M.view_time), (uint32_t)((M.chunk_first - M.view_time) / 4), 0u});

  /* block tables */
  std::vector<uint32_t> tab;
  size_t o_fill = tab.size();
  for (size_t r = 0; r < fills.size(); r++) addEntries(tab, (uint32_t)r, (fills[r].n + kFillBlock - 1) / kFillBlock);
  const size_t n_fill = tab.size() - o_fill, o_vert = tab.size();
  for (int v = 0; v < nv; v++) addEntries(tab, v, blocks(sh[v].V, 256));
  const size_t n_vert = tab.size() - o_vert, o_tri = tab.size();
  for (int v = 0; v < nv; v++) addEntries(tab, v, P[v].nblk);
  const size_t n_tri = tab.size() - o_tri, o_coarse = tab.size();
  for (int v = 0; v < nv; v++) addEntries(tab, v, (size_t)P[v].coarse_x * P[v].coarse_y);
  const size_t n_coarse = tab.size() - o_coarse, o_tiles = tab.size();
  for (int v = 0; v < nv; v++) addEntries(tab, v, P[v].ntiles);
  const size_t n_tiles = tab.size() - o_tiles, o_res = tab.size();
  if (b.samples == 4)
    for (int v = 0; v < nv; v++)
      if (P[v].color) addEntries(tab, v, blocks((size_t)P[v].fb_w * P[v].fb_h, 256));
  const size_t n_res = tab.size() - o_res;
  if (n_fill >= (1u << 31) || n_tiles >= (1u << 31)) return fail(RHIP_E_UNSUPPORTED, "batch too large");

  if ((rc = b.params.ensure(nv * sizeof(DrawParams))) || (rc = b.tabs.ensure(tab.size() * 4)) ||
      (rc = b.fills.ensure(fills.size() * sizeof(FillRange))))
    return rc;
  HIPCHK(hipMemcpy(b.params.p, P.data(), nv * sizeof(DrawParams), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.tabs.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.fills.p, fills.data(), fills.size() * sizeof(FillRange), hipMemcpyHostToDevice));

  const DrawParams *Pv = b.params.as<DrawParams>();
  const uint32_t *T = b.tabs.as<uint32_t>();
  unsigned long long *t0 = (unsigned long long *)misc, *vt = (unsigned long long *)(misc + M.view_time);
  int *queue = (int *)(misc + M.queue), *done = (int *)(misc + M.chunks_done), *cf = (int *)(misc + M.chunk_first);
  const int cus = cuCount(d->hip_device);
  int occ = 0, socc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, vk->fn, 64 * c.waves, 0) != hipSuccess || occ < 1) occ = 1;
  const ShadeKernel sk = shadeKernel(b.samples, b.shading);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&socc, sk.fn, 256, 0) != hipSuccess || socc < 1) socc = 1;
  const bool mx = b.func == 4 || b.func == 6;
  bool any_color = false;
  for (int v = 0; v < nv; v++) any_color = any_color || P[v].color;

  hipStream_t s = d->stream;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  kb_fill<<<(unsigned)n_fill, 256, 0, s>>>(b.fills.as<FillRange>(), T + o_fill, t0);
  if (n_vert) kb_vertex<<<(unsigned)n_vert, 256, 0, s>>>(Pv, T + o_vert);
  if (n_tri) {
    kb_setup<<<(unsigned)n_tri, 256, 0, s>>>(Pv, T + o_tri);
    kb_clip<<<(unsigned)n_tri, 256, 0, s>>>(Pv, T + o_tri);
  }
  kb_scan_fans<<<nv, 1024, 0, s>>>(Pv);
  if (n_tri) kb_clip_fans<<<(unsigned)n_tri, 256, 0, s>>>(Pv, T + o_tri);
  kb_bin_scan<<<nv, 1024, 0, s>>>(Pv);
  if (n_tri) kb_bin_emit<<<(unsigned)n_tri, 256, 0, s>>>(Pv, T + o_tri);
  if (b.samples == 4) {
    if (mx)
      kb_hiz_coarse<4, true><<<(unsigned)n_coarse, 256, 0, s>>>(Pv, T + o_coarse);
    else
      kb_hiz_coarse<4, false><<<(unsigned)n_coarse, 256, 0, s>>>(Pv, T + o_coarse);
  } else {
    if (mx)
      kb_hiz_coarse<1, true><<<(unsigned)n_coarse, 256, 0, s>>>(Pv, T + o_coarse);
    else
      kb_hiz_coarse<1, false><<<(unsigned)n_coarse, 256, 0, s>>>(Pv, T + o_coarse);
  }
  vk->launch(Pv, T + o_tiles, (int)n_tiles, queue, occ * cus, s);
  kb_job_chunks<<<1, 1024, 0, s>>>(Pv, nv, cf, vt);
  if (any_color) sk.launch(Pv, cf, nv, done, vt, socc * cus, s);
  if (n_res) kb_resolve<<<(unsigned)n_res, 256, 0, s>>>(Pv, T + o_res);
  const hipError_t le = hipGetLastError();
  const hipError_t ce = hipStreamEndCapture(s, &b.graph);
  if (le != hipSuccess) return fail(RHIP_E_HIP, "batch launch: %s", hipGetErrorString(le));
  if (ce != hipSuccess) return fail(RHIP_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
  HIPCHK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
  b.cfg = c;
  return RHIP_OK;
}

/* RHIP_OK if every resource the graph uses still has the storage it was built with. */
int checkRefs(rhip_device d, const Batch &b) {
  for (const auto &r : b.tex_refs) {
    Texture *t = getTex(d, r.first);
    if (!t || t->data != r.second) return fail(RHIP_E_INVALID, "batch: texture %d was destroyed", r.first);
  }
  for (const auto &r : b.vao_refs) {
    auto it = d->vao.find(r.first);
    if (it == d->vao.end() || it->second->verts != r.second)
      return fail(RHIP_E_INVALID, "batch: vertex array %d was destroyed", r.first);
  }
  return RHIP_OK;
}

Batch *getBatch(rhip_device d, int id) {
  auto it = d->batches.find(id);
  return it == d->batches.end() ? nullptr : it->second.get();
}

}  // namespace
}  // namespace rhip

using namespace rhip;

extern "C" {

int rhip_batch_create(rhip_device d, const rhip_batch_view *views, int n, const rhip_batch_config *config,
                      int *out) {
  int rc = checkDev(d);
  if (rc) return rc;
  if (!views || !out || n <= 0) return fail(RHIP_E_INVALID, "batch: no views");
  if (n > kMaxBatchViews) return fail(RHIP_E_UNSUPPORTED, "batch: at most %d views", kMaxBatchViews);
  auto b = std::make_shared<Batch>();
  b->nv = n;
  b->views.assign(views, views + n);
  b->ubytes.resize((size_t)n * 3);
  std::set<int> targets;
  char sig[160];
  for (int v = 0; v < n; v++) {
    rhip_batch_view &bv = b->views[v];
    const void **u[3] = {&bv.draw.uniform_model, &bv.draw.uniform_scene, &bv.draw.uniform_material};
    const size_t sz[3] = {256, 64, 48};
    for (int k = 0; k < 3; k++) {
      if (!*u[k]) continue;
      auto &dst = b->ubytes[(size_t)v * 3 + k];
      dst.assign((const uint8_t *)*u[k], (const uint8_t *)*u[k] + sz[k]);
      *u[k] = dst.data();
    }
    auto fit = d->fbo.find(bv.fbo);
    if (fit == d->fbo.end()) return fail(RHIP_E_INVALID, "batch view %d: invalid framebuffer %d", v, bv.fbo);
    for (int t : {fit->second.color, fit->second.depth})
      if (t && !targets.insert(t).second)
        return fail(RHIP_E_INVALID, "batch view %d: attachment %d is written by another view", v, t);
    auto pit = d->prog.find(bv.draw.program);
    if (pit == d->prog.end()) return fail(RHIP_E_INVALID, "batch view %d: invalid program", v);
    DrawParams P;
    DrawShape sh;
    if ((rc = drawSetup(d, bv.fbo, bv.viewport, bv.draw, kTile, P, sh))) return rc;
    if (sh.empty || !sh.two_phase || P.shading == RHS_SHADING_NEURAL_PBR)
      return fail(RHIP_E_UNSUPPORTED,
                  "batch view %d: views must be opaque draws that test and write depth with "
                  "LESS/LEQUAL/GREATER/GEQUAL, with a plain shader",
                  v);
    if (v == 0) {
      b->samples = P.samples;
      b->func = P.st.depthFunc;
      b->shading = P.shading;
    } else if (P.samples != b->samples || P.st.depthFunc != b->func || P.shading != b->shading) {
      return fail(RHIP_E_UNSUPPORTED, "batch view %d: views differ in sample count, depth test or shading", v);
    }
    snprintf(sig, sizeof sig, "%dx%d:%zu;", P.fb_w, P.fb_h, sh.T);
    b->signature += sig;
  }
  snprintf(sig, sizeof sig, "s%d f%d m%d", b->samples, b->func, b->shading);
  b->signature += sig;
  int khz = 0;
  HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, d->hip_device));
  b->clock_khz = khz > 0 ? khz : 100000;
  rhip_batch_config c = {kTile, 4};
  if (config) {
    c = *config;
  } else {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned.find(b->signature);
    if (it != g_tuned.end()) c = it->second;
  }
  if ((rc = buildBatch(d, *b, c))) return rc;
  const int id = d->next_id++;
  d->batches[id] = b;
  *out = id;
  return RHIP_OK;
}

int rhip_batch_run(rhip_device d, int id) {
  int rc = checkDev(d);
  if (rc) return rc;
  Batch *b = getBatch(d, id);
  if (!b || !b->exec) return fail(RHIP_E_INVALID, "invalid batch %d", id);
  if ((rc = checkRefs(d, *b))) return rc;
  HIPCHK(hipGraphLaunch(b->exec, d->stream));
  d->host_stats.draws += b->nv;
  return RHIP_OK;
}

int rhip_batch_tune(rhip_device d, int id, int reps, rhip_batch_timing *timings, int cap, int *count,
                    rhip_batch_config *best) {
  int rc = checkDev(d);
  if (rc) return rc;
  Batch *b = getBatch(d, id);
  if (!b) return fail(RHIP_E_INVALID, "invalid batch %d", id);
  if (reps < 1) return fail(RHIP_E_INVALID, "reps must be positive");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best_ms = 1e300;
  rhip_batch_config best_c = b->cfg;
  int k = 0;
  for (const rhip_batch_config &c : candidates(b->samples)) {
    if ((rc = buildBatch(d, *b, c))) break;
    if ((rc = rhip_batch_run(d, id))) break; /* warm-up */
    hipEventRecord(e0, d->stream);
    for (int r = 0; r < reps && !rc; r++) rc = rhip_batch_run(d, id);
    hipEventRecord(e1, d->stream);
    if (rc) break;
    if (hipEventSynchronize(e1) != hipSuccess) {
      rc = fail(RHIP_E_HIP, "batch tuning: replay failed");
      break;
    }
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / reps;
    if (timings && k < cap) timings[k] = {c, per};
    k++;
    if (per < best_ms) {
      best_ms = per;
      best_c = c;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (count) *count = k;
  if (rc) return rc;
  if ((rc = buildBatch(d, *b, best_c))) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned[b->signature] = best_c;
  }
  if (best) *best = best_c;
  return RHIP_OK;
}

int rhip_batch_get_config(rhip_device d, int id, rhip_batch_config *config) {
  if (!d || !config) return fail(RHIP_E_INVALID, "null argument");
  Batch *b = getBatch(d, id);
  if (!b) return fail(RHIP_E_INVALID, "invalid batch %d", id);
  *config = b->cfg;
  return RHIP_OK;
}

int rhip_batch_set_config(rhip_device d, int id, const rhip_batch_config *config) {
  int rc = checkDev(d);
  if (rc) return rc;
  Batch *b = getBatch(d, id);
  if (!b || !config) return fail(RHIP_E_INVALID, "invalid batch %d", id);
  return buildBatch(d, *b, *config);
}

int rhip_batch_view_times(rhip_device d, int id, double *ms, int n) {
  int rc = checkDev(d);
  if (rc) return rc;
  Batch *b = getBatch(d, id);
  if (!b) return fail(RHIP_E_INVALID, "invalid batch %d", id);
  if (!ms || n < b->nv) return fail(RHIP_E_INVALID, "need room for %d views", b->nv);
  HIPCHK(hipStreamSynchronize(d->stream));
  std::vector<unsigned long long> t(1 + (size_t)b->nv);
  HIPCHK(hipMemcpy(t.data(), b->misc.p, t.size() * 8, hipMemcpyDeviceToHost));
  for (int v = 0; v < b->nv; v++) ms[v] = (double)(long long)(t[1 + v] - t[0]) / b->clock_khz;
  return RHIP_OK;
}

int rhip_batch_destroy(rhip_device d, int id) {
  int rc = checkDev(d);
  if (rc) return rc;
  auto it = d->batches.find(id);
  if (it == d->batches.end()) return fail(RHIP_E_INVALID, "invalid batch %d", id);
  HIPCHK(hipStreamSynchronize(d->stream));
  d->batches.erase(it);
  return RHIP_OK;
}

}  // extern "C"
