/*
 * rhip.h -- C-ABI of librhip, the MI355X (gfx950) rasteriser behind
 * SoftGLRender's abstract Renderer interface.
 *
 * Each entry point replaces one call of the reference's interface
 * (src/Render/Renderer.h:24-59 and the resource classes it returns); the
 * citation is on each declaration. Plain pointers, sizes and POD structs only.
 *
 * Conventions
 *  - Every function returns RHIP_OK (0) or a negative RHIP_E_* code; the
 *    message of the last failure on the calling thread is rhip_last_error().
 *    (The reference logs LOGE and returns nullptr/false: RendererSoft.cpp:28-34,
 *    TextureSoft.h:116-124.)
 *  - Handles are positive ints, valid within the device that made them.
 *  - Work is queued on one HIP stream per device. Host buffers passed in are
 *    copied before the call returns (the reference copies vertex, index and
 *    uniform bytes too: VertexSoft.h:16-27, UniformSoft.h:31-37).
 *  - rhip_draw snapshots the uniform blocks it is given, so callers may
 *    rewrite them before the next draw (Viewer.cpp:684-716 does).
 *  - rhip_texture_read and rhip_wait_idle wait for queued work.
 */
#ifndef RHIP_H
#define RHIP_H

#include <stddef.h>
#include <stdint.h>

#include "rhip_stream.h" /* RhsStates, shading / define / slot enums */

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RHIP_OK = 0,
  RHIP_E_INVALID = -1,     /* bad argument or handle */
  RHIP_E_HIP = -2,         /* a HIP runtime call failed */
  RHIP_E_UNSUPPORTED = -3, /* a valid request this build does not implement */
  RHIP_E_OOM = -4,         /* device allocation failed */
  RHIP_E_DEVICE = -5,      /* a kernel reported an internal overflow */
};

/* Texel storage order of a sampled texture. The reference picks one for every
 * buffer at compile time (SOFTGL_TEXTURE_TILED / _MORTON, Buffer.h:215-227);
 * here it is per texture. Sampled values do not depend on it. */
enum { RHIP_LAYOUT_LINEAR = 0, RHIP_LAYOUT_TILED = 1, RHIP_LAYOUT_MORTON = 2 };

typedef struct rhip_device_s *rhip_device;

/* TextureDesc + SamplerDesc (Texture.h:17-46), enum values as the reference's. */
typedef struct rhip_texture_desc {
  int width, height;
  int type;        /* TextureType: 0 = 2D, 1 = CUBE */
  int format;      /* TextureFormat: 0 = RGBA8, 1 = FLOAT32 */
  uint32_t usage;  /* TextureUsage bits */
  int use_mipmaps; /* bool */
  int multi_sample;/* bool: 4 samples per texel */
  int layout;      /* RHIP_LAYOUT_* */
} rhip_texture_desc;

typedef struct rhip_sampler_desc {
  int filter_min, filter_mag; /* FilterMode */
  int wrap_s, wrap_t, wrap_r; /* WrapMode */
  int border_color;           /* BorderColor */
} rhip_sampler_desc;

/* ClearStates (RenderStates.h:65-70). */
typedef struct rhip_clear_states {
  int depth_flag, color_flag;
  float clear_color[4];
  float clear_depth;
} rhip_clear_states;

/* One draw: everything RendererSoft::draw reads from the bound state
 * (vao, program, resources, pipeline states; RendererSoft.cpp:115-164). */
typedef struct rhip_draw_desc {
  int vao;
  int program;
  RhsStates states;                        /* RenderStates (RenderStates.h:52-63) */
  const void *uniform_model;               /* UniformsModel, 256 B (Material.h:70-76) */
  const void *uniform_scene;               /* UniformsScene, 64 B */
  const void *uniform_material;            /* UniformsMaterial, 48 B */
  int sampler_tex[RHS_SLOT_COUNT];         /* texture per RhsSlot, 0 = unbound */
} rhip_draw_desc;

/* Per-frame counters (for Gfrag/s and diagnostics), summed since the last reset. */
typedef struct rhip_stats {
  long long fs_invocations; /* fragment-shader invocations (pixel granularity) */
  long long samples_written;
  long long triangles_binned;
  long long tile_pairs;
  long long clipped_triangles;
  long long draws;
  long long hiz_culled; /* (tile, triangle) pairs skipped by the per-tile hierarchical-Z test */
} rhip_stats;

/* Error of a shader against the FP64 evaluation of the plain PBR shader, per colour channel, in units of 1/255. */
typedef struct rhip_nn_error {
  double mean, p99, max;
  long long samples; /* fragments */
} rhip_nn_error;

/* ---------------------------------------------------------------- device (Renderer::create/destroy) */
int rhip_device_create(int hip_device, rhip_device *out);              /* Renderer::create  (Renderer.h:27) */
int rhip_device_destroy(rhip_device dev);                              /* Renderer::destroy (Renderer.h:28) */
const char *rhip_last_error(void);

/* ---------------------------------------------------------------- textures */
int rhip_texture_create(rhip_device dev, const rhip_texture_desc *desc, int *out_tex);       /* createTexture (Renderer.h:34) */
int rhip_texture_set_sampler(rhip_device dev, int tex, const rhip_sampler_desc *s);          /* Texture::setSamplerDesc (Texture.h:97) */
int rhip_texture_init(rhip_device dev, int tex);                                             /* Texture::initImageData (Texture.h:98) */
/* Texture::setImageData (Texture.h:99-100): level 0 of each layer, row-major linear,
 * 4 bytes per texel; mip levels are generated when the texture uses mipmaps. */
int rhip_texture_upload(rhip_device dev, int tex, int layers, const void *data, size_t bytes);
/* Writes one mip level of one layer (row-major linear); no reference counterpart:
 * the reference fills IBL mip chains by rendering into them (Environment.cpp:84-103). */
int rhip_texture_upload_level(rhip_device dev, int tex, int layer, int level, const void *data, size_t bytes);
/* Reads one level back, row-major linear. which = 0: the single-sample / resolved
 * image; which = 1: the raw 4x samples (4 consecutive values per pixel).
 * Texture::dumpImage (Texture.h:101) and ViewerSoftware::swapBuffer's read (ViewerSoftware.h:30-44). */
int rhip_texture_read(rhip_device dev, int tex, int layer, int level, int which, void *out, size_t bytes);
int rhip_texture_destroy(rhip_device dev, int tex);

/* ---------------------------------------------------------------- vertex arrays */
/* createVertexArrayObject (Renderer.h:37): float attributes, `stride` bytes per
 * vertex, int32 indices (read as points, lines or triangles by the draw's
 * primitive type; a trailing partial primitive is ignored, as the reference's
 * assembly does: RendererSoft.cpp:410-437). */
int rhip_vao_create(rhip_device dev, const void *vertices, int stride, int nverts, const int32_t *indices,
                    int nindices, int *out_vao);
int rhip_vao_update(rhip_device dev, int vao, const void *vertices, size_t bytes); /* VertexArrayObject::updateVertexData (Vertex.h:18) */
int rhip_vao_destroy(rhip_device dev, int vao);

/* ---------------------------------------------------------------- framebuffers (Framebuffer.h:20-91) */
int rhip_fbo_create(rhip_device dev, int *out_fbo);                                       /* createFrameBuffer (Renderer.h:31) */
int rhip_fbo_attach_color(rhip_device dev, int fbo, int tex, int layer, int level);       /* setColorAttachment (Framebuffer.h:27-39) */
int rhip_fbo_attach_depth(rhip_device dev, int fbo, int tex);                             /* setDepthAttachment (Framebuffer.h:41-46) */
int rhip_fbo_destroy(rhip_device dev, int fbo);

/* ---------------------------------------------------------------- programs */
/* createShaderProgram + addDefine(s) (Renderer.h:40, ShaderProgram.h:19-25): a program
 * is a shading model (RhsShading) plus a define bitmask (RhsDefine). */
int rhip_program_create(rhip_device dev, int shading, uint32_t defines, int *out_program);

/* ---------------------------------------------------------------- pipeline */
int rhip_begin_render_pass(rhip_device dev, int fbo, const rhip_clear_states *clear); /* beginRenderPass (Renderer.h:50) */
int rhip_set_viewport(rhip_device dev, int x, int y, int width, int height);          /* setViewPort (Renderer.h:51) */
/* set*(...) + draw (Renderer.h:52-56). Filled triangles draw with every
 * shading model; points, lines and triangles in point or line mode
 * (RendererSoft.cpp:560-718) with the base colour and vertex colour models
 * (others: RHIP_E_UNSUPPORTED), at point sizes and line widths up to 1024;
 * such a draw waits once for its fragment count before it sizes the list. */
int rhip_draw(rhip_device dev, const rhip_draw_desc *draw);
/* rhip_draw without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The draw's arena lives in a
 * static device buffer of rhip_draw_noalloc_capacity() bytes (64 MiB) that
 * every call on one GPU shares and that keeps the last call's contents: calls
 * that overlap on one GPU (two devices' streams, or another thread) race. A
 * draw whose worst-case arena exceeds it fails with RHIP_E_UNSUPPORTED (about
 * 4k triangles at 640x480 with 4x MSAA fit), and so does a draw of points or
 * lines (their fragment list is sized at draw time). Results equal rhip_draw's. */
int rhip_draw_noalloc(rhip_device dev, const rhip_draw_desc *draw);
size_t rhip_draw_noalloc_capacity(void);
int rhip_end_render_pass(rhip_device dev);                                            /* endRenderPass (Renderer.h:57) */
/* waitIdle (Renderer.h:58). Also returns RHIP_E_DEVICE, naming the stage, when a
 * kernel dropped work since the last wait or readback (a clipped polygon over
 * 9 vertices); rhip_texture_read reports the same. Reporting clears it. */
int rhip_wait_idle(rhip_device dev);
/* Test hook, no reference counterpart: a kernel marks device error `stage`
 * (0 = clipping overflow) as a dropping kernel would, so the reporting path
 * can be tested without provoking a real overflow. */
int rhip_debug_raise_device_error(rhip_device dev, int stage);

/* ---------------------------------------------------------------- measurement */
/* ---------------------------------------------------------------- learned PBR shader
 * No counterpart in SoftGLRender: a 24-64-64-3 MLP on BF16 MFMA that replaces
 * the image-based PBR shader (PbrSoft.h:225-322, point light off) for programs
 * created with RHS_SHADING_NEURAL_PBR. Records are fragments captured from
 * plain PBR draws (shading inputs, plain GPU colour, FP64 plain colour). */
int rhip_neural_param_count(void);                                      /* 5955 */
int rhip_neural_init(rhip_device dev, unsigned int seed);               /* fresh network */
int rhip_neural_set_weights(rhip_device dev, const float *w, int n);    /* W1 (64x24), b1, W2 (64x64), b2, W3 (3x64), b3 */
int rhip_neural_get_weights(rhip_device dev, float *w, int n);
int rhip_neural_records(rhip_device dev, long long capacity);          /* (re)allocates and empties the record buffer */
int rhip_neural_capture(rhip_device dev, int enable, int shift, int material); /* 1 in 2^shift fragments of PBR draws */
int rhip_neural_material(rhip_device dev, int material);               /* material (0..63) of the draws that follow; default 0 */
int rhip_neural_fallback(rhip_device dev, const int *materials, int n); /* these materials' learned-shader draws use the plain PBR shader */
int rhip_neural_record_count(rhip_device dev, long long *count);
int rhip_neural_read_records(rhip_device dev, long long first, long long count, float *out);   /* 32 floats each */
int rhip_neural_write_records(rhip_device dev, long long first, long long count, const float *in);
int rhip_neural_gradient(rhip_device dev, long long first, int count, float *grad, double *loss); /* summed over records */
/* training repeats bit for bit for the same weights, records and arguments */
int rhip_neural_train(rhip_device dev, long long first, long long count, int steps, int batch, float lr_start,
                      float lr_end, unsigned int seed, double *losses /* steps entries, may be NULL */);
int rhip_neural_evaluate(rhip_device dev, long long first, long long count, int material, rhip_nn_error *network,
                         rhip_nn_error *plain, float *outputs /* count x 3, may be NULL */);

/* ---------------------------------------------------------------- batched views
 * No counterpart in SoftGLRender, whose Viewer renders one view per frame
 * (Viewer.cpp:100-168): many independent views, each one render pass of one
 * draw into its own framebuffer, run as one captured HIP graph in which every
 * stage is one launch over all views and the raster is a persistent kernel
 * over (view, tile) work items. A view's result is the same as
 * beginRenderPass(fbo, clear) + setViewPort + draw + endRenderPass on it. */
typedef struct rhip_batch_view {
  int fbo;                 /* framebuffer; no two views of a batch share an attachment */
  int viewport[4];         /* x, y, width, height */
  rhip_clear_states clear; /* applied first, as beginRenderPass does */
  rhip_draw_desc draw;     /* uniform blocks are copied at creation */
} rhip_batch_view;

/* Screen tile (= fine bin) edge in pixels (8, 16, 32 or 64) and wavefronts
 * per tile workgroup of the raster. */
typedef struct rhip_batch_config {
  int tile, waves;
} rhip_batch_config;

typedef struct rhip_batch_timing {
  rhip_batch_config config;
  double ms; /* mean time of one batch */
} rhip_batch_timing;

/* Views must all take the two-phase raster (opaque, depth test and write with
 * LESS/LEQUAL/GREATER/GEQUAL) with one sample count and one shading model, and
 * the resources they use must outlive the batch. config NULL: the
 * configuration an earlier rhip_batch_tune found for the same workload in this
 * process, else tile 16 with 4 wavefronts. */
int rhip_batch_create(rhip_device dev, const rhip_batch_view *views, int n, const rhip_batch_config *config,
                      int *out_batch);
int rhip_batch_run(rhip_device dev, int batch); /* queues one replay of the batch's graph */
/* Autotuner: rebuilds the batch for every valid configuration, times `reps`
 * replays of each, keeps the fastest (also for later batches of the same
 * workload) and reports every timing (up to `cap`; `*count` gets the number). */
int rhip_batch_tune(rhip_device dev, int batch, int reps, rhip_batch_timing *timings, int cap, int *count,
                    rhip_batch_config *best);
int rhip_batch_get_config(rhip_device dev, int batch, rhip_batch_config *config);
int rhip_batch_set_config(rhip_device dev, int batch, const rhip_batch_config *config); /* rebuilds the graph */
/* Waits, then gives each view's completion time in the last replay, in
 * milliseconds from the replay's first kernel (device clock). */
int rhip_batch_view_times(rhip_device dev, int batch, double *ms, int n);
int rhip_batch_destroy(rhip_device dev, int batch);

int rhip_stats_get(rhip_device dev, rhip_stats *out); /* waits for queued work */
int rhip_stats_reset(rhip_device dev);
/* The host libm's single-precision functions on the device (no reference
 * counterpart; for tests). The shaders call glibc's log2f, exp2f and powf
 * (ShaderSoft.h:132, PbrSoft.h:201-235,313, BlinnPhongSoft.h:227); the device
 * reproduces them bit for bit from a table of the inputs where its own
 * rounding differs, built from the host libm at rhip_device_create.
 * fn: 0 log2f(x), 1 exp2f(-x), 2 powf(x, 5), 3 powf(x, 128), 4 powf(x, 1/2.2f),
 * 5 atan2f(y, x) (SkyboxSoft.h:84; no table: the algorithm glibc uses,
 * step for step), where x holds n (y, x) pairs, 2n floats.
 * rhip_libm_info (no device needed) builds fn's table on the host and gives
 * its listed inputs and gate (fn 0-4); rhip_libm_eval evaluates fn on the device. */
int rhip_libm_info(int fn, long long *entries, double *gate);
int rhip_libm_eval(rhip_device dev, int fn, const float *x, long long n, float *out);
/* The same evaluation on the host, from the host's copy of the table. */
int rhip_libm_eval_host(int fn, const float *x, long long n, float *out);
/* Hierarchical Z on (1, the default) or off (0) for later draws of the device.
 * No reference counterpart: the reference tests every triangle at every sample;
 * with it off the two-phase raster does too (no depth-bound culling), so tests
 * can check that culling never changes an output. */
int rhip_set_hierarchical_z(rhip_device dev, int enabled);
/* The HIP stream the device queues on (hipStream_t), for event timing. */
void *rhip_device_stream(rhip_device dev);
/* Per-kernel timing with HIP events recorded on the device's stream around every
 * launch while enabled; enabling (or disabling) clears what was recorded. */
int rhip_profile_enable(rhip_device dev, int on);
/* Entry `index` of the per-kernel totals since profiling was enabled: kernel
 * name, summed milliseconds and launch count; RHIP_E_INVALID past the end. */
int rhip_profile_query(rhip_device dev, int index, char *name, int name_cap, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif
#endif /* RHIP_H */
