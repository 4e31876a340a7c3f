/*
 * rhip_host.h -- host-side state of librhip shared by its translation units:
 * the device object, its resources and per-draw arenas.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "rhip.h"
#include "rhip_internal.h"

namespace rhip {

struct NnPacked;
struct Batch;

/* Records the message of a failure for rhip_last_error and returns `code`. */
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(RHIP_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

/* A grow-only device buffer. */
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return RHIP_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(bytes, (size_t)4096);
    want = want + std::min(want / 4, (size_t)1 << 30); /* room to grow, at most 1 GiB */
    if (hipMalloc(&p, want) != hipSuccess) return fail(RHIP_E_OOM, "device allocation of %zu bytes failed", want);
    cap = want;
    return RHIP_OK;
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

struct Texture {
  rhip_texture_desc desc{};
  rhip_sampler_desc samp{};
  TexDesc hd{};               /* host copy of the descriptor */
  TexDesc *dd = nullptr;      /* device copy */
  uint8_t *data = nullptr;    /* all levels of all layers */
  size_t bytes = 0;
  uint8_t *resolved = nullptr; /* resolve target of a multisampled colour texture */
  bool initialised = false;
  ~Texture() {
    if (dd) hipFree(dd);
    if (data) hipFree(data);
    if (resolved) hipFree(resolved);
  }
};

struct Vao {
  float *verts = nullptr;
  int32_t *idx = nullptr;
  int stride = 0, nverts = 0, nidx = 0;
  ~Vao() {
    if (verts) hipFree(verts);
    if (idx) hipFree(idx);
  }
};

struct Fbo {
  int color = 0, color_layer = 0, color_level = 0, depth = 0;
};

struct Program {
  int shading = 0;
  uint32_t defines = 0;
};

}  // namespace rhip

struct rhip_device_s {
  int hip_device = 0;
  hipStream_t stream = nullptr;
  int next_id = 1;
  std::map<int, std::unique_ptr<rhip::Texture>> tex;
  std::map<int, std::unique_ptr<rhip::Vao>> vao;
  std::map<int, rhip::Fbo> fbo;
  std::map<int, rhip::Program> prog;
  int cur_fbo = 0;
  int vp_x = 0, vp_y = 0, vp_w = 0, vp_h = 0;
  /* per-draw arenas */
  rhip::DevBuf arena;
  rhip::DevBuf lines; /* fragment lists of point and line draws, sized once their count is known */
  /* device address of this GPU's copy of the library's static draw scratch
   * (rhip_draw_noalloc), from hipGetSymbolAddress at creation */
  char *scratch = nullptr;
  bool no_alloc_call = false; /* inside rhip_draw_noalloc: no event timers */
  bool hiz_off = false;       /* rhip_set_hierarchical_z(dev, 0): no depth-bound culling */
  unsigned long long *counters = nullptr; /* device */
  uint32_t *errors = nullptr;             /* device: kErrSlots words set by kernels that drop work */
  float *gamma_lut = nullptr;             /* device: gamma tables (rhip_libm.hip) */
  rhip::LibmTables *libm = nullptr;       /* device: host libm tables (rhip_libm.h) */
  rhip::IblTables *ibl = nullptr;         /* device: sin / cos tables of the IBL shaders */
  uint32_t libm_ready = 0;                /* functions whose table is on this device */
  std::vector<void *> libm_bufs;
  /* optional per-kernel timing: an event pair around every launch */
  bool profiling = false;
  struct Timed {
    const char *name;
    hipEvent_t a, b;
  };
  std::vector<Timed> timed;
  std::vector<hipEvent_t> event_pool;
  std::map<int, std::shared_ptr<rhip::Batch>> batches; /* captured multi-view batches */
  rhip_stats host_stats{};
  /* learned shader: FP32 parameters, Adam moments, gradient, packed bf16 operands, captured records */
  struct Neural {
    float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    rhip::NnPacked *packed = nullptr;
    double *scratch = nullptr; /* losses and evaluation sums */
    float *partial = nullptr;  /* per-block gradient sums of a step (kNnMaxParts rows) */
    double *lossp = nullptr;   /* per-block squared errors of a step */
    unsigned int *hist = nullptr;
    int step = 0;
    bool ready = false;
    rhip::DevBuf records;
    long long rec_cap = 0;
    unsigned long long *rec_count = nullptr;
    bool capturing = false;
    int shift = 0, material = 0;
    int draw_material = 0; /* material of the draws that follow (rhip_neural_material) */
    uint64_t fallback = 0; /* materials whose learned-shader draws use the plain shader */
    uint32_t salt = 0;
  } nn;
  /* The multisample resolve of the last draw on a colour target, not yet launched: a draw's resolve is
   * overwritten by the next draw's on the same target unless something reads the resolved image in
   * between, so rhip_draw leaves it here; every other entry point and a draw that reads it or targets
   * another texture launch it first (flushResolve). */
  struct PendingResolve {
    bool on = false;
    int tex = 0; /* colour attachment (texture id) */
    unsigned grid = 0;
    rhip::DrawParams P;
  } resolve;
  bool defer_resolve = false; /* set while rhip_draw runs */
};

namespace rhip {

int flushResolve(rhip_device d);

/* The device made current, and a deferred multisample resolve launched (every entry point but rhip_draw). */
inline int checkDev(rhip_device d) {
  if (!d) return fail(RHIP_E_INVALID, "null device");
  hipError_t e = hipSetDevice(d->hip_device);
  if (e != hipSuccess) return fail(RHIP_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  return d->resolve.on ? flushResolve(d) : RHIP_OK;
}

inline Texture *getTex(rhip_device d, int id) {
  auto it = d->tex.find(id);
  return it == d->tex.end() ? nullptr : it->second.get();
}

inline unsigned blocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

/* What the host knows of a draw besides its DrawParams. */
struct DrawShape {
  size_t V = 0, T = 0; /* vertices, triangles */
  int stride = 0;      /* vertex bytes */
  bool empty = false;  /* no attachment: nothing to do */
  bool two_phase = false; /* visibility first, then shading (opaque draws that write depth) */
  bool listed = false;    /* ordered draws that write no depth: list, shade, then blend in order */
  bool lines = false;     /* points, lines, or triangles in point or line mode (rhip_lines.h) */
  long long units = 0;    /* lines: the points and lines the reference rasterizes */
};

/* Validates a draw into framebuffer `fbo` through viewport vp = (x, y, w, h)
 * with screen tiles of `tile` pixels, and fills P except for its arenas. */
int drawSetup(rhip_device d, int fbo, const int vp[4], const rhip_draw_desc &dd, int tile, DrawParams &P,
              DrawShape &sh);

/* Points P's arenas into memory from `base` on (sizes only when base is null);
 * returns the bytes they take. */
size_t carveArenas(DrawParams &P, const DrawShape &sh, char *base);

void decodeUniforms(const rhip_draw_desc &dd, Uniforms &u);

/* Puts the host libm tables of the functions in `fns` (bit per LibmFn) on the
 * device, building them first if this process has not (rhip_libm.hip). */
int libmEnsure(rhip_device d, uint32_t fns);
/* L with function fn's host-side table (the others empty; builds it if needed). */
int libmHostTables(int fn, LibmTables *L);
/* Listed inputs and gate of function fn's table (builds it if needed). */
int libmTableStats(int fn, long long *entries, double *gate);
/* t[0, 256): powf(k / 255, 2.2); t[256 + k]: smallest c >= 0 whose output byte is at least k. */
void gammaTablesHost(float *t);
float hostPowf(float x, float y);
/* The IBL shaders' loop tables from the host libm's sinf / cosf. */
void iblTablesHost(IblTables *t);

}  // namespace rhip
