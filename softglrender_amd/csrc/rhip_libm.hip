/*
 * rhip_libm.hip -- host side of rhip_libm.h: builds, once per process and per
 * function, the table of inputs where the host libm's result differs from the
 * device approximation's rounding, and uploads the tables, the albedo gamma
 * table and the output-gamma byte thresholds to each device.
 *
 * Building a table evaluates the host libm and the approximation on every
 * float of the function's range (up to 2^30 inputs), split over the host's
 * threads: about a second per function on 16 threads.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rhip_host.h"
#include "rhip_libm.h"

namespace rhip {
namespace {

/* The host libm, called through pointers the compiler cannot see through
 * (no constant folding, no pow(x, 5) -> x*x*x*x*x rewriting). */
float (*volatile g_log2f)(float) = log2f;
float (*volatile g_exp2f)(float) = exp2f;
float (*volatile g_powf)(float, float) = powf;

float hostLibm(int fn, float x) {
  switch (fn) {
    case kLmLog2:
      return g_log2f(x);
    case kLmExp2Neg:
      return g_exp2f(-x);
    case kLmPow5:
      return g_powf(x, 5.0f);
    case kLmPow128:
      return g_powf(x, 128.0f);
    default:
      return g_powf(x, 1.0f / 2.2f);
  }
}

struct HostTab {
  bool built = false;
  bool ok = false; /* every input within one ulp of the approximation's rounding */
  std::string why;
  uint32_t lo = 0, hi = 0;
  double gate = -1.0;
  std::vector<uint32_t> bucket, ent;
};

std::mutex g_mu;
HostTab g_tab[kLmCount];

void buildTab(int fn, HostTab &T) {
  lmRange(fn, &T.lo, &T.hi);
  const uint64_t n = (uint64_t)T.hi - T.lo + 1;
  int nt = (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, 16));
  struct Part {
    std::vector<uint32_t> xs; /* input bits | direction in bit 31 (inputs are positive) */
    double gate = -1.0;
    uint64_t bad = 0;
    uint32_t bad_x = 0;
    int first_byte = -1, last_byte = -1; /* kLmPowGamma: output bytes at the part's ends */
    bool monotone = true;
  };
  std::vector<Part> parts(nt);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      Part &p = parts[t];
      const uint64_t a = T.lo + n * t / nt, b = T.lo + n * (t + 1) / nt;
      for (uint64_t i = a; i < b; i++) {
        const float x = lmAsFloat((uint32_t)i);
        const double d = lmApprox(fn, x);
        const float h = hostLibm(fn, x);
        const uint32_t ab = lmBits((float)d), hb = lmBits(h);
        if (fn == kLmPowGamma) {
          const int byte = (int)(uint8_t)(int32_t)(std::min(std::max(h, 0.f), 1.f) * 255.f);
          if (p.first_byte < 0) p.first_byte = byte;
          if (byte < p.last_byte) p.monotone = false;
          p.last_byte = byte;
        }
        if (ab == hb) continue;
        if (hb == ab + 1 || hb == ab - 1) {
          p.xs.push_back((uint32_t)i | (hb == ab + 1 ? 0x80000000u : 0u));
          p.gate = std::max(p.gate, lmMidDist(d));
        } else if (!p.bad++) {
          p.bad_x = (uint32_t)i;
        }
      }
    });
  for (auto &t : th) t.join();
  T.ok = true;
  std::vector<uint32_t> xs;
  int prev_byte = -1;
  for (auto &p : parts) {
    if (fn == kLmPowGamma && (!p.monotone || p.first_byte < prev_byte) && T.ok) {
      /* the output-gamma byte thresholds (gammaTablesHost) need a monotonic byte */
      T.ok = false;
      T.why = "host powf(x, 1/2.2) gives a byte that is not monotonic in x on (0, 1]";
    }
    if (p.last_byte >= 0) prev_byte = p.last_byte;
    if (p.bad && T.ok) {
      T.ok = false;
      char buf[160];
      snprintf(buf, sizeof buf, "host libm function %d differs by more than one ulp at x = %a (and %llu more inputs)",
               fn, (double)lmAsFloat(p.bad_x), (unsigned long long)(p.bad - 1));
      T.why = buf;
    }
    T.gate = std::max(T.gate, p.gate);
    xs.insert(xs.end(), p.xs.begin(), p.xs.end()); /* parts are in input order */
  }
  const uint32_t b0 = T.lo >> 16, nb = (T.hi >> 16) - b0 + 1;
  T.bucket.assign(nb + 1, 0);
  T.ent.resize(xs.size());
  for (size_t i = 0; i < xs.size(); i++) {
    const uint32_t x = xs[i] & 0x7fffffffu;
    T.bucket[(x >> 16) - b0 + 1]++;
    T.ent[i] = (x & 0xffffu) | ((xs[i] >> 31) << 16);
  }
  for (uint32_t k = 0; k < nb; k++) T.bucket[k + 1] += T.bucket[k];
  if (T.ent.empty()) T.ent.push_back(0); /* keep a valid pointer; the gate (-1) never consults it */
  T.built = true;
}

/* The byte a PBR colour channel c becomes without an emissive map:
 * u8(clamp(powf(c, 1/2.2), 0, 1) * 255), truncating (PbrSoft.h:313, RendererSoft.cpp:949-954). */
int gammaOutByteHost(float c) {
  float g = g_powf(c, 1.0f / 2.2f);
  g = g < 0.f ? 0.f : g; /* glm::clamp keeps NaN as the reference's comparisons do */
  g = 1.f < g ? 1.f : g;
  return (int)(uint8_t)(int32_t)(g * 255.f);
}

}  // namespace

int libmEnsure(rhip_device d, uint32_t fns) {
  if (!d->libm) {
    HIPCHK(hipMalloc(&d->libm, sizeof(LibmTables)));
    LibmTables z;
    memset(&z, 0, sizeof z);
    for (int f = 0; f < kLmCount; f++) z.t[f].gate = -1.0;
    HIPCHK(hipMemcpy(d->libm, &z, sizeof z, hipMemcpyHostToDevice));
  }
  for (int f = 0; f < kLmCount; f++) {
    if (!(fns & (1u << f)) || (d->libm_ready & (1u << f))) continue;
    HostTab *T = &g_tab[f];
    {
      std::lock_guard<std::mutex> lk(g_mu);
      if (!T->built) buildTab(f, *T);
    }
    if (!T->ok) return fail(RHIP_E_UNSUPPORTED, "%s", T->why.c_str());
    uint32_t *bk = nullptr, *en = nullptr;
    HIPCHK(hipMalloc(&bk, T->bucket.size() * 4));
    HIPCHK(hipMalloc(&en, T->ent.size() * 4));
    d->libm_bufs.push_back(bk);
    d->libm_bufs.push_back(en);
    HIPCHK(hipMemcpy(bk, T->bucket.data(), T->bucket.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(en, T->ent.data(), T->ent.size() * 4, hipMemcpyHostToDevice));
    LibmTab dt;
    dt.lo = T->lo;
    dt.hi = T->hi;
    dt.gate = T->gate;
    dt.bucket = bk;
    dt.ent = en;
    HIPCHK(hipMemcpy(&d->libm->t[f], &dt, sizeof dt, hipMemcpyHostToDevice));
    d->libm_ready |= 1u << f;
  }
  return RHIP_OK;
}

int libmHostTables(int fn, LibmTables *L) {
  std::lock_guard<std::mutex> lk(g_mu);
  HostTab &T = g_tab[fn];
  if (!T.built) buildTab(fn, T);
  if (!T.ok) return fail(RHIP_E_UNSUPPORTED, "%s", T.why.c_str());
  memset(L, 0, sizeof *L);
  for (int f = 0; f < kLmCount; f++) L->t[f].gate = -1.0;
  L->t[fn].lo = T.lo;
  L->t[fn].hi = T.hi;
  L->t[fn].gate = T.gate;
  L->t[fn].bucket = T.bucket.data();
  L->t[fn].ent = T.ent.data();
  return RHIP_OK;
}

int libmTableStats(int fn, long long *entries, double *gate) {
  if (fn < 0 || fn >= kLmCount) return fail(RHIP_E_INVALID, "no libm table %d", fn);
  std::lock_guard<std::mutex> lk(g_mu);
  HostTab &T = g_tab[fn];
  if (!T.built) buildTab(fn, T);
  if (!T.ok) return fail(RHIP_E_UNSUPPORTED, "%s", T.why.c_str());
  if (entries) *entries = (long long)T.bucket.back();
  if (gate) *gate = T.gate;
  return RHIP_OK;
}

void gammaTablesHost(float *t) {
  for (int k = 0; k < 256; k++) t[k] = g_powf((float)k / 255.f, 2.2f);
  for (int k = 0; k < 256; k++) {
    uint32_t lo = 0, hi = 0x7f800000u; /* +inf gives byte 255 */
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (gammaOutByteHost(lmAsFloat(mid)) >= k)
        hi = mid;
      else
        lo = mid + 1;
    }
    t[256 + k] = lmAsFloat(lo);
  }
}

float hostPowf(float x, float y) { return g_powf(x, y); }

float (*volatile g_sinf)(float) = sinf;
float (*volatile g_cosf)(float) = cosf;

void iblTablesHost(IblTables *t) {
  memset(t, 0, sizeof *t);
  /* IBLIrradianceSoft.h:84-85: float loop variables stepped by 0.025f while below 2 PI and PI / 2 */
  const float pi = 3.14159265359f, delta = 0.025f;
  int n = 0;
  for (float phi = 0.0f; phi < 2.0f * pi && n < 256; phi += delta, n++) {
    t->sphi[n] = g_sinf(phi);
    t->cphi[n] = g_cosf(phi);
  }
  t->nphi = n;
  n = 0;
  for (float theta = 0.0f; theta < 0.5f * pi && n < 64; theta += delta, n++) {
    t->stheta[n] = g_sinf(theta);
    t->ctheta[n] = g_cosf(theta);
  }
  t->ntheta = n;
  /* IBLPrefilterSoft.h:102-112: phi = 2 PI * (i / 1024) */
  for (int i = 0; i < 1024; i++) {
    const float xi = (float)i / (float)1024u;
    const float phi = 2.0f * pi * xi;
    t->hcos[i] = g_cosf(phi);
    t->hsin[i] = g_sinf(phi);
  }
}

}  // namespace rhip

namespace {

/* Function fn at v as the shaders call it (exp2 of -v). */
__host__ __device__ inline float evalFn(const rhip::LibmTables *L, int fn, float v) {
  switch (fn) {
    case rhip::kLmLog2:
      return rhip::hlog2f(L, v);
    case rhip::kLmExp2Neg:
      return rhip::hexp2f(L, -v);
    case rhip::kLmPow5:
      return rhip::hpowi<5>(L, v);
    case rhip::kLmPow128:
      return rhip::hpowi<128>(L, v);
    default:
      return rhip::hpowGamma(L, v);
  }
}

__global__ void k_libm_eval(const rhip::LibmTables *L, int fn, const float *x, long long n, float *out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = fn == rhip::kLmAtan2 ? rhip::fdAtan2f(x[2 * i], x[2 * i + 1]) : evalFn(L, fn, x[i]);
}

}  // namespace

extern "C" {

int rhip_libm_info(int fn, long long *entries, double *gate) { return rhip::libmTableStats(fn, entries, gate); }

int rhip_libm_eval_host(int fn, const float *x, long long n, float *out) {
  using namespace rhip;
  if (fn < 0 || fn > kLmAtan2 || n < 0 || (n && (!x || !out))) return fail(RHIP_E_INVALID, "libm eval arguments");
  if (fn == kLmAtan2) {
    for (long long i = 0; i < n; i++) out[i] = fdAtan2f(x[2 * i], x[2 * i + 1]);
    return RHIP_OK;
  }
  LibmTables L;
  int rc = libmHostTables(fn, &L);
  if (rc) return rc;
  for (long long i = 0; i < n; i++) out[i] = evalFn(&L, fn, x[i]);
  return RHIP_OK;
}

int rhip_libm_eval(rhip_device d, int fn, const float *x, long long n, float *out) {
  using namespace rhip;
  int rc = checkDev(d);
  if (rc) return rc;
  if (fn < 0 || fn > kLmAtan2 || n < 0 || (n && (!x || !out))) return fail(RHIP_E_INVALID, "libm eval arguments");
  if (!n) return RHIP_OK;
  if (fn < kLmCount && (rc = libmEnsure(d, 1u << fn))) return rc;
  const long long nin = fn == kLmAtan2 ? 2 * n : n;
  float *dx = nullptr, *dout = nullptr;
  HIPCHK(hipMalloc(&dx, nin * 4));
  if (hipMalloc(&dout, n * 4) != hipSuccess) {
    hipFree(dx);
    return fail(RHIP_E_OOM, "libm eval buffers");
  }
  hipError_t e = hipMemcpyAsync(dx, x, nin * 4, hipMemcpyHostToDevice, d->stream);
  if (e == hipSuccess) {
    const long long blocksNeeded = (n + 255) / 256;
    k_libm_eval<<<(unsigned)std::min<long long>(blocksNeeded, 8192), 256, 0, d->stream>>>(d->libm, fn, dx, n, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, d->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
  hipFree(dx);
  hipFree(dout);
  if (e != hipSuccess) return fail(RHIP_E_HIP, "libm eval: %s", hipGetErrorString(e));
  return RHIP_OK;
}

}  // extern "C"
