/*
 * rhip_libm.h -- the host libm's single-precision log2f, exp2f and powf, as
 * the shaders call them, reproduced bit for bit on the device.
 *
 * The reference shades with glibc's powf, log2f and exp2f
 * (ShaderSoft.h:132 for the LOD, PbrSoft.h:201,206 pow(x, 5), :215
 * exp2(-9.28 NdotV), :235 pow(albedo, 2.2), :313 pow(colour, 1/2.2),
 * BlinnPhongSoft.h:227 pow(x, 128)). Those are accurate to about 0.5-0.8 ulp
 * but not correctly rounded, so a device implementation of its own differs from
 * them on a few inputs in a million, and a differing LOD or gamma moves a
 * truncated byte.
 *
 * Method. For each function a deterministic approximation in double precision
 * made of IEEE basic operations only (+, -, *, /, rint, ldexp: correctly
 * rounded, so host and device compute the same double), rounded to float, is
 * within one ulp of the host libm's result everywhere in the function's input
 * range. At first use the library evaluates both on every float of that range
 * on the host (rhip_libm.hip) and keeps the inputs where they differ, bucketed
 * by the top 16 bits of the input, with the direction of the one-ulp step. The
 * device evaluates the approximation and consults the table only when the
 * double is within `gate` ulps of a rounding midpoint, `gate` being the largest
 * such distance of any listed input (measured by the same enumeration), so
 * every listed input is found. Inputs outside the range fall back to the
 * correctly rounded value (documented per function below).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhip {

enum LibmFn {
  kLmLog2 = 0,   /* log2f(x), x in [1, FLT_MAX] (the LOD: below 1 the LOD is clamped to 0 anyway) */
  kLmExp2Neg,    /* exp2f(-u), u in [2^-27, 16] (EnvBRDFApprox, u = 9.28 NdotV; below: both are 1) */
  kLmPow5,       /* powf(x, 5), x in [2^-31, 1 + 2^-10] (Fresnel; below: both are 0) */
  kLmPow128,     /* powf(x, 128), x in [0.4, 1 + 2^-10] (Blinn-Phong specular; below: both are 0) */
  kLmPowGamma,   /* powf(x, 1/2.2f), x in [2^-149, 1] (output gamma; from 1 on the result is >= 1 and clamped) */
  kLmCount
};

/* One function's table (device memory): inputs whose host result is one ulp
 * away from the approximation's rounding. */
struct LibmTab {
  uint32_t lo, hi;        /* input bit patterns covered, inclusive (positive floats) */
  double gate;            /* midpoint distance (ulps) at or below which an input may be listed */
  const uint32_t *bucket; /* ((hi >> 16) - (lo >> 16) + 2) offsets into ent, by input bits >> 16 */
  const uint32_t *ent;    /* low 16 input bits | 1 << 16 when the host result is one ulp above */
};

struct LibmTables {
  LibmTab t[kLmCount];
};

/* ------------------------------------------------------------- approximations (host and device) */

__host__ __device__ inline float lmAsFloat(uint32_t u) {
  union {
    uint32_t u;
    float f;
  } v;
  v.u = u;
  return v.f;
}
__host__ __device__ inline uint32_t lmBits(float f) {
  union {
    float f;
    uint32_t u;
  } v;
  v.f = f;
  return v.u;
}

/* log2 of a positive finite float: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, series to s^21.
 * Subnormals are normalised with integer operations (no float arithmetic on
 * them, whatever the denormal mode). */
__host__ __device__ inline double lmLog2(float x) {
  uint32_t u = lmBits(x);
  int e = (int)((u >> 23) & 255u) - 127;
  if (e == -127) {
    const int sh = __builtin_clz(u & 0x7fffffu) - 8; /* moves the leading bit to bit 23 */
    u = (u & 0x7fffffu) << sh;
    e = -126 - sh;
  }
  double m = (double)lmAsFloat((u & 0x7fffffu) | 0x3f800000u);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0), z = s * s;
  double p = 1.0 / 21;
  p = p * z + 1.0 / 19;
  p = p * z + 1.0 / 17;
  p = p * z + 1.0 / 15;
  p = p * z + 1.0 / 13;
  p = p * z + 1.0 / 11;
  p = p * z + 1.0 / 9;
  p = p * z + 1.0 / 7;
  p = p * z + 1.0 / 5;
  p = p * z + 1.0 / 3;
  p = p * z + 1.0;
  return (double)e + (2.0 * s * p) * 1.4426950408889634074;
}

/* 2^t for |t| < 1000: t = k + r with k = rint(t), |r| <= 1/2, e^(r ln 2) to the 13th power. */
__host__ __device__ inline double lmExp2(double t) {
  const double k = rint(t), u = (t - k) * 0.69314718055994530942;
  double p = 1.0 / 6227020800.0;
  p = p * u + 1.0 / 479001600.0;
  p = p * u + 1.0 / 39916800.0;
  p = p * u + 1.0 / 3628800.0;
  p = p * u + 1.0 / 362880.0;
  p = p * u + 1.0 / 40320.0;
  p = p * u + 1.0 / 5040.0;
  p = p * u + 1.0 / 720.0;
  p = p * u + 1.0 / 120.0;
  p = p * u + 1.0 / 24.0;
  p = p * u + 1.0 / 6.0;
  p = p * u + 0.5;
  p = p * u + 1.0;
  p = p * u + 1.0;
  return ldexp(p, (int)k);
}

/* The approximation of function fn at x, in double (x inside the function's range). */
__host__ __device__ inline double lmApprox(int fn, float x) {
  switch (fn) {
    case kLmLog2:
      return lmLog2(x);
    case kLmExp2Neg:
      return lmExp2(-(double)x);
    case kLmPow5: {
      const double a = x, a2 = a * a;
      return a2 * a2 * a;
    }
    case kLmPow128: {
      double r = x;
      for (int i = 0; i < 7; i++) r = r * r;
      return r;
    }
    default:
      return lmExp2(lmLog2(x) * (double)(1.0f / 2.2f));
  }
}

/* Distance in ulps of d from the nearest midpoint between floats, measured
 * against the neighbour on d's side of its rounding (for d >= 0). */
__host__ __device__ inline double lmMidDist(double d) {
  const float f = (float)d;
  const double fl = (double)f, err = d - fl;
  const uint32_t b = lmBits(f);
  double ulp;
  if (err >= 0)
    ulp = (double)lmAsFloat(b + 1) - fl;
  else
    ulp = b == 0 ? 0x1p-149 : fl - (double)lmAsFloat(b - 1);
  return 0.5 - (err >= 0 ? err : -err) / ulp;
}

/* Input bit ranges (inclusive) of the tables. */
__host__ __device__ inline void lmRange(int fn, uint32_t *lo, uint32_t *hi) {
  switch (fn) {
    case kLmLog2:
      *lo = 0x3f800000u; /* 1 */
      *hi = 0x7f7fffffu; /* FLT_MAX */
      break;
    case kLmExp2Neg:
      *lo = 0x32000000u; /* 2^-27 */
      *hi = 0x41800000u; /* 16 */
      break;
    case kLmPow5:
      *lo = 0x30000000u; /* 2^-31 */
      *hi = 0x3f802000u; /* 1 + 2^-10 */
      break;
    case kLmPow128:
      *lo = 0x3ecccccdu; /* 0.4 */
      *hi = 0x3f802000u;
      break;
    default:
      *lo = 0x00000001u; /* 2^-149 */
      *hi = 0x3f800000u; /* 1 */
      break;
  }
}

/* ------------------------------------------------------------- evaluation (the device's; the host runs the
 * same code on its own copy of the tables to test the lookup) */

/* The host libm's result for x in the table's range. */
__host__ __device__ inline float lmEval(const LibmTab &T, int fn, float x) {
  const double d = lmApprox(fn, x);
  float f = (float)d;
  if (lmMidDist(d) <= T.gate) {
    const uint32_t xb = lmBits(x), key = xb & 0xffffu;
    const uint32_t bk = (xb >> 16) - (T.lo >> 16);
    int lo = (int)T.bucket[bk], hi = (int)T.bucket[bk + 1] - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const uint32_t e = T.ent[mid], k = e & 0xffffu;
      if (k == key) {
        f = lmAsFloat(lmBits(f) + ((e >> 16) ? 1u : 0xffffffffu));
        break;
      }
      if (k < key)
        lo = mid + 1;
      else
        hi = mid - 1;
    }
  }
  return f;
}

__host__ __device__ inline bool lmInRange(const LibmTab &T, float x) {
  const uint32_t b = lmBits(x);
  return x > 0.f && b >= T.lo && b <= T.hi;
}

/* log2f as the host libm computes it (x >= 1; smaller x: correctly rounded). */
__host__ __device__ inline float hlog2f(const LibmTables *L, float x) {
  if (lmInRange(L->t[kLmLog2], x)) return lmEval(L->t[kLmLog2], kLmLog2, x);
  return (float)log2((double)x);
}

/* exp2f(t) as the host libm computes it for t in [-16, -2^-27]; elsewhere correctly rounded. */
__host__ __device__ inline float hexp2f(const LibmTables *L, float t) {
  const float u = -t;
  if (lmInRange(L->t[kLmExp2Neg], u)) return lmEval(L->t[kLmExp2Neg], kLmExp2Neg, u);
  return (float)exp2((double)t);
}

/* powf(x, N), N = 5 or 128, as the host libm computes it. */
template <int N>
__host__ __device__ inline float hpowi(const LibmTables *L, float x) {
  static_assert(N == 5 || N == 128, "exponent");
  constexpr int fn = N == 5 ? kLmPow5 : kLmPow128;
  if (lmInRange(L->t[fn], x)) return lmEval(L->t[fn], fn, x);
  return (float)pow((double)x, (double)N);
}

/* powf(x, 1/2.2f) as the host libm computes it for x in (0, 1]; elsewhere correctly rounded. */
__host__ __device__ inline float hpowGamma(const LibmTables *L, float x) {
  if (lmInRange(L->t[kLmPowGamma], x)) return lmEval(L->t[kLmPowGamma], kLmPowGamma, x);
  return (float)pow((double)x, (double)(1.0f / 2.2f));
}

}  // namespace rhip
