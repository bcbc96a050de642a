/*
 * Deterministic binary32 power function for the exner recovery
 *   exner = pow(zz*R/p0*rtheta, R/cv)      (cuda_driver.py:1466-1472, driver.py:800-815)
 *
 * Method: binary64 evaluation of exp(y*ln x) rounded once to binary32.
 *   ln x:  x = m*2^e with m in [1/sqrt2, sqrt2); ln m = 2*atanh(s),
 *          s = (m-1)/(m+1), by the odd Taylor series of atanh to s^23
 *          (|s| <= 0.1716, truncation < 1e-19); e*ln2 with a Cody-Waite
 *          split of ln2 (hi part has 32 significant bits, so e*hi is exact).
 *   exp t: t = k*ln2 + r, k = nearest integer of t/ln2, |r| <= ln2/2; exp r by
 *          its Taylor series to degree 13 (truncation < 5e-18), then an exact
 *          scaling by 2^k built from the exponent bits.
 * Only IEEE-754 binary64 +, -, *, /, floor and conversions are used, so the
 * same source gives the same bits on the host (SSE2, no contraction) and on
 * gfx950 (no contraction, IEEE division).  The double result carries a
 * relative error of a few 1e-16, so the binary32 rounding equals the
 * correctly rounded power except within that distance of a rounding tie.
 *
 * The reference's v8.4.1 exner uses the GNU C library's powf, which is
 * within 0.82 ulp, not correctly rounded: for the exner exponent R/cv it
 * returns the other neighbour of the exact value on 74 241 of the
 * 117 440 512 binary32 arguments in [2^-12, 4).  hexdyn_gnu_powf_fix turns
 * the correctly rounded result into the GNU one from a bitmap of those
 * arguments, which the library builds once per process by calling the host's
 * powf (the only source of the GNU bits used here) and comparing.
 */
#ifndef HEXDYN_POWF_H
#define HEXDYN_POWF_H

#include <stdint.h>

#if defined(__HIPCC__)
#define HEXDYN_HD __host__ __device__
#else
#define HEXDYN_HD
#endif

HEXDYN_HD static inline uint32_t hexdyn_f32_bits(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, sizeof(u));
  return u;
}

HEXDYN_HD static inline float hexdyn_f32_from_bits(uint32_t u) {
  float x;
  __builtin_memcpy(&x, &u, sizeof(x));
  return x;
}

HEXDYN_HD static inline double hexdyn_f64_from_bits(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, sizeof(x));
  return x;
}

/* Arguments of the GNU correction bitmap: [2^-12, 4) as binary32 bits. */
#define HEXDYN_POW_LO 0x39800000u
#define HEXDYN_POW_HI 0x40800000u

/* x^y in binary64, before the final rounding to binary32 (special values
 * are returned exactly). */
HEXDYN_HD static inline double hexdyn_pow_d(float xf, float yf) {
  const uint32_t ix = hexdyn_f32_bits(xf);
  if (!(xf > 0.0f) || ix >= 0x7f800000u) {
    if (xf == 0.0f) return yf > 0.0f ? 0.0 : (double)hexdyn_f32_from_bits(0x7f800000u);
    if (ix == 0x7f800000u) return yf > 0.0f ? (double)xf : 0.0;
    return (double)hexdyn_f32_from_bits(0x7fc00000u);
  }
  int e;
  uint32_t mbits;
  if (ix < 0x00800000u) {
    const uint32_t jx = hexdyn_f32_bits(xf * 8388608.0f); /* exact for subnormals */
    e = (int)((jx >> 23) & 0xffu) - 127 - 23;
    mbits = (jx & 0x007fffffu) | 0x3f800000u;
  } else {
    e = (int)((ix >> 23) & 0xffu) - 127;
    mbits = (ix & 0x007fffffu) | 0x3f800000u;
  }
  double m = (double)hexdyn_f32_from_bits(mbits);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  const double ln2_hi = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
  const double ln2_lo = 1.90821492927058770002e-10;
  const double ed = (double)e;
  const double lnx = ed * ln2_hi + (ed * ln2_lo + (2.0 * s) * p);
  const double t = (double)yf * lnx;
  if (t > 89.0) return (double)hexdyn_f32_from_bits(0x7f800000u);
  if (t < -104.0) return 0.0;
  const double kd = __builtin_floor(t * 1.4426950408889634 + 0.5);
  const double r = (t - kd * ln2_hi) - kd * ln2_lo;
  double q = 1.0 / 6227020800.0;          /* 1/13! */
  q = q * r + 1.0 / 479001600.0;          /* 1/12! */
  q = q * r + 1.0 / 39916800.0;
  q = q * r + 1.0 / 3628800.0;
  q = q * r + 1.0 / 362880.0;
  q = q * r + 1.0 / 40320.0;
  q = q * r + 1.0 / 5040.0;
  q = q * r + 1.0 / 720.0;
  q = q * r + 1.0 / 120.0;
  q = q * r + 1.0 / 24.0;
  q = q * r + 1.0 / 6.0;
  q = q * r + 0.5;
  q = q * r + 1.0;
  q = q * r + 1.0;
  const int k = (int)kd;
  const double scale = hexdyn_f64_from_bits((uint64_t)(k + 1023) << 52);
  return q * scale;
}

HEXDYN_HD static inline float hexdyn_powf_eval(float xf, float yf) { return (float)hexdyn_pow_d(xf, yf); }

/* The GNU powf bits of a positive argument x whose correctly rounded power
 * is r (from the binary64 value d): where the bitmap marks x, the GNU result
 * is the neighbour of r on the side of the exact value.  Every marked
 * argument lies within a hair of a rounding midpoint (|d - r| >= 0.4983 ulp
 * over the whole range for the exner exponent), so the bitmap is read only
 * where |d - r| >= near_tie ulp, the smallest such distance the host found
 * when it built the bitmap (about 0.4 % of arguments). */
HEXDYN_HD static inline float hexdyn_gnu_powf_fix(float x, float r, double d, const uint32_t *ties, double near_tie) {
  const uint32_t ix = hexdyn_f32_bits(x);
  if (ix < HEXDYN_POW_LO || ix >= HEXDYN_POW_HI) return r;
  const double rd = (double)r;
  const double ulp = (double)hexdyn_f32_from_bits(hexdyn_f32_bits(r) + 1u) - rd;
  if (__builtin_fabs(d - rd) < near_tie * ulp) return r;
  const uint32_t i = ix - HEXDYN_POW_LO;
  if (!((ties[i >> 5] >> (i & 31u)) & 1u)) return r;
  const uint32_t ir = hexdyn_f32_bits(r);
  return hexdyn_f32_from_bits(d > (double)r ? ir + 1u : ir - 1u);
}

#endif /* HEXDYN_POWF_H */
