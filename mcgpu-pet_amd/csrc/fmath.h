/*
 * fmath.h — deterministic elementary functions for the photon-transport path.
 *
 * Every function here is built only from IEEE-754 operations that are correctly
 * rounded on both the host (x86-64 SSE, g++/gcc) and the device (gfx950, hipcc):
 * + - * /, sqrt, fma, rint, fabs, copysign and bit manipulation. Compiled with
 * floating-point contraction disabled (-ffp-contract=off) on both sides, the same
 * input therefore produces the same bits on the CPU and on the MI355X. This is
 * what lets a whole photon history on the GPU be compared bit-for-bit with a CPU
 * restatement of the same algorithm.
 *
 * The float functions evaluate in double and round once, so their error is
 * below 0.5 ulp(float) + 2^-50 relative: as accurate as (or more than) the
 * correctly rounded single-precision libm results.
 *
 * Methods (all textbook):
 *   exp:   Cody-Waite reduction x = k ln2 + r with a split ln2, Taylor series to
 *          r^13 on |r| <= ln2/2 (truncation < 5e-18), exact scaling by 2^k.
 *   log:   x = 2^e m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1),
 *          odd series to s^23 (|s| <= 0.1716, truncation < 1e-18).
 *   sin/cos: Cody-Waite reduction by pi/2 with a three-part split (33+33+53 bits),
 *          Taylor series to r^19 / r^20 on |r| <= pi/4.
 *   atan:  reciprocal for |x|>1, two half-angle reductions t -> t/(1+sqrt(1+t^2))
 *          (|t| <= tan(pi/16)), odd Taylor series to t^27.
 *   atan2, acos: derived from atan with IEEE quadrant/sign-of-zero conventions.
 *
 * The header is plain C99 so that it can be included from C, C++ and HIP code.
 */
#ifndef MCPET_FMATH_H
#define MCPET_FMATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define MCF_HD static inline __host__ __device__
#else
#define MCF_HD static inline
#endif

#define MCF_INF (__builtin_huge_val())
#define MCF_NAN (__builtin_nan(""))

MCF_HD uint64_t mcf_bits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
MCF_HD double mcf_from_bits(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }

/* 2^k for -1022 <= k <= 1023 (exact). */
MCF_HD double mcf_pow2i(int k) { return mcf_from_bits((uint64_t)(k + 1023) << 52); }

/* ---------------------------------------------------------------- exp */
MCF_HD double mcf_exp(double x)
{
  if (x != x) return x;
  if (x > 0x1.62e42fefa39efp+9) return MCF_INF;       /* > log(DBL_MAX) */
  if (x < -745.2) return 0.0;
  const double inv_ln2 = 0x1.71547652b82fep+0;
  const double ln2_hi  = 0x1.62e42ff000000p-1;        /* 32 significant bits */
  const double ln2_lo  = -0x1.718432a1b0e26p-35;
  double kd = rint(x * inv_ln2);
  int k = (int)kd;
  double r = fma(-kd, ln2_hi, x);
  r = fma(-kd, ln2_lo, r);
  double p = 0x1.6124613a86d09p-33;                   /* 1/13! */
  p = fma(p, r, 0x1.1eed8eff8d898p-29);               /* 1/12! */
  p = fma(p, r, 0x1.ae64567f544e4p-26);               /* 1/11! */
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);               /* 1/10! */
  p = fma(p, r, 0x1.71de3a556c734p-19);               /* 1/9!  */
  p = fma(p, r, 0x1.a01a01a01a01ap-16);               /* 1/8!  */
  p = fma(p, r, 0x1.a01a01a01a01ap-13);               /* 1/7!  */
  p = fma(p, r, 0x1.6c16c16c16c17p-10);               /* 1/6!  */
  p = fma(p, r, 0x1.1111111111111p-7);                /* 1/5!  */
  p = fma(p, r, 0x1.5555555555555p-5);                /* 1/4!  */
  p = fma(p, r, 0x1.5555555555555p-3);                /* 1/3!  */
  p = fma(p, r, 0x1.0000000000000p-1);                /* 1/2!  */
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  /* scale by 2^k in two exact steps (k in [-1075, 1024]) */
  int k1 = k / 2;
  int k2 = k - k1;
  return (p * mcf_pow2i(k1)) * mcf_pow2i(k2);
}

/* ---------------------------------------------------------------- log */
MCF_HD double mcf_log(double x)
{
  if (x != x) return x;
  if (x < 0.0) return MCF_NAN;
  if (x == 0.0) return -MCF_INF;
  if (x == MCF_INF) return x;
  uint64_t b = mcf_bits(x);
  int e = (int)(b >> 52);
  if (e == 0) {                                        /* subnormal */
    x *= 0x1.0p54;
    b = mcf_bits(x);
    e = (int)(b >> 52) - 54;
  }
  e -= 1023;
  double m = mcf_from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp+0) { m *= 0.5; e += 1; }  /* m in [sqrt(1/2), sqrt(2)] */
  double f = m - 1.0;                                  /* exact (Sterbenz) */
  double s = f / (2.0 + f);
  double z = s * s;
  double R = 0x1.642c8590b2164p-5;                     /* 1/23 */
  R = fma(R, z, 0x1.8618618618618p-5);                 /* 1/21 */
  R = fma(R, z, 0x1.af286bca1af28p-5);                 /* 1/19 */
  R = fma(R, z, 0x1.e1e1e1e1e1e1ep-5);                 /* 1/17 */
  R = fma(R, z, 0x1.1111111111111p-4);                 /* 1/15 */
  R = fma(R, z, 0x1.3b13b13b13b14p-4);                 /* 1/13 */
  R = fma(R, z, 0x1.745d1745d1746p-4);                 /* 1/11 */
  R = fma(R, z, 0x1.c71c71c71c71cp-4);                 /* 1/9  */
  R = fma(R, z, 0x1.2492492492492p-3);                 /* 1/7  */
  R = fma(R, z, 0x1.999999999999ap-3);                 /* 1/5  */
  R = fma(R, z, 0x1.5555555555555p-2);                 /* 1/3  */
  double s2 = s + s;
  double logm = fma(s2 * z, R, s2);
  const double ln2_hi = 0x1.62e42ff000000p-1;
  const double ln2_lo = -0x1.718432a1b0e26p-35;
  double ed = (double)e;
  return ed * ln2_hi + (ed * ln2_lo + logm);
}

/* ---------------------------------------------------------------- sin / cos */
MCF_HD void mcf_sincos(double x, double* sn, double* cs)
{
  if (!(fabs(x) < 0x1.0p30)) {                         /* NaN / inf / out of supported range */
    *sn = MCF_NAN; *cs = MCF_NAN; return;
  }
  const double two_over_pi = 0x1.45f306dc9c883p-1;
  const double pio2_1 = 0x1.921fb54400000p+0;
  const double pio2_2 = 0x1.0b4611a600000p-34;
  const double pio2_3 = 0x1.3198a2e037073p-69;
  double kd = rint(x * two_over_pi);
  int q = ((int)kd) & 3;
  double r = fma(-kd, pio2_1, x);
  r = fma(-kd, pio2_2, r);
  r = fma(-kd, pio2_3, r);
  double z = r * r;
  /* sin r = r + r^3 * S(z),  S(z) = -1/3! + z/5! - z^2/7! + ... - z^8/19! */
  double sp = -0x1.2f49b46814157p-57;                  /* -1/19! */
  sp = fma(sp, z,  0x1.952c77030ad4ap-49);             /* +1/17! */
  sp = fma(sp, z, -0x1.ae7f3e733b81fp-41);             /* -1/15! */
  sp = fma(sp, z,  0x1.6124613a86d09p-33);             /* +1/13! */
  sp = fma(sp, z, -0x1.ae64567f544e4p-26);             /* -1/11! */
  sp = fma(sp, z,  0x1.71de3a556c734p-19);             /* +1/9!  */
  sp = fma(sp, z, -0x1.a01a01a01a01ap-13);             /* -1/7!  */
  sp = fma(sp, z,  0x1.1111111111111p-7);              /* +1/5!  */
  sp = fma(sp, z, -0x1.5555555555555p-3);              /* -1/3!  */
  double s = fma(r * z, sp, r);
  /* cos r = 1 + z * C(z),  C(z) = -1/2! + z/4! - ... + z^9/20! */
  double cp = 0x1.e542ba4020225p-62;                   /* +1/20! */
  cp = fma(cp, z, -0x1.6827863b97d97p-53);             /* -1/18! */
  cp = fma(cp, z,  0x1.ae7f3e733b81fp-45);             /* +1/16! */
  cp = fma(cp, z, -0x1.93974a8c07c9dp-37);             /* -1/14! */
  cp = fma(cp, z,  0x1.1eed8eff8d898p-29);             /* +1/12! */
  cp = fma(cp, z, -0x1.27e4fb7789f5cp-22);             /* -1/10! */
  cp = fma(cp, z,  0x1.a01a01a01a01ap-16);             /* +1/8!  */
  cp = fma(cp, z, -0x1.6c16c16c16c17p-10);             /* -1/6!  */
  cp = fma(cp, z,  0x1.5555555555555p-5);              /* +1/4!  */
  cp = fma(cp, z, -0x1.0000000000000p-1);              /* -1/2!  */
  double c = fma(z, cp, 1.0);
  switch (q) {
    case 0:  *sn = s;  *cs = c;  break;
    case 1:  *sn = c;  *cs = -s; break;
    case 2:  *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s;  break;
  }
}

/* ---------------------------------------------------------------- atan */
MCF_HD double mcf_atan(double x)
{
  if (x != x) return x;
  double ax = fabs(x);
  if (ax == MCF_INF) return copysign(0x1.921fb54442d18p+0, x);
  int inv = ax > 1.0;
  double t = inv ? 1.0 / ax : ax;
  t = t / (1.0 + sqrt(fma(t, t, 1.0)));
  t = t / (1.0 + sqrt(fma(t, t, 1.0)));
  double z = t * t;
  double p = -0x1.2f684bda12f68p-5;                    /* -1/27 */
  p = fma(p, z,  0x1.47ae147ae147bp-5);                /* +1/25 */
  p = fma(p, z, -0x1.642c8590b2164p-5);                /* -1/23 */
  p = fma(p, z,  0x1.8618618618618p-5);                /* +1/21 */
  p = fma(p, z, -0x1.af286bca1af28p-5);                /* -1/19 */
  p = fma(p, z,  0x1.e1e1e1e1e1e1ep-5);                /* +1/17 */
  p = fma(p, z, -0x1.1111111111111p-4);                /* -1/15 */
  p = fma(p, z,  0x1.3b13b13b13b14p-4);                /* +1/13 */
  p = fma(p, z, -0x1.745d1745d1746p-4);                /* -1/11 */
  p = fma(p, z,  0x1.c71c71c71c71cp-4);                /* +1/9  */
  p = fma(p, z, -0x1.2492492492492p-3);                /* -1/7  */
  p = fma(p, z,  0x1.999999999999ap-3);                /* +1/5  */
  p = fma(p, z, -0x1.5555555555555p-2);                /* -1/3  */
  double a = 4.0 * fma(t * z, p, t);
  if (inv) a = 0x1.921fb54442d18p+0 - a;
  return copysign(a, x);
}

MCF_HD double mcf_atan2(double y, double x)
{
  const double pi = 0x1.921fb54442d18p+1;
  const double pio2 = 0x1.921fb54442d18p+0;
  if (x != x || y != y) return x + y;
  if (y == 0.0) {
    if (x > 0.0 || (x == 0.0 && !signbit(x))) return y;   /* +-0 */
    return copysign(pi, y);
  }
  if (x == 0.0) return copysign(pio2, y);
  if (fabs(x) == MCF_INF || fabs(y) == MCF_INF) {
    if (fabs(x) == MCF_INF && fabs(y) == MCF_INF)
      return copysign(x > 0.0 ? 0.25 * pi : 0.75 * pi, y);
    if (fabs(y) == MCF_INF) return copysign(pio2, y);
    return x > 0.0 ? copysign(0.0, y) : copysign(pi, y);
  }
  double a = mcf_atan(y / x);
  if (x > 0.0) return a;
  return y > 0.0 ? a + pi : a - pi;
}

MCF_HD double mcf_acos(double x)
{
  if (x != x || fabs(x) > 1.0) return MCF_NAN;
  return mcf_atan2(sqrt((1.0 - x) * (1.0 + x)), x);
}

/* ---------------------------------------------------------------- float wrappers */
MCF_HD float mcf_logf(float x) { return (float)mcf_log((double)x); }
MCF_HD float mcf_expf(float x) { return (float)mcf_exp((double)x); }
MCF_HD float mcf_powf(float x, float y)
{
  /* used for x in (0,1], finite y: exp(y*log x) */
  if (x == 1.0f || y == 0.0f) return 1.0f;
  return (float)mcf_exp((double)y * mcf_log((double)x));
}
MCF_HD void mcf_sincosf(float x, float* s, float* c)
{
  double sd, cd;
  mcf_sincos((double)x, &sd, &cd);
  *s = (float)sd; *c = (float)cd;
}
MCF_HD float mcf_cosf(float x) { double s, c; mcf_sincos((double)x, &s, &c); return (float)c; }
MCF_HD float mcf_atan2f(float y, float x) { return (float)mcf_atan2((double)y, (double)x); }
MCF_HD float mcf_acosf(float x) { return (float)mcf_acos((double)x); }

#endif /* MCPET_FMATH_H */
