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
 * The double functions are accurate to a few ulp(double). The float functions are
 * evaluated in float with explicit fmaf and are accurate to 2.5 ulp(float).
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
 *   float log/exp/sincos/atan2/acos: see the float section below.
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

/* ---------------------------------------------------------------- float functions
 * Evaluated in float with explicit fmaf (a single rounding on both sides). Polynomial
 * coefficients were fitted (weighted least squares iterated towards minimax) to the
 * relative error of each kernel on its reduced interval; the fit error is below 1e-9
 * relative, so the result error is dominated by float rounding (<= 2.5 ulp, measured
 * against a double-precision reference on 2*10^7 inputs per function). */

#define MCF_INFF (__builtin_huge_valf())
#define MCF_NANF (__builtin_nanf(""))

MCF_HD uint32_t mcf_fbits(float x) { uint32_t u; __builtin_memcpy(&u, &x, 4); return u; }
MCF_HD float mcf_ffrom(uint32_t u) { float x; __builtin_memcpy(&x, &u, 4); return x; }
/* 2^k for -126 <= k <= 127 (exact). */
MCF_HD float mcf_pow2if(int k) { return mcf_ffrom((uint32_t)(k + 127) << 23); }

#define MCF_LN2F_HI 0x1.62e4p-1f          /* 16 significant bits: e*LN2F_HI exact */
#define MCF_LN2F_LO 0x1.7f7d1cp-20f
#define MCF_PIF_HI 0x1.921fb6p+1f
#define MCF_PIF_LO -0x1.777a5cp-24f
#define MCF_PIO2F_HI 0x1.921fb6p+0f
#define MCF_PIO2F_LO -0x1.777a5cp-25f
#define MCF_PIO4F_HI 0x1.921fb6p-1f
#define MCF_PIO4F_LO -0x1.777a5cp-26f

/* log: x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
 * log(1+f) = f - f^2/2 + f^3 P(f), P of degree 8 on [sqrt(1/2)-1, sqrt(2)-1]. */
MCF_HD float mcf_logf(float x)
{
  if (!(x > 0.0f)) return (x == 0.0f) ? -MCF_INFF : MCF_NANF;
  if (x == MCF_INFF) return x;
  uint32_t ix = mcf_fbits(x);
  int e = 0;
  if (ix < 0x00800000u) {                              /* subnormal */
    x *= 0x1.0p25f;
    ix = mcf_fbits(x);
    e = -25;
  }
  e += (int)(ix >> 23) - 127;
  uint32_t mi = (ix & 0x007fffffu) | 0x3f800000u;
  if (mi > 0x3fb504f3u) {                              /* m > sqrt(2): use m/2 */
    mi -= 0x00800000u;
    e += 1;
  }
  const float f = mcf_ffrom(mi) - 1.0f;
  const float z = f * f;
  float p = 0x1.1455b2p-4f;
  p = fmaf(p, f, -0x1.de3bdcp-4f);
  p = fmaf(p, f, 0x1.e6f132p-4f);
  p = fmaf(p, f, -0x1.fc221ap-4f);
  p = fmaf(p, f, 0x1.23375ap-3f);
  p = fmaf(p, f, -0x1.555c4cp-3f);
  p = fmaf(p, f, 0x1.99a4b6p-3f);
  p = fmaf(p, f, -0x1.000006p-2f);
  p = fmaf(p, f, 0x1.555548p-2f);
  const float r = fmaf(f * z, p, -0.5f * z) + f;
  const float ef = (float)e;
  return ef * MCF_LN2F_HI + fmaf(ef, MCF_LN2F_LO, r);
}

/* exp: x = k ln2 + r (split ln2, |r| <= ln2/2), e^r = 1 + r + r^2 Q(r), Q of degree 5,
 * exact scaling by 2^k in two steps (gradual underflow rounds once). */
MCF_HD float mcf_expf(float x)
{
  if (x != x) return x;
  if (x > 0x1.62e42ep+6f) return MCF_INFF;             /* > log(FLT_MAX) */
  if (x < -0x1.9fe368p+6f) return 0.0f;                /* < log(2^-150)  */
  const float kf = rintf(x * 0x1.715476p+0f);
  const int k = (int)kf;
  float r = fmaf(-kf, MCF_LN2F_HI, x);
  r = fmaf(-kf, MCF_LN2F_LO, r);
  float q = 0x1.9f08cep-13f;
  q = fmaf(q, r, 0x1.6d8cfap-10f);
  q = fmaf(q, r, 0x1.111270p-7f);
  q = fmaf(q, r, 0x1.55548ep-5f);
  q = fmaf(q, r, 0x1.555554p-3f);
  q = fmaf(q, r, 0x1.0p-1f);
  const float p = fmaf(r * r, q, r) + 1.0f;
  const int k1 = k / 2;
  const int k2 = k - k1;
  return (p * mcf_pow2if(k1)) * mcf_pow2if(k2);
}

/* x^y = exp(y log x) for x in (0, 1] and finite y. The error grows with |y log x|
 * (about 3 ulp for |y log x| <= 2, the range of its use: x in [1/3, 1], y in (0, 1)). */
MCF_HD float mcf_powf(float x, float y)
{
  if (x == 1.0f || y == 0.0f) return 1.0f;
  return mcf_expf(y * mcf_logf(x));
}

/* sin/cos: x = k pi/2 + r, the reduction done in double with a two-part pi/2 (33 + 53
 * bits, exact product for |k| < 2^13) and r rounded once to float; sin r = r + r z S(z),
 * cos r = 1 - z/2 + z^2 C(z), z = r^2, S and C of degree 3 on |r| <= pi/4.
 * |x| >= 8192 (and inf/NaN) go through the double routine. */
MCF_HD void mcf_sincosf(float x, float* sn, float* cs)
{
  if (!(fabsf(x) < 8192.0f)) {
    double sd, cd;
    mcf_sincos((double)x, &sd, &cd);
    *sn = (float)sd;
    *cs = (float)cd;
    return;
  }
  const float kf = rintf(x * 0x1.45f306p-1f);
  const int q = ((int)kf) & 3;
  const double kd = (double)kf;
  double rd = fma(-kd, 0x1.921fb54400000p+0, (double)x);  /* exact: kd*hi has <= 46 bits */
  rd = fma(-kd, 0x1.0b4611a626331p-34, rd);
  const float r = (float)rd;
  const float z = r * r;
  float sp = 0x1.6cd1eep-19f;
  sp = fmaf(sp, z, -0x1.a00f80p-13f);
  sp = fmaf(sp, z, 0x1.111108p-7f);
  sp = fmaf(sp, z, -0x1.555556p-3f);
  const float s = fmaf(r * z, sp, r);
  float cp = -0x1.241e96p-22f;
  cp = fmaf(cp, z, 0x1.a010dep-16f);
  cp = fmaf(cp, z, -0x1.6c16b8p-10f);
  cp = fmaf(cp, z, 0x1.555556p-5f);
  const float hz = 0.5f * z;
  const float w = 1.0f - hz;
  const float c = w + (((1.0f - w) - hz) + (z * z) * cp);
  switch (q) {
    case 0:  *sn = s;  *cs = c;  break;
    case 1:  *sn = c;  *cs = -s; break;
    case 2:  *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s;  break;
  }
}
MCF_HD float mcf_cosf(float x) { float s, c; mcf_sincosf(x, &s, &c); return c; }

/* atan2: octant reduction to t = min/max in [0, 1]; for t > tan(pi/8) the angle is
 * pi/4 + atan((n - d)/(n + d)), so one division serves both cases;
 * atan t = t + t z A(z), A of degree 5 on |t| <= tan(pi/8). */
MCF_HD float mcf_atan2f(float y, float x)
{
  if (x != x || y != y) return x + y;
  if (y == 0.0f) {
    if (x > 0.0f || (x == 0.0f && !signbit(x))) return y;
    return copysignf(MCF_PIF_HI, y);
  }
  const float ax = fabsf(x), ay = fabsf(y);
  if (x == 0.0f) return copysignf(MCF_PIO2F_HI, y);
  if (ax == MCF_INFF || ay == MCF_INFF) {
    if (ax == MCF_INFF && ay == MCF_INFF) return copysignf(x > 0.0f ? MCF_PIO4F_HI : 0x1.2d97c8p+1f, y);
    if (ay == MCF_INFF) return copysignf(MCF_PIO2F_HI, y);
    return x > 0.0f ? copysignf(0.0f, y) : copysignf(MCF_PIF_HI, y);
  }
  const int swap = ay > ax;
  const float n = swap ? ax : ay, d = swap ? ay : ax;
  const int oct = n > 0x1.a8279ap-2f * d;              /* t > tan(pi/8) */
  const float t = oct ? (n - d) / (n + d) : n / d;
  const float z = t * t;
  float p = 0x1.84a122p-5f;
  p = fmaf(p, z, -0x1.5b2feap-4f);
  p = fmaf(p, z, 0x1.c463f2p-4f);
  p = fmaf(p, z, -0x1.247edap-3f);
  p = fmaf(p, z, 0x1.999918p-3f);
  p = fmaf(p, z, -0x1.555554p-2f);
  float a = fmaf(t * z, p, t);
  if (oct) a = MCF_PIO4F_HI + (a + MCF_PIO4F_LO);
  if (swap) a = MCF_PIO2F_HI - (a - MCF_PIO2F_LO);
  if (x < 0.0f) a = MCF_PIF_HI - (a - MCF_PIF_LO);
  return copysignf(a, y);
}

/* acos: |x| <= 1/2: pi/2 - asin x; otherwise 2 asin(sqrt((1-|x|)/2)) (reflected for x < 0);
 * asin s = s + s z P(z), P of degree 5 on z = s^2 <= 1/4. */
MCF_HD float mcf_acosf(float x)
{
  if (!(fabsf(x) <= 1.0f)) return MCF_NANF;
  const float ax = fabsf(x);
  float z, s;
  if (ax <= 0.5f) {
    z = x * x;
    s = x;
  } else {
    z = (1.0f - ax) * 0.5f;                             /* exact */
    s = sqrtf(z);
  }
  float p = 0x1.335568p-5f;
  p = fmaf(p, z, 0x1.d92094p-7f);
  p = fmaf(p, z, 0x1.049192p-5f);
  p = fmaf(p, z, 0x1.6caed4p-5f);
  p = fmaf(p, z, 0x1.33387cp-4f);
  p = fmaf(p, z, 0x1.55554cp-3f);
  const float as = fmaf(s * z, p, s);
  if (ax <= 0.5f) return MCF_PIO2F_HI - (as - MCF_PIO2F_LO);
  if (x > 0.0f) return 2.0f * as;
  return MCF_PIF_HI - (2.0f * as - MCF_PIF_LO);
}

#endif /* MCPET_FMATH_H */
