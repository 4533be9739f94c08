/*
 * ranecu.h — RANECU (L'Ecuyer 1988 combined MLCG) with sequence-splitting skip-ahead,
 * as used by the reference transport (MCGPU-PET_kernel.cu:898-1083).
 *
 * Stream layout (reference K:251-253, K:931): stream i starts at the position
 * (i+1) * 555555 * 256 of both MLCGs, both seeded with the same input seed.
 * The multiplier a^((i+1)L) mod m is computed here from a host-side table of
 * a^(L 2^k) mod m (k = 0..62), one modular product per set bit of i+1, instead of
 * the reference's square-and-multiply over the full leap; the value is the same
 * integer. The final product with the input seed uses the reference's own
 * Russian-peasant/approximate-factoring routine so that any int seed maps exactly
 * as in the reference.
 */
#ifndef MCPET_RANECU_H
#define MCPET_RANECU_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MCR_HD static inline __host__ __device__
#else
#define MCR_HD static inline
#endif

#define MCR_A1 40014
#define MCR_M1 2147483563
#define MCR_A2 40692
#define MCR_M2 2147483399
#define MCR_LEAP_DISTANCE 256
#define MCR_HIST_PER_THREAD 555555
#define MCR_JUMP_BITS 63

/* (a*s) mod m for 0 < a, s < m by L'Ecuyer & Cote's Russian-peasant + approximate
 * factoring (exact in 32-bit arithmetic; identical to the reference's abMODm for all int inputs). */
MCR_HD int mcr_abmodm(int m, int a, int s)
{
  int q, k;
  int p = -m;
  while (a > 32768) {
    if (0 != (a & 1)) {
      p += s;
      if (p > 0) p -= m;
    }
    a >>= 1;
    s = (s - m) + s;
    if (s < 0) s += m;
  }
  q = m / a;
  k = s / q;
  s = a * (s - k * q) - k * (m - q * a);
  while (s < 0) s += m;
  p += s;
  if (p < 0) p += m;
  return p;
}

MCR_HD uint32_t mcr_mulmod(uint32_t a, uint32_t b, uint32_t m) { return (uint32_t)(((uint64_t)a * b) % m); }

/* (a*b) mod m for m = 2^31 - c (both RANECU moduli), a, b < m: two folds of the high
 * part (2^31 = c mod m) and one conditional subtraction; the same integer as the
 * generic remainder, without a 64-bit division. */
MCR_HD uint32_t mcr_mulmod_pm(uint32_t a, uint32_t b, uint32_t c)
{
  const uint64_t x = (uint64_t)a * b;                       /* < 2^62 */
  const uint64_t y = (x >> 31) * c + (x & 0x7fffffffull);   /* < 250 * 2^31 */
  uint32_t z = (uint32_t)((y >> 31) * c + (y & 0x7fffffffull)); /* < 2^31 + 62250 */
  const uint32_t m = 0x80000000u - c;
  if (z >= m) z -= m;
  return z;
}

/* Host: table[k] = a^(L*2^k) mod m, L = 555555*256. */
static inline void mcr_build_jump_table(int a, int m, uint32_t* table)
{
  uint64_t L = (uint64_t)MCR_HIST_PER_THREAD * MCR_LEAP_DISTANCE;
  /* a^L by square-and-multiply */
  uint64_t base = (uint64_t)a, r = 1, e = L;
  while (e) {
    if (e & 1) r = (r * base) % (uint64_t)m;
    base = (base * base) % (uint64_t)m;
    e >>= 1;
  }
  table[0] = (uint32_t)r;
  for (int k = 1; k < MCR_JUMP_BITS; k++) table[k] = mcr_mulmod(table[k - 1], table[k - 1], (uint32_t)m);
}

/* Seeds of stream `id` (reference thread id) from the jump tables. */
MCR_HD void mcr_init_stream(int64_t id, int seed_input, const uint32_t* j1, const uint32_t* j2, int* sx, int* sy)
{
  uint64_t n = (uint64_t)(id + 1);
  uint32_t y1 = 1, y2 = 1;
  for (int k = 0; k < MCR_JUMP_BITS && n; k++, n >>= 1) {
    if (n & 1ull) {
      y1 = mcr_mulmod_pm(y1, j1[k], 0x80000000u - (uint32_t)MCR_M1);
      y2 = mcr_mulmod_pm(y2, j2[k], 0x80000000u - (uint32_t)MCR_M2);
    }
  }
  if (seed_input > 0 && seed_input < MCR_M2) {
    /* abMODm(m, a, s) = a*s mod m for 0 < a, s < m */
    *sx = (int)mcr_mulmod_pm((uint32_t)seed_input, y1, 0x80000000u - (uint32_t)MCR_M1);
    *sy = (int)mcr_mulmod_pm((uint32_t)seed_input, y2, 0x80000000u - (uint32_t)MCR_M2);
  } else {
    *sx = mcr_abmodm(MCR_M1, seed_input, (int)y1);
    *sy = mcr_abmodm(MCR_M2, seed_input, (int)y2);
  }
}

/* Products of operands below 2^24 whose result fits in 31 bits: one full-rate 24-bit
 * multiply on the device instead of a quarter-rate 32-bit one; the same integer. */
#if defined(__HIP_DEVICE_COMPILE__)
#define MCR_MUL24(a, b) __mul24((a), (b))
#else
#define MCR_MUL24(a, b) ((a) * (b))
#endif

/* One step of both MLCGs (Schrage decomposition), combined output in [1, m1-1].
 * The states lie in [1, m-1], so the quotients are taken unsigned (no sign fix-up), and
 * every product has operands below 2^16 (quotient < 40015, remainder < q) and a result
 * below 2^31: 24-bit multiplies. */
MCR_HD int mcr_step(int* sx, int* sy)
{
  int i1 = (int)((unsigned)*sx / 53668u);
  *sx = MCR_MUL24(40014, *sx - MCR_MUL24(i1, 53668)) - MCR_MUL24(i1, 12211);
  int i2 = (int)((unsigned)*sy / 52774u);
  *sy = MCR_MUL24(40692, *sy - MCR_MUL24(i2, 52774)) - MCR_MUL24(i2, 3791);
  if (*sx < 0) *sx += 2147483563;
  if (*sy < 0) *sy += 2147483399;
  i2 = *sx - *sy;
  if (i2 < 1) i2 += 2147483562;
  return i2;
}

/* float_rn(i2) * 2^-31 (the reference constant 4.65661305739e-10f is exactly 2^-31 in binary32). */
MCR_HD float mcr_ranecu(int* sx, int* sy) { return (float)mcr_step(sx, sy) * 4.65661305739e-10f; }
MCR_HD double mcr_ranecu_double(int* sx, int* sy) { return (double)mcr_step(sx, sy) * 4.6566130573917692e-10; }

#endif /* MCPET_RANECU_H */
