/*
 * pet_oracle.c — CPU restatement of the MCGPU-PET photon-transport path.
 *
 * TEST INFRASTRUCTURE ONLY (see pet_oracle.h): used by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg as the checker; never by the product.
 *
 * Citations: K = MCGPU-PET_kernel.cu, C = MCGPU-PET.cu, H = MCGPU-PET.h of the
 * reference tree. Expressions keep the reference's float/double promotion order
 * (C semantics with FLT_EVAL_METHOD 0, no contraction), because the GPU library is
 * checked against this file bit-for-bit in the shared-elementary-function build.
 *
 * Deviations from the reference, all where the reference reads indeterminate memory
 * or writes out of bounds:
 *   - GCOa: rn[] of shells with binding energy >= E is 0 (uninitialised in K:1338).
 *   - PSF: a record is only written if its own slot is < psf_size (K:598 writes
 *     index+1 == psf_size when psf_size is odd).
 *   - Stream ids are 64-bit (K:251 overflows int for Nvox*tpb >= 2^31).
 *   - locate_voxel treats NaN positions as outside and clamps indices to the grid;
 *     a sinogram angle index >= NANGLES is rejected; the RITA interval index is clamped.
 *   - The GRAa and GCOa rejection loops stop after 2^20 trials (never reached with
 *     valid tables; it bounds the run time on corrupt input).
 */
#define _GNU_SOURCE
#include "pet_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <zlib.h>

#ifdef ORACLE_SHARED_FMATH
#include "../mcgpu-pet_amd/csrc/fmath.h"
static inline float o_logf(float x) { return mcf_logf(x); }
static inline float o_expf(float x) { return mcf_expf(x); }
static inline float o_powf(float x, float y) { return mcf_powf(x, y); }
static inline void o_sincosf(float x, float* s, float* c) { mcf_sincosf(x, s, c); }
static inline float o_cosf(float x) { return mcf_cosf(x); }
static inline float o_atan2f(float y, float x) { return mcf_atan2f(y, x); }
static inline double o_log(double x) { return mcf_log(x); }
static inline double o_exp(double x) { return mcf_exp(x); }
static inline void o_sincos(double x, double* s, double* c) { mcf_sincos(x, s, c); }
static inline float o_acosf(float x) { return mcf_acosf(x); }
#else
static inline float o_logf(float x) { return logf(x); }
static inline float o_expf(float x) { return expf(x); }
static inline float o_powf(float x, float y) { return powf(x, y); }
static inline void o_sincosf(float x, float* s, float* c) { sincosf(x, s, c); }
static inline float o_cosf(float x) { return cosf(x); }
static inline float o_atan2f(float y, float x) { return atan2f(y, x); }
static inline double o_log(double x) { return log(x); }
static inline double o_exp(double x) { return exp(x); }
static inline void o_sincos(double x, double* s, double* c) { sincos(x, s, c); }
static inline float o_acosf(float x) { return acosf(x); }
#endif

#define O_PI 3.14159265358979323846            /* H:105 */
#define O_EPS_SOURCE 0.00015f                  /* H:119 */
#define O_INV_SPEEDOFLIGHT (1.0 / 29979245800e-12) /* H:127 */
#define O_E511 510998.95f                      /* H:128 */
#define O_LEAP_DISTANCE 256                    /* K:898 */
#define O_A1 40014                             /* K:900-903 */
#define O_M1 2147483563
#define O_A2 40692
#define O_M2 2147483399
#define O_MAXMAT 15
#define O_MAX_TRIALS (1 << 20)   /* rejection-loop bound shared with the library (no-op for valid tables) */

/* __float2int_rd: floor, with NaN -> INT_MIN and saturation (host and device agree). */
static inline int o_f2i_rd(float x)
{
  float f = floorf(x);
  if (!(f >= -2147483648.0f)) return (int)0x80000000;
  if (f >= 2147483648.0f) return 0x7fffffff;
  return (int)f;
}
/* __double2ull_rn with saturation. */
static inline unsigned long long o_d2ull_rn(double x)
{
  double r = rint(x);
  if (!(r > 0.0)) return 0ull;
  if (r >= 18446744073709551616.0) return ~0ull;
  return (unsigned long long)r;
}
static inline unsigned long long o_f2ull_rn(float x)
{
  float r = rintf(x);
  if (!(r > 0.0f)) return 0ull;
  if (r >= 18446744073709551616.0f) return ~0ull;
  return (unsigned long long)r;
}

/* ------------------------------------------------------------------ RANECU (K:898-1083) */

int orc_abMODm(int m, int a, int s)     /* K:1000-1032 */
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
  q = (int)m / a;
  k = (int)s / q;
  s = a * (s - k * q) - k * (m - q * a);
  while (s < 0) s += m;
  p += s;
  if (p < 0) p += m;
  return p;
}

static int o_powmod_leap(int m, int a, unsigned long long leap)   /* K:932-949 */
{
  int y = 1, z = a;
  if (leap == 0) return 1;
  for (;;) {
    if (0 != (leap & 1ull)) {
      leap >>= 1;
      y = orc_abMODm(m, z, y);
      if (0 == leap) break;
    } else {
      leap >>= 1;
    }
    z = orc_abMODm(m, z, z);
  }
  return y;
}

void orc_init_PRNG(long long history_batch, int histories_per_thread, int seed_input, int* sx, int* sy)  /* K:924-978 */
{
  unsigned long long leap = ((unsigned long long)(history_batch + 1)) * (unsigned long long)(histories_per_thread * O_LEAP_DISTANCE);
  int y1 = o_powmod_leap(O_M1, O_A1, leap);
  *sx = orc_abMODm(O_M1, seed_input, y1);
  int y2 = o_powmod_leap(O_M2, O_A2, leap);
  *sy = orc_abMODm(O_M2, seed_input, y2);
}

float orc_ranecu(int* sx, int* sy)       /* K:1044-1061 */
{
  int i1 = (int)(*sx / 53668);
  *sx = 40014 * (*sx - i1 * 53668) - i1 * 12211;
  int i2 = (int)(*sy / 52774);
  *sy = 40692 * (*sy - i2 * 52774) - i2 * 3791;
  if (*sx < 0) *sx += 2147483563;
  if (*sy < 0) *sy += 2147483399;
  i2 = *sx - *sy;
  if (i2 < 1) i2 += 2147483562;
  return ((float)i2 * 4.65661305739e-10f);
}

double orc_ranecu_double(int* sx, int* sy) /* K:1067-1083 */
{
  int i1 = (int)(*sx / 53668);
  *sx = 40014 * (*sx - i1 * 53668) - i1 * 12211;
  int i2 = (int)(*sy / 52774);
  *sy = 40692 * (*sy - i2 * 52774) - i2 * 3791;
  if (*sx < 0) *sx += 2147483563;
  if (*sy < 0) *sy += 2147483399;
  i2 = *sx - *sy;
  if (i2 < 1) i2 += 2147483562;
  return ((double)i2 * 4.6566130573917692e-10);
}

void orc_update_seed_PRNG(int batch_number, unsigned long long total_histories, int* seed)  /* C:2906-2935 */
{
  if (0 == batch_number) return;
  unsigned long long leap = total_histories * (unsigned long long)(batch_number * O_LEAP_DISTANCE);
  int y = o_powmod_leap(O_M1, O_A1, leap);
  *seed = orc_abMODm(O_M1, *seed, y);
}

void orc_ranecu_block(int seed, const long long* ids, int nstreams, int n, float* out, double* out_d, int as_double)
{
  for (int i = 0; i < nstreams; i++) {
    int sx, sy;
    orc_init_PRNG(ids[i], 555555, seed, &sx, &sy);
    for (int j = 0; j < n; j++) {
      if (as_double) out_d[(size_t)i * n + j] = orc_ranecu_double(&sx, &sy);
      else out[(size_t)i * n + j] = orc_ranecu(&sx, &sy);
    }
  }
}

/* ------------------------------------------------------------------ geometry and physics */

typedef struct o_rng { int x, y; unsigned long long draws; } o_rng;
static inline float RN(o_rng* r) { r->draws++; return orc_ranecu(&r->x, &r->y); }
static inline double RND(o_rng* r) { r->draws++; return orc_ranecu_double(&r->x, &r->y); }

static inline int o_locate(const orc_sim* s, const float* p)     /* K:1097-1124 */
{
  if ((p[1] < O_EPS_SOURCE) || (p[1] > (s->size_bbox[1] - O_EPS_SOURCE)) || (p[0] < O_EPS_SOURCE) ||
      (p[0] > (s->size_bbox[0] - O_EPS_SOURCE)) || (p[2] < O_EPS_SOURCE) || (p[2] > (s->size_bbox[2] - O_EPS_SOURCE)))
    return -100;
  /* NaN positions count as outside and indices are clamped to the grid (the
   * reference reads out of bounds in both cases). */
  if (!(p[0] == p[0] && p[1] == p[1] && p[2] == p[2])) return -100;
  int vx = o_f2i_rd(p[0] * s->inv_voxel_size[0]);
  int vy = o_f2i_rd(p[1] * s->inv_voxel_size[1]);
  int vz = o_f2i_rd(p[2] * s->inv_voxel_size[2]);
  if (vx > s->nx - 1) vx = s->nx - 1;
  if (vy > s->ny - 1) vy = s->ny - 1;
  if (vz > s->nz - 1) vz = s->nz - 1;
  return (vx + vy * (s->nx) + vz * (s->nx) * (s->ny));
}

int orc_locate_voxel(const orc_sim* s, float x, float y, float z)
{
  float p[3] = {x, y, z};
  return o_locate(s, p);
}

static void o_rotate_double(float* d, double costh, double phi)  /* K:1159-1200 */
{
  double DXY, NORM, cosphi, sinphi, SDT;
  DXY = (double)(d[0] * d[0] + d[1] * d[1]);
  o_sincos(phi, &sinphi, &cosphi);
  NORM = DXY + (double)(d[2] * d[2]);
  if (fabs(NORM - 1.0) > 1.0e-14) {
    NORM = 1.0 / sqrt(NORM);
    d[0] = (float)(NORM * d[0]);
    d[1] = (float)(NORM * d[1]);
    d[2] = (float)(NORM * d[2]);
    DXY = (double)(d[0] * d[0] + d[1] * d[1]);
  }
  if (DXY > 1.0e-28) {
    SDT = sqrt((1.0 - costh * costh) / DXY);
    float xin = d[0];
    float yz = d[1] * d[2];
    float xz = xin * d[2];
    d[0] = (float)(d[0] * costh + SDT * ((double)xz * cosphi - d[1] * sinphi));
    d[1] = (float)(d[1] * costh + SDT * ((double)yz * cosphi + xin * sinphi));
    d[2] = (float)(d[2] * costh - DXY * SDT * cosphi);
  } else {
    SDT = sqrt(1.0 - costh * costh);
    d[1] = (float)(SDT * sinphi);
    if (d[2] > 0.0) {
      d[0] = (float)(SDT * cosphi);
      d[2] = (float)costh;
    } else {
      d[0] = (float)(-SDT * cosphi);
      d[2] = (float)(-costh);
    }
  }
}

void orc_probe_rotate(int n, const float* dir_in, const double* costh, const double* phi, float* dir_out)
{
  for (int i = 0; i < n; i++) {
    float d[3] = {dir_in[3 * i], dir_in[3 * i + 1], dir_in[3 * i + 2]};
    o_rotate_double(d, costh[i], phi[i]);
    dir_out[3 * i] = d[0]; dir_out[3 * i + 1] = d[1]; dir_out[3 * i + 2] = d[2];
  }
}

/* The Rayleigh (GRAa) and Compton (GCOa) samplers below follow the PENELOPE 2006 algorithms
 * as the reference states them (MCGPU-PET_kernel.cu:1216-1228, 1318-1330), which carry this
 * notice:
 *   PENELOPE/PENGEOM (version 2006)
 *     Copyright (c) 2001-2006
 *     Universitat de Barcelona
 *   Permission to use, copy, modify, distribute and sell this software and its documentation
 *   for any purpose is hereby granted without fee, provided that the above copyright notice
 *   appears in all copies and that both that copyright notice and this permission notice
 *   appear in all supporting documentation. The Universitat de Barcelona makes no
 *   representations about the suitability of this software for any purpose. It is provided
 *   "as is" without express or implied warranty. */

static void o_GRAa(float energy, double* costh, int mat, float pmax_current, o_rng* r, const orc_rayleigh* cgra)  /* K:1230-1296 */
{
  double xmax = ((double)energy) * 8.065535669099010e-5;
  double x2lim = (double)cgra->xco[(mat + 1) * ORC_NP_RAYLEIGH - 1];
  double x2max = (xmax * xmax) < x2lim ? (xmax * xmax) : x2lim;
  if (xmax < 0.01) {
    int trials = 0;
    do {
      *costh = 1.0 - RND(r) * 2.0;
    } while (RND(r) > (((*costh) * (*costh) + 1.0) * 0.5) && ++trials < O_MAX_TRIALS);
    return;
  }
  *costh = 1.0;
  for (int trial = 0; trial < O_MAX_TRIALS; trial++) {
    double ru = RND(r) * (double)pmax_current;
    int itn = (int)(ru * (ORC_NP_RAYLEIGH - 1));
    if (itn < 0) itn = 0;                         /* clamp: no-op for valid tables */
    if (itn > ORC_NP_RAYLEIGH - 1) itn = ORC_NP_RAYLEIGH - 1;
    int i__ = (int)cgra->itlco[itn + mat * ORC_NP_RAYLEIGH];
    int j = (int)cgra->ituco[itn + mat * ORC_NP_RAYLEIGH];
    if ((j - i__) > 1) {
      do {
        int k = (i__ + j) >> 1;
        if (ru > cgra->pco[k - 1 + mat * ORC_NP_RAYLEIGH]) i__ = k;
        else j = k;
      } while ((j - i__) > 1);
    }
    int index = i__ - 1 + mat * ORC_NP_RAYLEIGH;
    double rr = ru - cgra->pco[index];
    double xx;
    if (rr > 1e-16) {
      double d__ = (double)(cgra->pco[index + 1] - cgra->pco[index]);
      float aco_index = cgra->aco[index], bco_index = cgra->bco[index], xco_index = cgra->xco[index];
      xx = (double)xco_index + (double)(aco_index + 1.0f + bco_index) * d__ * rr /
                                   (d__ * d__ + (aco_index * d__ + bco_index * rr) * rr) *
                                   (double)(cgra->xco[index + 1] - xco_index);
    } else {
      xx = cgra->xco[index];
    }
    if (xx < x2max) {
      *costh = 1.0 - 2.0 * xx / x2max;
      if (RND(r) < (((*costh) * (*costh) + 1.0) * 0.5)) break;
    }
  }
}

static void o_GCOa(float* energy, double* costh, int mat, o_rng* r, const orc_compton* cg, unsigned long long* shell_evals)  /* K:1334-1563 */
{
  float s, a1, s0, af, ek, ek2, ek3, tau, pzomc, taumin;
  float rn[ORC_MAX_SHELLS];
  double cdt1;
  int i__;
  int my_noscco = cg->noscco[mat];
  for (i__ = 0; i__ < ORC_MAX_SHELLS; i__++) rn[i__] = 0.0f;

  ek = *energy * 1.956951306108245e-6f;
  ek2 = ek * 2.f + 1.f;
  ek3 = ek * ek;
  taumin = 1.f / ek2;
  a1 = o_logf(ek2);

  s0 = 0.0f;
  for (i__ = 0; i__ < my_noscco; i__++) {
    float temp = cg->uico[mat + i__ * O_MAXMAT];
    if (temp < *energy) {
      float aux = *energy * (*energy - temp) * 2.f;
      pzomc = cg->fj0[mat + i__ * O_MAXMAT] * (aux - temp * 510998.918f) / (sqrtf(aux + aux + temp * temp) * 510998.918f);
      if (pzomc > 0.0f)
        temp = (0.707106781186545f + pzomc * 1.4142135623731f) * (0.707106781186545f + pzomc * 1.4142135623731f);
      else
        temp = (0.707106781186545f - pzomc * 1.4142135623731f) * (0.707106781186545f - pzomc * 1.4142135623731f);
      temp = 0.5f * o_expf(0.5f - temp);
      if (pzomc > 0.0f) temp = 1.0f - temp;
      s0 += cg->fco[mat + i__ * O_MAXMAT] * temp;
    }
  }

  int trials = 0;
  do {
    if (RN(r) * (a1 + 2. * ek * (ek + 1.f) * taumin * taumin) < a1)
      tau = o_powf(taumin, RN(r));
    else
      tau = sqrtf(1.f + RN(r) * (taumin * taumin - 1.f));
    cdt1 = (double)(1.f - tau) / (((double)tau) * ((double)*energy) * 1.956951306108245e-6);
    if (cdt1 > 2.0) cdt1 = 1.99999999;
    s = 0.0f;
    *shell_evals += (unsigned long long)my_noscco;
    for (i__ = 0; i__ < my_noscco; i__++) {
      float temp = cg->uico[mat + i__ * O_MAXMAT];
      if (temp < *energy) {
        float aux = (*energy) * (*energy - temp) * ((float)cdt1);
        if ((aux > 1.0e-12f) || (temp > 1.0e-12f))
          pzomc = cg->fj0[mat + i__ * O_MAXMAT] * (aux - temp * 510998.918f) / (sqrtf(aux + aux + temp * temp) * 510998.918f);
        else
          pzomc = 0.002f;
        temp = pzomc * 1.4142135623731f;
        if (pzomc > 0.0f)
          temp = 0.5f - (temp + 0.70710678118654502f) * (temp + 0.70710678118654502f);
        else
          temp = 0.5f - (0.70710678118654502f - temp) * (0.70710678118654502f - temp);
        temp = 0.5f * o_expf(temp);
        if (pzomc > 0.0f) temp = 1.0f - temp;
        s += cg->fco[mat + i__ * O_MAXMAT] * temp;
        rn[i__] = temp;
      }
    }
  } while ((RN(r) * s0) > (s * (1.0f + tau * ((ek3 - ek2 - 1.0f) + tau * (ek2 + tau * ek3))) / (ek3 * tau * (tau * tau + 1.0f))) &&
           ++trials < O_MAX_TRIALS);

  *costh = 1.0 - cdt1;

  pzomc = 0.0f;
  for (int trial = 0; trial < O_MAX_TRIALS; trial++) {
    float temp = s * RN(r);
    float pac = 0.0f;
    int ishell = my_noscco - 1;
    for (i__ = 0; i__ < (my_noscco - 1); i__++) {
      pac += cg->fco[mat + i__ * O_MAXMAT] * rn[i__];
      if (pac > temp) {
        ishell = i__;
        break;
      }
    }
    temp = RN(r) * rn[ishell];
    if (temp < 0.5f)
      pzomc = (0.70710678118654502f - sqrtf(0.5f - o_logf(temp + temp))) / (cg->fj0[mat + ishell * O_MAXMAT] * 1.4142135623731f);
    else
      pzomc = (sqrtf(0.5f - o_logf(2.0f - 2.0f * temp)) - 0.70710678118654502f) / (cg->fj0[mat + ishell * O_MAXMAT] * 1.4142135623731f);
    if (pzomc < -1.0f) continue;
    temp = (float)(tau * (tau - (*costh) * 2.f) + 1.f);
    if (temp > 1.0e-20f)
      af = sqrtf(temp) * (tau * (tau - ((float)(*costh))) / temp + 1.f);
    else
      af = 0.00200f;
    if (af > 0.0f) temp = af * 0.2f + 1.f;
    else temp = 1.f - af * 0.2f;
    float pzc = pzomc < 0.2f ? pzomc : 0.2f;
    pzc = pzc > -0.2f ? pzc : -0.2f;
    if (RN(r) * temp < (af * pzc + 1.f)) break;
  }
  {
    float t, b1, b2, temp;
    t = pzomc * pzomc;
    b1 = 1.f - t * tau * tau;
    b2 = 1.f - t * tau * ((float)(*costh));
    temp = sqrtf(fabsf(b2 * b2 - b1 * (1.0f - t)));
    if (pzomc < 0.0f) temp *= -1.0f;
    t = (tau / b1) * (b2 + temp);
    if (t > 1.0f) t = 1.0f;
    (*energy) *= t;
  }
}

/* Photon 1 / photon 2 of a pair (K:665-773). d1x is direction1.x used as the pair flag. */
static void o_source(const orc_sim* s, int ix, int iy, int iz, float* pos, float* dir, float* pos1, float* dir1,
                     float* energyFirst, o_rng* r)
{
  if (dir1[0] > 10.0f) {
    dir[2] = 1.0f - 2.0f * RN(r);
    float phi_sampled = 6.28318530716f * RN(r);
    float sin_theta = sqrtf(1.0f - dir[2] * dir[2]);
    float sp, cp;
    o_sincosf(phi_sampled, &sp, &cp);
    dir[1] = sin_theta * sp;
    dir[0] = sin_theta * cp;
    float r1 = RN(r);
    float r2 = RN(r);
    float r3 = RN(r);
    r1 = (O_EPS_SOURCE > r1 - O_EPS_SOURCE) ? O_EPS_SOURCE : r1 - O_EPS_SOURCE;
    r2 = (O_EPS_SOURCE > r2 - O_EPS_SOURCE) ? O_EPS_SOURCE : r2 - O_EPS_SOURCE;
    r3 = (O_EPS_SOURCE > r3 - O_EPS_SOURCE) ? O_EPS_SOURCE : r3 - O_EPS_SOURCE;
    pos[0] = ((float)ix + r1) / s->inv_voxel_size[0];
    pos[1] = ((float)iy + r2) / s->inv_voxel_size[1];
    pos[2] = ((float)iz + r3) / s->inv_voxel_size[2];
    dir1[0] = dir[0]; dir1[1] = dir[1]; dir1[2] = dir[2];
    pos1[0] = pos[0]; pos1[1] = pos[1]; pos1[2] = pos[2];
    *energyFirst = -10.0f;
    return;
  }
  if (s->bg_mode) {
    dir[2] = 1.0f - 2.0f * RN(r);
    float phi_sampled = 6.28318530716f * RN(r);
    float sin_theta = sqrtf(1.0f - dir[2] * dir[2]);
    float sp, cp;
    o_sincosf(phi_sampled, &sp, &cp);
    dir[1] = sin_theta * sp;
    dir[0] = sin_theta * cp;
  } else {
    float phi_sampled2 = (float)((double)o_atan2f(dir1[1], dir1[0]) + O_PI);
    float theta_sampled2 = (float)(O_PI - (double)o_acosf(dir1[2]));
    float V1 = RN(r);
    float V2 = RN(r);
    float THETA_NC = (float)((2.0 * 0.21276 * O_PI / 180.0) * (double)sqrtf(-2.0f * o_logf(V2 + 1e-8f)) *
                             (double)o_cosf((float)((2.0 * O_PI) * (double)V1)));
    float PHI_NC = (float)((2.0 * O_PI) * (double)RN(r));
    float ST, CT, CP, SP, ST1, CT1, CP1, SP1;
    o_sincosf(THETA_NC, &ST, &CT);
    o_sincosf(PHI_NC, &SP, &CP);
    o_sincosf(theta_sampled2, &ST1, &CT1);
    o_sincosf(phi_sampled2, &SP1, &CP1);
    float GX = CP * ST;
    float GY = SP * ST;
    float GZ = CT;
    float GX1 = GX * CT1 * CP1 - GY * SP1 + GZ * ST1 * CP1;
    float GY1 = GX * CT1 * SP1 + GY * CP1 + GZ * ST1 * SP1;
    float GZ1 = -GX * ST1 + GZ * CT1;
    dir[0] = GX1; dir[1] = GY1; dir[2] = GZ1;
  }
  pos[0] = pos1[0]; pos[1] = pos1[1]; pos[2] = pos1[2];
  dir1[0] = 11.1f;
}

/* Cylinder intersection of tally_Sinogram part 1 (K:1645-1700): returns 1 if both
 * photons hit within the height; writes phi, z of both. */
static int o_cylinder_pair(const orc_sim* s, const float* pos, const float* dir, const float* posf, const float* dirf,
                           float* phi, float* zn, float* phi1, float* zfn)
{
  float x0 = pos[0] - s->det_center[0];
  float y0 = pos[1] - s->det_center[1];
  float A = dir[0] * dir[0] + dir[1] * dir[1];
  float B = 2.0f * (x0 * dir[0] + y0 * dir[1]);
  float C = x0 * x0 + y0 * y0 - s->det_radius * s->det_radius;
  float dist;
  if (A == 0.0f) dist = 99999999.9f;
  else dist = (-B + sqrtf(B * B - 4.0f * A * C)) / (2.0f * A);
  float pxn = x0 + dist * dir[0];
  float pyn = y0 + dist * dir[1];
  float pzn = (pos[2] - s->det_center[2]) + dist * dir[2];
  *phi = o_atan2f(pyn, pxn);
  *zn = pzn;
  if (!(fabsf(pzn) < 0.5f * s->det_height)) return 0;
  x0 = posf[0] - s->det_center[0];
  y0 = posf[1] - s->det_center[1];
  A = dirf[0] * dirf[0] + dirf[1] * dirf[1];
  B = 2.0f * (x0 * dirf[0] + y0 * dirf[1]);
  C = x0 * x0 + y0 * y0 - s->det_radius * s->det_radius;
  if (A == 0.0f) dist = 99999999.9f;
  else dist = (-B + sqrtf(B * B - 4.0f * A * C)) / (2.0f * A);
  float pfxn = x0 + dist * dirf[0];
  float pfyn = y0 + dist * dirf[1];
  float pfzn = (posf[2] - s->det_center[2]) + dist * dirf[2];
  *phi1 = o_atan2f(pfyn, pfxn);
  *zfn = pfzn;
  if (!(fabsf(pfzn) < 0.5f * s->det_height)) return 0;
  return 1;
}

/* Crystal indices and Michelogram bin (K:1749-1788). */
static int o_bin_from_crystals(const orc_sim* s, int ix1, int iz1, int ix2, int iz2)
{
  if (iz1 < 0 || iz2 < 0 || iz1 >= s->nrows || iz2 >= s->nrows) return -1;
  if (abs(iz2 - iz1) > s->mrd) return -1;
  float delta = (float)abs(iz2 - iz1);
  int incl = o_f2i_rd((delta - 1 + ((s->span + 1) / 2)) / (float)s->span);
  int iseg = 2 * incl;
  if (iz2 < iz1 && iseg > 0) iseg = iseg - 1;
  int ofseg = iseg * s->nzs;
  for (int kka = 0; kka <= iseg; kka = kka + 1) {
    float ofk = (float)((kka - 2) * 0.5);
    if (kka > 1) ofseg = ofseg - (s->span + 1);
    if (kka > 3) ofseg = ofseg - o_f2i_rd(ofk) * 2 * s->span;
  }
  int ofz = 0;
  if (incl > 0) ofz = ofz + ((s->span + 1) / 2) + (incl - 1) * s->span;
  int izm = ofseg + (iz1 + iz2) - ofz;
  int ixtemp = ix2;
  int iztemp = iz2;
  if (ix2 < ix1) { ix2 = ix1; ix1 = ixtemp; iz2 = iz1; iz1 = iztemp; }
  (void)iz1; (void)iz2;
  int ith = ((ix1 + ix2 + s->nangles + 1) % s->ncrystals) / 2;
  int ir = abs(ix2 - ix1 - s->nangles);
  if ((ix1 < ith) || (ix2 >= (ith + s->nangles))) ir = -ir;
  ir = ir + s->nrad / 2;
  /* ith >= nangles only happens for NANGLES < NCRYSTALS/2, where the reference writes
   * out of bounds; rejected here. */
  if (izm < 0 || izm >= s->nsinos || ir < 0 || ir >= s->nrad || ith >= s->nangles) return -1;
  return izm * s->nangles * s->nrad + ith * s->nrad + ir;
}

int orc_sinogram_bin(const orc_sim* s, int ix1, int iz1, int ix2, int iz2) { return o_bin_from_crystals(s, ix1, iz1, ix2, iz2); }

static int o_detector_bin_core(const orc_sim* s, float phi, float zn, float phi1, float zfn)
{
  int ix1 = o_f2i_rd((float)(((double)phi / (2.0 * O_PI) + 0.5) * (double)s->ncrystals));
  int ix2 = o_f2i_rd((float)(((double)phi1 / (2.0 * O_PI) + 0.5) * (double)s->ncrystals));
  int iz1 = o_f2i_rd((float)(((double)(zn / s->fovz) + 0.5) * (double)s->nrows));
  int iz2 = o_f2i_rd((float)(((double)(zfn / s->fovz) + 0.5) * (double)s->nrows));
  return o_bin_from_crystals(s, ix1, iz1, ix2, iz2);
}

int orc_detector_bin(const orc_sim* s, const float* pos, const float* dir, const float* posf, const float* dirf)
{
  float phi, zn, phi1, zfn;
  if (!o_cylinder_pair(s, pos, dir, posf, dirf, &phi, &zn, &phi1, &zfn)) return -2;
  return o_detector_bin_core(s, phi, zn, phi1, zfn);
}

typedef struct o_local {
  unsigned long long c[ORC_C_N];
  long long img_t, img_sc;
} o_local;

static void o_tally_sinogram(orc_sim* s, float energy, float energyFirst, const float* pos, const float* posf,
                             const float* dir, const float* dirf, int ss, int ssf, int emit_vox, o_rng* r, o_local* L)  /* K:1623-1836 */
{
  float phi, zn, phi1, zfn;
  if (!o_cylinder_pair(s, pos, dir, posf, dirf, &phi, &zn, &phi1, &zfn)) return;
  int ibin = o_detector_bin_core(s, phi, zn, phi1, zfn);
  if (ibin < 0) return;
  float randno1 = RN(r);
  float randno2 = RN(r);
  float gaussian_var1 = sqrtf((float)(-2.0 * (double)o_logf(randno1 + 1e-8f))) * o_cosf((float)(2.0f * O_PI * (double)randno2));
  float randno3 = RN(r);
  float randno4 = RN(r);
  float gaussian_var2 = sqrtf((float)(-2.0 * (double)o_logf(randno3 + 1.0e-8f))) * o_cosf((float)(2.0f * O_PI * (double)randno4));
  float Energia1 = energy + ((s->e_resol) / 2.35f) * energy * gaussian_var1;
  float Energia2 = energyFirst + ((s->e_resol) / 2.35f) * energyFirst * gaussian_var2;
  int iE1 = o_f2i_rd(Energia1 / 1000.0f);
  int iE2 = o_f2i_rd(Energia2 / 1000.0f);
  if ((iE1 < s->ne) && (iE1 >= 0)) __atomic_fetch_add(&s->spectrum[iE1], 1, __ATOMIC_RELAXED);
  if ((iE2 < s->ne) && (iE2 >= 0)) __atomic_fetch_add(&s->spectrum[iE2], 1, __ATOMIC_RELAXED);
  L->c[ORC_C_PAIRS]++;
  if ((Energia1) > (s->e_low) && (Energia1) < (s->e_high) && (Energia2) > (s->e_low) && (Energia2) < (s->e_high)) {
    if (ss == 0 && ssf == 0) {
      __atomic_fetch_add(&s->trues[ibin], 1, __ATOMIC_RELAXED);
      L->img_t++;
      L->c[ORC_C_COINC]++;
    }
    if (ss > 0 || ssf > 0) {
      __atomic_fetch_add(&s->scatter[ibin], 1, __ATOMIC_RELAXED);
      L->img_sc++;
      L->c[ORC_C_COINC]++;
    }
  }
}

static void o_write_psf(unsigned char* base, int i, unsigned long long t_ps, float travel_ps, int vox, float e, float z,
                        float phi, const float* v, short idx1)
{
  unsigned char* p = base + (size_t)i * 44;
  short idx2 = 0;
  memcpy(p + 0, &t_ps, 8);
  memcpy(p + 8, &travel_ps, 4);
  memcpy(p + 12, &vox, 4);
  memcpy(p + 16, &e, 4);
  memcpy(p + 20, &z, 4);
  memcpy(p + 24, &phi, 4);
  memcpy(p + 28, &v[0], 4);
  memcpy(p + 32, &v[1], 4);
  memcpy(p + 36, &v[2], 4);
  memcpy(p + 40, &idx1, 2);
  memcpy(p + 42, &idx2, 2);
}

static void o_tally_psf(orc_sim* s, float energy, float energyFirst, float* position, float* positionFirst,
                        const float* direction, const float* directionFirst, int ss, int ssf, unsigned long long remaining,
                        float travel, float travelFirst, int emit_vox, o_rng* r)  /* K:530-644 */
{
  float x0 = position[0] - s->det_center[0];
  float y0 = position[1] - s->det_center[1];
  float A = direction[0] * direction[0] + direction[1] * direction[1];
  float B = 2.0f * (x0 * direction[0] + y0 * direction[1]);
  float C = x0 * x0 + y0 * y0 - s->det_radius * s->det_radius;
  float dist;
  if (A == 0.0f) dist = 99999999.9f;
  else dist = (-B + sqrtf(B * B - 4.0f * A * C)) / (2.0f * A);
  travel += dist;
  position[0] = x0 + dist * direction[0];
  position[1] = y0 + dist * direction[1];
  position[2] = (position[2] - s->det_center[2]) + dist * direction[2];
  float phi = o_atan2f(position[1], position[0]);
  if (fabsf(position[2]) < 0.5f * s->det_height) {
    x0 = positionFirst[0] - s->det_center[0];
    y0 = positionFirst[1] - s->det_center[1];
    A = directionFirst[0] * directionFirst[0] + directionFirst[1] * directionFirst[1];
    B = 2.0f * (x0 * directionFirst[0] + y0 * directionFirst[1]);
    C = x0 * x0 + y0 * y0 - s->det_radius * s->det_radius;
    if (A == 0.0f) dist = 99999999.9f;
    else dist = (-B + sqrtf(B * B - 4.0f * A * C)) / (2.0f * A);
    travelFirst += dist;
    positionFirst[0] = x0 + dist * directionFirst[0];
    positionFirst[1] = y0 + dist * directionFirst[1];
    positionFirst[2] = (positionFirst[2] - s->det_center[2]) + dist * directionFirst[2];
    float phi1 = o_atan2f(positionFirst[1], positionFirst[0]);
    float randno1 = RN(r);
    float randno2 = RN(r);
    float gaussian_var1 = sqrtf(-2.0f * o_logf(randno1 + 1e-8f)) * o_cosf((float)(2.0f * O_PI * (double)randno2));
    float randno3 = RN(r);
    float randno4 = RN(r);
    float gaussian_var2 = sqrtf(-2.0f * o_logf(randno3 + 1.0e-8f)) * o_cosf((float)(2.0f * O_PI * (double)randno4));
    float Energia1 = energy + ((s->e_resol) * (1.0f / 2.35f)) * energy * gaussian_var1;
    float Energia2 = energyFirst + ((s->e_resol) * (1.0f / 2.35f)) * energyFirst * gaussian_var2;
    if (fabsf(positionFirst[2]) < 0.5f * s->det_height && (Energia1 > s->e_low) && (Energia1 < s->e_high) &&
        (Energia2 > s->e_low) && (Energia2 < s->e_high)) {
      int index = __atomic_fetch_add(&s->psf_index, 2, __ATOMIC_RELAXED);
      if (index > 2000000000) __atomic_store_n(&s->psf_index, 2000000000, __ATOMIC_RELAXED);
      if (index < s->psf_size && s->psf) {
        int i0, i1;
        if (travel < travelFirst) { i0 = index; i1 = index + 1; }
        else { i0 = index + 1; i1 = index; }
        unsigned long long t_ps = s->acquisition_time_ps - remaining;
        if (i0 < s->psf_size)
          o_write_psf(s->psf, i0, t_ps, (float)(travel * O_INV_SPEEDOFLIGHT), emit_vox, Energia1, position[2], phi, direction, (short)ss);
        if (i1 < s->psf_size)
          o_write_psf(s->psf, i1, t_ps, (float)(travelFirst * O_INV_SPEEDOFLIGHT), emit_vox, Energia2, positionFirst[2], phi1,
                      directionFirst, (short)ssf);
      }
    }
  }
}

static void o_tally_dose(orc_sim* s, float Edep, int mat, const float* pos)   /* K:396-410, K:490-514, K:1591-1609 */
{
  if (s->material_dose && s->material_edep) {
    __atomic_fetch_add(&s->material_edep[2 * mat], o_f2ull_rn(Edep * 1.0f), __ATOMIC_RELAXED);
    __atomic_fetch_add(&s->material_edep[2 * mat + 1], o_f2ull_rn(Edep * Edep), __ATOMIC_RELAXED);
  }
  if (s->roi[1] > -1 && s->voxel_edep) {
    short vx = (short)o_f2i_rd(pos[0] * s->inv_voxel_size[0]);
    short vy = (short)o_f2i_rd(pos[1] * s->inv_voxel_size[1]);
    short vz = (short)o_f2i_rd(pos[2] * s->inv_voxel_size[2]);
    if ((vx < s->roi[0]) || (vx > s->roi[1]) || (vy < s->roi[2]) || (vy > s->roi[3]) || (vz < s->roi[4]) || (vz > s->roi[5]))
      return;
    int DX = 1 + (int)(s->roi[1] - s->roi[0]);
    int num = (int)(vx - s->roi[0]) + ((int)(vy - s->roi[2])) * DX + ((int)(vz - s->roi[4])) * DX * (1 + (int)(s->roi[3] - s->roi[2]));
    __atomic_fetch_add(&s->voxel_edep[2 * (size_t)num], o_f2ull_rn(Edep * 1.0f), __ATOMIC_RELAXED);
    __atomic_fetch_add(&s->voxel_edep[2 * (size_t)num + 1], o_f2ull_rn(Edep * Edep), __ATOMIC_RELAXED);
  }
}

/* Decay streams of a voxel and their random-number stream ids: the reference's layout
 * (threads_per_block per voxel, id = voxel * threads_per_block + thread, K:251) or the
 * balanced layout (stream_decays > 0: stream_first[v] .. stream_first[v + 1] - 1). */
static long long o_streams_of_voxel(const orc_sim* s, int v)
{
  if (s->stream_decays > 0 && s->stream_first) return s->stream_first[v + 1] - s->stream_first[v];
  return s->threads_per_block;
}
static long long o_stream_id(const orc_sim* s, int v, long long t)
{
  if (s->stream_decays > 0 && s->stream_first) return s->stream_first[v] + t;
  return (long long)v * s->threads_per_block + t;
}
static int o_is_last_stream(const orc_sim* s, int v, long long t)
{
  const long long nvox = (long long)s->nx * s->ny * s->nz;
  if (s->stream_decays > 0 && s->stream_first) return o_stream_id(s, v, t) == s->stream_first[nvox] - 1;
  return v == nvox - 1 && t == s->threads_per_block - 1;
}
/* Streams of an emitting voxel in the balanced layout: max(1, ceil(E / stream_decays)), E the
 * expected decays activity * tau (1 - exp(-T / tau)) over the acquisition T (at most 1e9). */
static long long o_balanced_count(const orc_sim* s, float activity)
{
  const double tau = (double)s->mean_life, T = (double)s->acquisition_time_ps * 1.0e-12;
  const double e = (double)activity * (tau * (1.0 - exp(-T / tau)));
  const double n = ceil(e / (double)s->stream_decays);
  if (n < 1.0) return 1;
  if (n > 1.0e9) return 1000000000LL;
  return (long long)n;
}

/* One reference GPU thread: K:223-458 for block (voxel absvox) and threadIdx tid. */
static void o_thread(orc_sim* s, int absvox_block, long long tid, o_local* L)
{
  const float* vmd = s->voxel_mat_dens;
  float voxel_activity = vmd[3 * (size_t)absvox_block + 2];
  unsigned long long acq_thread = s->acquisition_time_ps;
  unsigned long long histories_thread = 0;
  double inv_activity_thread;
  const long long per_voxel = o_streams_of_voxel(s, absvox_block);
  if (voxel_activity > 1.0e-7f) inv_activity_thread = ((double)per_voxel / (double)voxel_activity);
  else return;
  const double inv_mean_life = 1.0 / (double)s->mean_life;
  const int ix = absvox_block % s->nx;
  const int iy = (absvox_block / s->nx) % s->ny;
  const int iz = absvox_block / (s->nx * s->ny);
  o_rng r;
  r.draws = 0;
  long long thread_id = o_stream_id(s, absvox_block, tid);
  orc_init_PRNG(thread_id, 555555, s->seed, &r.x, &r.y);
  L->c[ORC_C_STREAMS]++;

  float position[3], direction[3], position1[3], direction1[3];
  float energy, step, prob = 0.0f, randno = 0.0f, mfp_density = 0.0f, mfp_Woodcock;
  float ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
  int index, material0 = 0, material_old, absvox;
  int scatter_state;
  float energyFirst = -99.9f, travel_distanceFirst = -99.9f;
  float positionFirst[3] = {0, 0, 0}, directionFirst[3] = {0, 0, 0};
  int scatter_stateFirst = 0;
  long long img_t0 = L->img_t, img_sc0 = L->img_sc;
  direction1[0] = 11.0f;
  direction1[1] = direction1[2] = 0.0f;
  position1[0] = position1[1] = position1[2] = 0.0f;

  for (;;) {
    float travel_distance = 0.0f;
    if (direction1[0] > 10.0f) {
      double Dt = -o_log(RND(&r)) * inv_activity_thread;
      unsigned long long Dt_ps = o_d2ull_rn(Dt * 1e12);
      if (Dt_ps < acq_thread) {
        acq_thread = acq_thread - Dt_ps;
        histories_thread = histories_thread + 2;
        L->c[ORC_C_DECAYS]++;
        L->c[ORC_C_EMIT_PAIR]++;
        inv_activity_thread = inv_activity_thread * o_exp(Dt * inv_mean_life);
      } else {
        break;
      }
    }
    o_source(s, ix, iy, iz, position, direction, position1, direction1, &energyFirst, &r);
    energy = O_E511;
    scatter_state = 0;
    index = o_f2i_rd((energy - s->e0) * s->ide);
    mfp_Woodcock = s->woodcock[2 * index] + energy * s->woodcock[2 * index + 1];
    material_old = -1;
    for (;;) {
      absvox = o_locate(s, position);
      if (absvox < 0) break;
      do {
        step = -(mfp_Woodcock)*o_logf(RN(&r));
        L->c[ORC_C_WOODCOCK]++;
        travel_distance += step;
        position[0] += step * direction[0];
        position[1] += step * direction[1];
        position[2] += step * direction[2];
        absvox = o_locate(s, position);
        if (absvox < 0) break;
        float mdx = vmd[3 * (size_t)absvox], mdy = vmd[3 * (size_t)absvox + 1];
        material0 = (int)(mdx - 1);
        if (material0 != material_old) {
          const size_t b = 3 * ((size_t)index * O_MAXMAT + material0);
          ra[0] = s->mfp_a[b]; ra[1] = s->mfp_a[b + 1]; ra[2] = s->mfp_a[b + 2];
          rb[0] = s->mfp_b[b]; rb[1] = s->mfp_b[b + 1]; rb[2] = s->mfp_b[b + 2];
          material_old = material0;
          L->c[ORC_C_MATCHANGE]++;
        }
        mfp_density = mfp_Woodcock * mdy;
        prob = 1.0f - mfp_density * (ra[0] + energy * rb[0]);
        randno = RN(&r);
      } while (randno < prob);
      if (absvox < 0) break;
      prob += mfp_density * (ra[1] + energy * rb[1]);
      if (randno < prob) {
        double costh_Compton;
        randno = energy;
        L->c[ORC_C_COMPTON]++;
        o_GCOa(&energy, &costh_Compton, material0, &r, s->compton, &L->c[ORC_C_SHELLEVAL]);
        o_rotate_double(direction, costh_Compton, 6.28318530717958647693 * RND(&r));
        randno = energy - randno;
        index = o_f2i_rd((energy - s->e0) * s->ide);
        if (index > -1) {
          mfp_Woodcock = s->woodcock[2 * index] + energy * s->woodcock[2 * index + 1];
          material_old = -2;
          scatter_state = (scatter_state == 0) ? 1 : 3;
        }
      } else {
        prob += mfp_density * (ra[2] + energy * rb[2]);
        if (randno < prob) {
          double costh_Rayleigh;
          L->c[ORC_C_RAYLEIGH]++;
          float pmax_current = s->rayleigh->pmax[(index + 1) * O_MAXMAT + material0];
          o_GRAa(energy, &costh_Rayleigh, material0, pmax_current, &r, s->rayleigh);
          o_rotate_double(direction, costh_Rayleigh, 6.28318530717958647693 * RND(&r));
          scatter_state = (scatter_state == 0) ? 2 : 3;
        } else {
          L->c[ORC_C_PHOTO]++;
          randno = -energy;
          index = -11;
        }
      }
      if (randno < -0.001f) {
        float Edep = -1.0f * randno;
        o_tally_dose(s, Edep, material0, position);
      }
      if (index < 0) break;
    }
    if (index > -1) {
      if (direction1[0] > 10.0f) {
        if (energyFirst > -0.5f) {
          if (s->tally_type == 0 || (s->tally_type == 1 && scatter_state == 0 && scatter_stateFirst == 0) ||
              (s->tally_type == 2 && (scatter_state != 0 || scatter_stateFirst != 0))) {
            if ((s->tally_psf_sinogram == 0) || (s->tally_psf_sinogram == 2))
              o_tally_sinogram(s, energy, energyFirst, position, positionFirst, direction, directionFirst, scatter_state,
                               scatter_stateFirst, absvox_block, &r, L);
            if ((s->tally_psf_sinogram == 0) || (s->tally_psf_sinogram == 1)) {
              float p2[3] = {position[0], position[1], position[2]};
              float pf[3] = {positionFirst[0], positionFirst[1], positionFirst[2]};
              o_tally_psf(s, energy, energyFirst, p2, pf, direction, directionFirst, scatter_state, scatter_stateFirst,
                          acq_thread, travel_distance, travel_distanceFirst, absvox_block, &r);
            }
          }
        }
      } else {
        energyFirst = energy;
        positionFirst[0] = position[0]; positionFirst[1] = position[1]; positionFirst[2] = position[2];
        directionFirst[0] = direction[0]; directionFirst[1] = direction[1]; directionFirst[2] = direction[2];
        scatter_stateFirst = scatter_state;
        travel_distanceFirst = travel_distance;
      }
    }
  }
  if (o_is_last_stream(s, absvox_block, tid)) {  /* K:454-455 */
    s->final_seed = r.x;
    s->final_seed_written = 1;
  }
  __atomic_fetch_add(&s->total_histories, histories_thread, __ATOMIC_RELAXED);
  if (L->img_t != img_t0) __atomic_fetch_add(&s->image_trues[absvox_block], (int)(L->img_t - img_t0), __ATOMIC_RELAXED);
  if (L->img_sc != img_sc0) __atomic_fetch_add(&s->image_scatter[absvox_block], (int)(L->img_sc - img_sc0), __ATOMIC_RELAXED);
  L->c[ORC_C_RNDRAWS] += r.draws;
}

/* ------------------------------------------------------------------ extended decay model
 * Restatement of the library's extensions (no reference counterpart): positron and
 * prompt-gamma branching per decay and the singles list. The photon transport and the
 * pair tallies are the reference's (K:297-442) as in o_thread; the random-number order
 * per decay is: decay time, positron branch (only if positron_fraction < 1), prompt
 * branch (only if prompt_fraction > 0), source() of photon 1, [photon 1, its single,
 * photon 2's direction, photon 2, its single or the pair tallies], [prompt gamma's
 * isotropic direction, prompt gamma, its single]. */

/* One photon from pos/dir with the given energy until it escapes or is absorbed
 * (K:297-415); the body of o_thread's inner loop. */
static void o_transport(orc_sim* s, float* position, float* direction, float* energy_io, int* index_io, int* ss_io,
                        float* travel_io, o_rng* r, o_local* L)
{
  const float* vmd = s->voxel_mat_dens;
  float energy = *energy_io, step, prob = 0.0f, randno = 0.0f, mfp_density = 0.0f, mfp_Woodcock;
  float ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
  int material0 = 0, material_old, absvox;
  int scatter_state = 0;
  float travel_distance = 0.0f;
  int index = o_f2i_rd((energy - s->e0) * s->ide);
  mfp_Woodcock = s->woodcock[2 * index] + energy * s->woodcock[2 * index + 1];
  material_old = -1;
  for (;;) {
    absvox = o_locate(s, position);
    if (absvox < 0) break;
    do {
      step = -(mfp_Woodcock)*o_logf(RN(r));
      L->c[ORC_C_WOODCOCK]++;
      travel_distance += step;
      position[0] += step * direction[0];
      position[1] += step * direction[1];
      position[2] += step * direction[2];
      absvox = o_locate(s, position);
      if (absvox < 0) break;
      float mdx = vmd[3 * (size_t)absvox], mdy = vmd[3 * (size_t)absvox + 1];
      material0 = (int)(mdx - 1);
      if (material0 != material_old) {
        const size_t b = 3 * ((size_t)index * O_MAXMAT + material0);
        ra[0] = s->mfp_a[b]; ra[1] = s->mfp_a[b + 1]; ra[2] = s->mfp_a[b + 2];
        rb[0] = s->mfp_b[b]; rb[1] = s->mfp_b[b + 1]; rb[2] = s->mfp_b[b + 2];
        material_old = material0;
        L->c[ORC_C_MATCHANGE]++;
      }
      mfp_density = mfp_Woodcock * mdy;
      prob = 1.0f - mfp_density * (ra[0] + energy * rb[0]);
      randno = RN(r);
    } while (randno < prob);
    if (absvox < 0) break;
    prob += mfp_density * (ra[1] + energy * rb[1]);
    if (randno < prob) {
      double costh_Compton;
      randno = energy;
      L->c[ORC_C_COMPTON]++;
      o_GCOa(&energy, &costh_Compton, material0, r, s->compton, &L->c[ORC_C_SHELLEVAL]);
      o_rotate_double(direction, costh_Compton, 6.28318530717958647693 * RND(r));
      randno = energy - randno;
      index = o_f2i_rd((energy - s->e0) * s->ide);
      if (index > -1) {
        mfp_Woodcock = s->woodcock[2 * index] + energy * s->woodcock[2 * index + 1];
        material_old = -2;
        scatter_state = (scatter_state == 0) ? 1 : 3;
      }
    } else {
      prob += mfp_density * (ra[2] + energy * rb[2]);
      if (randno < prob) {
        double costh_Rayleigh;
        L->c[ORC_C_RAYLEIGH]++;
        float pmax_current = s->rayleigh->pmax[(index + 1) * O_MAXMAT + material0];
        o_GRAa(energy, &costh_Rayleigh, material0, pmax_current, r, s->rayleigh);
        o_rotate_double(direction, costh_Rayleigh, 6.28318530717958647693 * RND(r));
        scatter_state = (scatter_state == 0) ? 2 : 3;
      } else {
        L->c[ORC_C_PHOTO]++;
        randno = -energy;
        index = -11;
      }
    }
    if (randno < -0.001f) o_tally_dose(s, -1.0f * randno, material0, position);
    if (index < 0) break;
  }
  *energy_io = energy;
  *index_io = index;
  *ss_io = scatter_state;
  *travel_io = travel_distance;
}

/* A photon that left the phantom: detector hit within the height, energy blur (the PSF
 * path's form, K:580-587), singles spectrum, energy window, and the 32-byte record with
 * detection time = emission time + flight time to the cylinder. */
static void o_record_single(orc_sim* s, const float* pos, const float* dir, float energy, float travel, int ss, int kind,
                            unsigned long long remaining, unsigned long long event, o_rng* r)
{
  float x0 = pos[0] - s->det_center[0];
  float y0 = pos[1] - s->det_center[1];
  float A = dir[0] * dir[0] + dir[1] * dir[1];
  float B = 2.0f * (x0 * dir[0] + y0 * dir[1]);
  float Cq = x0 * x0 + y0 * y0 - s->det_radius * s->det_radius;
  float dist;
  if (A == 0.0f) dist = 99999999.9f;
  else dist = (-B + sqrtf(B * B - 4.0f * A * Cq)) / (2.0f * A);
  float z = (pos[2] - s->det_center[2]) + dist * dir[2];
  if (!(fabsf(z) < 0.5f * s->det_height)) return;
  float randno1 = RN(r);
  float randno2 = RN(r);
  float g = sqrtf(-2.0f * o_logf(randno1 + 1e-8f)) * o_cosf((float)(2.0f * O_PI * (double)randno2));
  float e = energy + ((s->e_resol) * (1.0f / 2.35f)) * energy * g;
  int iE = o_f2i_rd(e / 1000.0f);
  if ((iE < s->ne) && (iE >= 0)) __atomic_fetch_add(&s->spectrum[iE], 1, __ATOMIC_RELAXED);
  if (!(e > s->e_low && e < s->e_high)) return;
  float phi = o_atan2f(y0 + dist * dir[1], x0 + dist * dir[0]);
  unsigned long long flight = o_d2ull_rn((double)(travel + dist) * O_INV_SPEEDOFLIGHT);
  if (flight >= (1ull << 40)) flight = (1ull << 40) - 1ull;
  unsigned long long t = (s->acquisition_time_ps - remaining) + flight;
  long long k = __atomic_fetch_add(&s->singles_n, 1, __ATOMIC_RELAXED);
  if (k < s->singles_cap && s->singles) {
    unsigned char* p = s->singles + (size_t)k * 32;
    short s16 = (short)ss, k16 = (short)kind;
    memcpy(p + 0, &t, 8);
    memcpy(p + 8, &event, 8);
    memcpy(p + 16, &e, 4);
    memcpy(p + 20, &z, 4);
    memcpy(p + 24, &phi, 4);
    memcpy(p + 28, &s16, 2);
    memcpy(p + 30, &k16, 2);
  }
}

static void o_thread_ext(orc_sim* s, int absvox_block, long long tid, o_local* L)
{
  const float* vmd = s->voxel_mat_dens;
  float voxel_activity = vmd[3 * (size_t)absvox_block + 2];
  unsigned long long acq_thread = s->acquisition_time_ps;
  unsigned long long histories_thread = 0;
  double inv_activity_thread;
  const long long per_voxel = o_streams_of_voxel(s, absvox_block);
  if (voxel_activity > 1.0e-7f) inv_activity_thread = ((double)per_voxel / (double)voxel_activity);
  else return;
  const double inv_mean_life = 1.0 / (double)s->mean_life;
  const int ix = absvox_block % s->nx;
  const int iy = (absvox_block / s->nx) % s->ny;
  const int iz = absvox_block / (s->nx * s->ny);
  const int singles = s->tally_psf_sinogram == 3;
  o_rng r;
  r.draws = 0;
  long long thread_id = o_stream_id(s, absvox_block, tid);
  orc_init_PRNG(thread_id, 555555, s->seed, &r.x, &r.y);
  L->c[ORC_C_STREAMS]++;
  unsigned long long event = (unsigned long long)thread_id << 32;
  long long img_t0 = L->img_t, img_sc0 = L->img_sc;
  for (;;) {
    double Dt = -o_log(RND(&r)) * inv_activity_thread;
    unsigned long long Dt_ps = o_d2ull_rn(Dt * 1e12);
    if (!(Dt_ps < acq_thread)) break;
    acq_thread = acq_thread - Dt_ps;
    histories_thread = histories_thread + 2;
    L->c[ORC_C_DECAYS]++;
    inv_activity_thread = inv_activity_thread * o_exp(Dt * inv_mean_life);
    event++;
    int pair = 1, prompt = 0;
    if (s->positron_fraction < 1.0f) pair = RN(&r) < s->positron_fraction;
    if (s->prompt_fraction > 0.0f) prompt = RN(&r) < s->prompt_fraction;
    float position[3], direction[3], position1[3], direction1[3] = {11.0f, 0.0f, 0.0f}, energyFirst;
    o_source(s, ix, iy, iz, position, direction, position1, direction1, &energyFirst, &r);
    if (prompt) L->c[ORC_C_EMIT_PROMPT]++;
    if (pair) {
      L->c[ORC_C_EMIT_PAIR]++;
      float energy = O_E511, travel, travelFirst = 0.0f;
      int index, ss, ssFirst = 0;
      float positionFirst[3] = {0, 0, 0}, directionFirst[3] = {0, 0, 0};
      o_transport(s, position, direction, &energy, &index, &ss, &travel, &r, L);
      if (index > -1) {
        if (singles) {
          o_record_single(s, position, direction, energy, travel, ss, 0, acq_thread, event, &r);
        } else {
          energyFirst = energy;
          memcpy(positionFirst, position, sizeof position);
          memcpy(directionFirst, direction, sizeof direction);
          ssFirst = ss;
          travelFirst = travel;
        }
      }
      o_source(s, ix, iy, iz, position, direction, position1, direction1, &energyFirst, &r);
      energy = O_E511;
      o_transport(s, position, direction, &energy, &index, &ss, &travel, &r, L);
      if (index > -1) {
        if (singles) {
          o_record_single(s, position, direction, energy, travel, ss, 1, acq_thread, event, &r);
        } else if (energyFirst > -0.5f &&
                   (s->tally_type == 0 || (s->tally_type == 1 && ss == 0 && ssFirst == 0) ||
                    (s->tally_type == 2 && (ss != 0 || ssFirst != 0)))) {
          if ((s->tally_psf_sinogram == 0) || (s->tally_psf_sinogram == 2))
            o_tally_sinogram(s, energy, energyFirst, position, positionFirst, direction, directionFirst, ss, ssFirst,
                             absvox_block, &r, L);
          if ((s->tally_psf_sinogram == 0) || (s->tally_psf_sinogram == 1)) {
            float p2[3] = {position[0], position[1], position[2]};
            float pf[3] = {positionFirst[0], positionFirst[1], positionFirst[2]};
            o_tally_psf(s, energy, energyFirst, p2, pf, direction, directionFirst, ss, ssFirst, acq_thread, travel,
                        travelFirst, absvox_block, &r);
          }
        }
      }
    }
    if (prompt) {
      float dir3[3], pos3[3] = {position1[0], position1[1], position1[2]};
      dir3[2] = 1.0f - 2.0f * RN(&r);
      float phi_sampled = 6.28318530716f * RN(&r);
      float sin_theta = sqrtf(1.0f - dir3[2] * dir3[2]);
      float sp, cp;
      o_sincosf(phi_sampled, &sp, &cp);
      dir3[1] = sin_theta * sp;
      dir3[0] = sin_theta * cp;
      float energy = s->prompt_energy, travel;
      int index, ss;
      o_transport(s, pos3, dir3, &energy, &index, &ss, &travel, &r, L);
      if (index > -1 && singles) o_record_single(s, pos3, dir3, energy, travel, ss, 2, acq_thread, event, &r);
    }
  }
  if (o_is_last_stream(s, absvox_block, tid)) {  /* K:454-455 */
    s->final_seed = r.x;
    s->final_seed_written = 1;
  }
  __atomic_fetch_add(&s->total_histories, histories_thread, __ATOMIC_RELAXED);
  if (L->img_t != img_t0) __atomic_fetch_add(&s->image_trues[absvox_block], (int)(L->img_t - img_t0), __ATOMIC_RELAXED);
  if (L->img_sc != img_sc0) __atomic_fetch_add(&s->image_scatter[absvox_block], (int)(L->img_sc - img_sc0), __ATOMIC_RELAXED);
  L->c[ORC_C_RNDRAWS] += r.draws;
}

typedef struct o_worker {
  orc_sim* s;
  long long* next;
  o_local L;
} o_worker;

static void* o_worker_main(void* arg)
{
  o_worker* w = (o_worker*)arg;
  orc_sim* s = w->s;
  const long long nvox = (long long)s->nx * s->ny * s->nz;
  const int stride = s->voxel_stride > 0 ? s->voxel_stride : 1;
  const long long nsel = (nvox - s->voxel_offset + stride - 1) / stride;
  for (;;) {
    long long k = __atomic_fetch_add(w->next, 1, __ATOMIC_RELAXED);
    if (k >= nsel) break;
    long long v = s->voxel_offset + k * stride;
    if (!(s->voxel_mat_dens[3 * v + 2] > 1.0e-7f)) continue;
    const int ext = s->tally_psf_sinogram == 3 || s->positron_fraction < 1.0f || s->prompt_fraction > 0.0f;
    const long long nst = o_streams_of_voxel(s, (int)v);
    for (long long t = 0; t < nst; t++) {
      if (ext) o_thread_ext(s, (int)v, t, &w->L);
      else o_thread(s, (int)v, t, &w->L);
    }
  }
  return NULL;
}

int orc_run(orc_sim* s)
{
  const int nt = s->nthreads > 0 ? s->nthreads : 1;
  o_worker* ws = (o_worker*)calloc((size_t)nt, sizeof(o_worker));
  pthread_t* th = (pthread_t*)calloc((size_t)nt, sizeof(pthread_t));
  if (!ws || !th) { free(ws); free(th); return -1; }
  long long next = 0;
  s->psf_index = 0;
  s->singles_n = 0;
  s->total_histories = 0;
  s->final_seed = 0;
  s->final_seed_written = 0;
  memset(s->counters, 0, sizeof s->counters);
  long long* first = NULL;
  if (s->stream_decays > 0) {   /* balanced layout: stream ids in voxel order */
    const long long nvox = (long long)s->nx * s->ny * s->nz;
    first = (long long*)malloc((size_t)(nvox + 1) * sizeof(long long));
    if (!first) { free(ws); free(th); return -1; }
    long long id = 0;
    for (long long v = 0; v < nvox; v++) {
      first[v] = id;
      const float a = s->voxel_mat_dens[3 * v + 2];
      if (a > 1.0e-7f) id += o_balanced_count(s, a);
    }
    first[nvox] = id;
  }
  s->stream_first = first;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int i = 0; i < nt; i++) {
    ws[i].s = s;
    ws[i].next = &next;
    if (nt == 1) o_worker_main(&ws[i]);
    else pthread_create(&th[i], NULL, o_worker_main, &ws[i]);
  }
  if (nt > 1)
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  s->wall_seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  for (int i = 0; i < nt; i++)
    for (int c = 0; c < ORC_C_N; c++) s->counters[c] += ws[i].L.c[c];
  free(ws);
  free(th);
  free(first);
  s->stream_first = NULL;
  return 0;
}

/* ------------------------------------------------------------------ probes */

static void o_probe_rng(o_rng* r, int seed, long long stream)
{
  r->draws = 0;
  orc_init_PRNG(stream, 555555, seed, &r->x, &r->y);
}

void orc_probe_gcoa(const orc_sim* s, int seed, long long stream, int n, float energy, int mat, float* e_out, double* costh_out)
{
  o_rng r;
  unsigned long long ev = 0;
  o_probe_rng(&r, seed, stream);
  for (int i = 0; i < n; i++) {
    float e = energy;
    o_GCOa(&e, &costh_out[i], mat, &r, s->compton, &ev);
    e_out[i] = e;
  }
}

void orc_probe_graa(const orc_sim* s, int seed, long long stream, int n, float energy, int mat, double* costh_out)
{
  o_rng r;
  o_probe_rng(&r, seed, stream);
  int index = o_f2i_rd((energy - s->e0) * s->ide);
  float pmax = s->rayleigh->pmax[(index + 1) * O_MAXMAT + mat];
  for (int i = 0; i < n; i++) o_GRAa(energy, &costh_out[i], mat, pmax, &r, s->rayleigh);
}

void orc_probe_source(const orc_sim* s, int seed, long long stream, int n, int ix, int iy, int iz, float* pos1, float* dir1,
                      float* dir2)
{
  o_rng r;
  o_probe_rng(&r, seed, stream);
  for (int i = 0; i < n; i++) {
    float pos[3], dir[3], p1[3], d1[3] = {11.0f, 0, 0}, ef;
    o_source(s, ix, iy, iz, pos, dir, p1, d1, &ef, &r);
    memcpy(pos1 + 3 * i, pos, 12);
    memcpy(dir1 + 3 * i, dir, 12);
    o_source(s, ix, iy, iz, pos, dir, p1, d1, &ef, &r);
    memcpy(dir2 + 3 * i, dir, 12);
  }
}

void orc_probe_first_interaction(const orc_sim* s, int seed, long long stream, int n, const float* pos0, const float* dir0,
                                 float energy, float* dist, int* type)
{
  o_rng r;
  o_probe_rng(&r, seed, stream);
  const float* vmd = s->voxel_mat_dens;
  for (int i = 0; i < n; i++) {
    float position[3] = {pos0[3 * i], pos0[3 * i + 1], pos0[3 * i + 2]};
    float direction[3] = {dir0[3 * i], dir0[3 * i + 1], dir0[3 * i + 2]};
    int index = o_f2i_rd((energy - s->e0) * s->ide);
    float mfp_Woodcock = s->woodcock[2 * index] + energy * s->woodcock[2 * index + 1];
    int material_old = -1, material0 = 0, absvox;
    float ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0}, prob = 0, randno = 0, mfp_density = 0, travel = 0;
    dist[i] = -1.0f;
    type[i] = 0;
    absvox = o_locate(s, position);
    if (absvox < 0) continue;
    do {
      float step = -(mfp_Woodcock)*o_logf(RN(&r));
      travel += step;
      position[0] += step * direction[0];
      position[1] += step * direction[1];
      position[2] += step * direction[2];
      absvox = o_locate(s, position);
      if (absvox < 0) break;
      material0 = (int)(vmd[3 * (size_t)absvox] - 1);
      if (material0 != material_old) {
        const size_t b = 3 * ((size_t)index * O_MAXMAT + material0);
        ra[0] = s->mfp_a[b]; ra[1] = s->mfp_a[b + 1]; ra[2] = s->mfp_a[b + 2];
        rb[0] = s->mfp_b[b]; rb[1] = s->mfp_b[b + 1]; rb[2] = s->mfp_b[b + 2];
        material_old = material0;
      }
      mfp_density = mfp_Woodcock * vmd[3 * (size_t)absvox + 1];
      prob = 1.0f - mfp_density * (ra[0] + energy * rb[0]);
      randno = RN(&r);
    } while (randno < prob);
    if (absvox < 0) continue;
    dist[i] = travel;
    prob += mfp_density * (ra[1] + energy * rb[1]);
    if (randno < prob) type[i] = 1;
    else {
      prob += mfp_density * (ra[2] + energy * rb[2]);
      type[i] = (randno < prob) ? 2 : 3;
    }
  }
}

/* ------------------------------------------------------------------ input restatements */

void orc_sinogram_sizes(int mrd, int span, int nzs, int nrad, int nangles, int* nseg, int* nsinos, int* nbins)  /* C:1176-1188 */
{
  *nseg = 2 * (int)(floor(mrd / span)) + 1;
  *nsinos = (*nseg) * nzs;
  for (int aaa = 1; aaa <= (*nseg); aaa = aaa + 1) {
    if (aaa > 1) *nsinos = *nsinos - (span + 1);
    if (aaa > 3) *nsinos = *nsinos - 2 * (int)floor((aaa - 2) / 2) * span;
  }
  *nbins = nrad * nangles * (*nsinos);
}

void orc_free(void* p) { free(p); }

/* Library extension (no reference counterpart): the energy grid continued past the files'
 * last point E_L up to e_max. Added points, as raw mean free paths: Compton scaled by the
 * Klein-Nishina total cross section (sigma_KN(E_L) / sigma_KN(E)); Rayleigh and photoelectric
 * as power laws whose exponents are the log-log slopes between E_L and the file point nearest
 * E_L / 2; total = 1 / (sum of the inverse partial mean free paths); pmax = the last value. */
static double o_klein_nishina(double e)
{
  double k = e / 510998.95, l = log(1.0 + 2.0 * k);
  return (1.0 + k) / (k * k) * (2.0 * (1.0 + k) / (1.0 + 2.0 * k) - l / k) + l / (2.0 * k) -
         (1.0 + 3.0 * k) / ((1.0 + 2.0 * k) * (1.0 + 2.0 * k));
}

static void o_extend_rows(double* rows, int n_file, int n_total, double e0, double delta_e)
{
  const double* last = rows + 6 * (size_t)(n_file - 1);
  int h = (int)floor((0.5 * last[0] - e0) / delta_e + 0.5);
  if (h > n_file - 2) h = n_file - 2;
  if (h < 0) h = 0;
  const double* half = rows + 6 * (size_t)h;
  double s_ray = log(last[1] / half[1]) / log(last[0] / half[0]);
  double s_pho = log(last[3] / half[3]) / log(last[0] / half[0]);
  double kn_last = o_klein_nishina(last[0]);
  for (int i = n_file; i < n_total; i++) {
    double* r = rows + 6 * (size_t)i;
    double e = e0 + i * delta_e;
    r[0] = e;
    r[1] = last[1] * pow(e / last[0], s_ray);
    r[2] = last[2] * kn_last / o_klein_nishina(e);
    r[3] = last[3] * pow(e / last[0], s_pho);
    r[4] = 1.0 / (1.0 / r[1] + 1.0 / r[2] + 1.0 / r[3]);
    r[5] = last[5];
  }
}

int orc_load_material_ext(const char* names, const float* density_max_in, double e_max, float* density_nominal, int* num_values,
                          float* e0, float* ide, float** woodcock, float** mfp_a, float** mfp_b, orc_rayleigh* ray,
                          orc_compton* cmp)  /* C:1565-1831 (+ the grid extension above) */
{
  char new_line[256];
  int mat, i, bin, input_num_values = 0, input_rayleigh_values = 0, input_num_shells = 0;
  double delta_e = -99999.0;
  float density_max[O_MAXMAT];
  float* W = NULL;
  float* Ta = NULL;
  float* Tb = NULL;
  double* rows = NULL;
  int N = 0, NT = 0;
  for (mat = 0; mat < O_MAXMAT; mat++) {
    density_max[mat] = density_max_in[mat];
    density_nominal[mat] = -1.0f;
    cmp->noscco[mat] = 0;
  }
  for (mat = 0; mat < O_MAXMAT; mat++) {
    const char* name = names + (size_t)mat * 250;
    if (name[0] == '\0' || name[0] == '\n') continue;
    gzFile f = gzopen(name, "rb");
    if (!f) return -2;
    do {
      if (!gzgets(f, new_line, 250)) { gzclose(f); return -3; }
    } while (strstr(new_line, "[NOMINAL DENSITY") == NULL);
    gzgets(f, new_line, 250);
    sscanf(new_line, "# %f", &density_nominal[mat]);
    if (!(density_max[mat] > 0)) {
      if (0 == mat) density_max[mat] = 0.01f * density_nominal[mat];
      else { gzclose(f); continue; }
    }
    gzgets(f, new_line, 250);
    gzgets(f, new_line, 250);
    sscanf(new_line, "# %d", &input_num_values);
    if (0 == mat) {
      N = input_num_values;
    } else if (input_num_values != N) {
      gzclose(f);
      return -4;
    }
    gzgets(f, new_line, 250);
    gzgets(f, new_line, 250);
    if (0 == mat) rows = (double*)calloc(6 * (size_t)N, sizeof(double));
    double e_last = -1.0;
    for (i = 0; i < N; i++) {
      double* r = rows + 6 * (size_t)i;
      gzgets(f, new_line, 250);
      sscanf(new_line, "  %le  %le  %le  %le  %le  %le", &r[0], &r[1], &r[2], &r[3], &r[4], &r[5]);
      if (0 == i && 0 == mat) *e0 = r[0];
      if (1 == i) delta_e = r[0] - e_last;
      e_last = r[0];
    }
    if (0 == mat) {
      NT = N;
      if (e_max > e_last) NT = N + (int)ceil((e_max - e_last) / delta_e) + 1;
      if (NT > ORC_MAX_ENERGYBINS_RAYLEIGH) { gzclose(f); free(rows); return -10; }
      rows = (double*)realloc(rows, 6 * (size_t)NT * sizeof(double));
      W = (float*)calloc(2 * (size_t)NT, sizeof(float));
      Ta = (float*)calloc(3 * (size_t)NT * O_MAXMAT, sizeof(float));
      Tb = (float*)calloc(3 * (size_t)NT * O_MAXMAT, sizeof(float));
      for (i = 0; i < NT; i++) W[2 * i] = 99999999.99f;
    }
    if (NT > N) o_extend_rows(rows, N, NT, *e0, delta_e);
    for (i = 0; i < NT; i++) {
      const double* r = rows + 6 * (size_t)i;
      double d_rayleigh = r[1], d_compton = r[2], d_total_mfp = r[4], d_pmax = r[5];
      float temp_mfp = d_total_mfp * (density_nominal[mat]) / (density_max[mat]);
      if (temp_mfp < W[2 * i]) W[2 * i] = temp_mfp;
      Ta[3 * (i * O_MAXMAT + mat) + 0] = 1.0 / (d_total_mfp * density_nominal[mat]);
      Ta[3 * (i * O_MAXMAT + mat) + 1] = 1.0 / (d_compton * density_nominal[mat]);
      Ta[3 * (i * O_MAXMAT + mat) + 2] = 1.0 / (d_rayleigh * density_nominal[mat]);
      ray->pmax[i * O_MAXMAT + mat] = d_pmax;
    }
    *ide = 1.0f / delta_e;
    for (i = 0; i < (NT - 1); i++) {
      bin = i * O_MAXMAT + mat;
      for (int c = 0; c < 3; c++) Tb[3 * bin + c] = (Ta[3 * (bin + O_MAXMAT) + c] - Ta[3 * bin + c]) / delta_e;
    }
    for (int c = 0; c < 3; c++) Tb[3 * ((NT - 1) * O_MAXMAT + mat) + c] = Tb[3 * ((NT - 2) * O_MAXMAT + mat) + c];
    for (i = 0; i < NT; i++) {
      double d_energy = *e0 + i * delta_e;
      bin = i * O_MAXMAT + mat;
      for (int c = 0; c < 3; c++) Ta[3 * bin + c] = Ta[3 * bin + c] - d_energy * Tb[3 * bin + c];
    }
    do {
      gzgets(f, new_line, 250);
      if (gzeof(f) != 0) { gzclose(f); return -5; }
    } while (strstr(new_line, "[DATA VALUES") == NULL);
    gzgets(f, new_line, 250);
    sscanf(new_line, "# %d", &input_rayleigh_values);
    if (input_rayleigh_values != ORC_NP_RAYLEIGH) { gzclose(f); return -6; }
    gzgets(f, new_line, 250);
    for (i = 0; i < input_rayleigh_values; i++) {
      int itl, itu;
      bin = ORC_NP_RAYLEIGH * mat + i;
      gzgets(f, new_line, 250);
      sscanf(new_line, "  %e  %e  %e  %e  %d  %d", &ray->xco[bin], &ray->pco[bin], &ray->aco[bin], &ray->bco[bin], &itl, &itu);
      ray->itlco[bin] = (unsigned char)itl;
      ray->ituco[bin] = (unsigned char)itu;
    }
    do {
      gzgets(f, new_line, 250);
      if (gzeof(f) != 0) { gzclose(f); return -7; }
    } while (strstr(new_line, "[NUMBER OF SHELLS") == NULL);
    gzgets(f, new_line, 250);
    sscanf(new_line, "# %d", &input_num_shells);
    if (input_num_shells > ORC_MAX_SHELLS) { gzclose(f); return -8; }
    cmp->noscco[mat] = input_num_shells;
    gzgets(f, new_line, 250);
    for (i = 0; i < input_num_shells; i++) {
      int kz, ks;
      bin = mat + i * O_MAXMAT;
      gzgets(f, new_line, 250);
      sscanf(new_line, " %e  %e  %e  %d  %d", &cmp->fco[bin], &cmp->uico[bin], &cmp->fj0[bin], &kz, &ks);
    }
    gzclose(f);
  }
  free(rows);
  if (!W) return -9;
  for (i = 0; i < (NT - 1); i++) W[2 * i + 1] = (W[2 * (i + 1)] - W[2 * i]) / delta_e;
  W[2 * (NT - 1) + 1] = 0.0f;  /* not set by the reference (heap memory) */
  for (i = 0; i < NT; i++) W[2 * i] = W[2 * i] - (*e0 + i * delta_e) * W[2 * i + 1];
  *num_values = NT;
  *woodcock = W;
  *mfp_a = Ta;
  *mfp_b = Tb;
  return 0;
}


int orc_load_material(const char* names, const float* density_max_in, float* density_nominal, int* num_values, float* e0,
                      float* ide, float** woodcock, float** mfp_a, float** mfp_b, orc_rayleigh* ray, orc_compton* cmp)  /* C:1565-1831 */
{
  return orc_load_material_ext(names, density_max_in, 0.0, density_nominal, num_values, e0, ide, woodcock, mfp_a, mfp_b, ray, cmp);
}

#ifdef ORACLE_SHARED_FMATH
/* Host evaluation (gcc) of the same elementary-function set as the library's device
 * self-test kernel, for the host/device bit-equality check. */
void orc_fmath_eval(const double* x, const double* y, int n, double* od, float* of)
{
  for (int i = 0; i < n; i++) {
    double a = x[i], b = y[i], s, c;
    mcf_sincos(a, &s, &c);
    od[0 * (size_t)n + i] = mcf_log(fabs(a) + 1e-300);
    od[1 * (size_t)n + i] = mcf_exp(a);
    od[2 * (size_t)n + i] = s;
    od[3 * (size_t)n + i] = c;
    od[4 * (size_t)n + i] = mcf_atan2(a, b);
    od[5 * (size_t)n + i] = mcf_acos(fmin(fmax(b, -1.0), 1.0));
    od[6 * (size_t)n + i] = sqrt(fabs(a));
    float af = (float)a, bf = (float)b, sf, cf;
    mcf_sincosf(af, &sf, &cf);
    of[0 * (size_t)n + i] = mcf_logf(fabsf(af) + 1e-30f);
    of[1 * (size_t)n + i] = mcf_expf(af);
    of[2 * (size_t)n + i] = mcf_powf(fminf(fabsf(bf), 1.0f), fabsf(af));
    of[3 * (size_t)n + i] = sf;
    of[4 * (size_t)n + i] = cf;
    of[5 * (size_t)n + i] = mcf_atan2f(af, bf);
    of[6 * (size_t)n + i] = sqrtf(fabsf(af)) / (bf + 3.0f);
    of[7 * (size_t)n + i] = mcf_acosf(fminf(fmaxf(bf, -1.0f), 1.0f));
  }
}
#endif
