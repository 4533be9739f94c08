// kernels.hip — gfx950 photon-transport kernels.
//
// One lane runs one decay stream at a time: the reference's per-thread Poisson
// process (MCGPU-PET_kernel.cu:229-292) with its own RANECU stream. The kernels
// differ only in how lanes get streams:
//   plain      : lane i of the grid owns stream i (one pass, no queue);
//   persistent : a fixed grid of waves pulls stream indices from a device counter and
//                refills a lane as soon as its stream ends (streams are ordered by
//                decreasing voxel activity, so the long streams start first).
// Per lane the sequence of random numbers and the floating-point operations are those
// of the reference thread with the same id; integer tallies are order-independent, so
// both kernels produce the same tallies bit for bit.
//
// Memory layout (MI355X-first):
//   - voxels: a u8 (or u16) index per voxel into a palette of (density, material)
//     words when the phantom has few distinct pairs (the volume then stays in each XCD's
//     L2), else one 8-byte word per voxel; activity is only read once per stream from a
//     compacted, activity-sorted list of emitting voxels;
//   - cross sections: total (a,b) pairs in an 8-byte table read in the Woodcock loop,
//     Compton/Rayleigh (a,b) in a separate 16-byte table read only at real events;
//   - Compton shell table (5460 B), the energy spectrum and the per-material dose
//     are kept in LDS and flushed once per workgroup;
//   - per-lane state touched by only some phases of the persistent kernel (the pair kept
//     between the two photons, the Compton sampler between rejection trials) lives in
//     LDS slots, so that registers hold only the stream and the current photon.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>

#include "devutil.h"
#include "fmath.h"
#include "kparams.h"

#define K_PI 3.14159265358979323846
#define K_EPS 0.00015f
#define K_INV_C (1.0 / 29979245800e-12)
#define K_E511 510998.95f
#define K_MAXMAT 15
// LDS address space: the wave's stream-queue range is passed to a non-inlined function, which
// would otherwise see a generic pointer and use flat accesses.
#define K_LDS __attribute__((address_space(3)))
// Upper bound on the iterations of each rejection loop (GRAa, GCOa). Valid tables accept
// within a few trials; the bound only guarantees that no lane can loop forever on
// corrupt input. The CPU checker applies the same bound.
#define K_MAX_TRIALS (1 << 20)

// Occupancy target of the transport kernels (waves per SIMD; 0 leaves it to the compiler).
// The persistent kernel is bound by the latency of dependent gathers and transcendental
// chains; 4 waves per SIMD (128 VGPRs, ~10 dwords of spill) hides more of it than 3 (+17%
// on the head workload) or 5 (spills grow), and than the compiler's own 2.
// Woodcock steps a lane may take in one STEP phase while its steps stay virtual, and GCOa
// rejection trials (tau loop, p_z loop) in one CTAU / CPZ phase: the phase selection's ballots
// are paid once for several. 3 and 2 measured best on the head workload (1-4 steps, 1-3
// trials); more inflate the register pressure of every phase.
#ifndef MCPET_STEPS_PER_PHASE
#define MCPET_STEPS_PER_PHASE 3
#endif
#ifndef MCPET_TRIALS_PER_PHASE
#define MCPET_TRIALS_PER_PHASE 2
#endif
// Blocks chained after the selected phase's (bit flags): 1 GCOa's tau trials after an
// interaction, 2 its p_z trials after the tau trials, 4 Woodcock steps after the p_z trials,
// 8 Woodcock steps after an emission, 16 the next decay after a pair tally. All five measured
// best on the head workload (6.8 % fewer cycles than none; results are the same bit for bit,
// since a lane's own sequence of operations does not depend on when its blocks run).
#ifndef MCPET_CHAIN
#define MCPET_CHAIN 31
#endif
#ifndef MCPET_WAVES_PER_EU
#define MCPET_WAVES_PER_EU 4
#endif
#if MCPET_WAVES_PER_EU > 0
#define K_OCCUPANCY __attribute__((amdgpu_waves_per_eu(MCPET_WAVES_PER_EU, 8)))
#else
#define K_OCCUPANCY
#endif
// Per-phase execution statistics of the persistent kernel (diagnostic builds only).
#ifndef MCPET_SM_STATS
#define MCPET_SM_STATS 0
#endif
// Rarely executed physics (Compton, Rayleigh, tallies) as real calls instead of inlined code.
#ifndef MCPET_NOINLINE_RARE
#define MCPET_NOINLINE_RARE 0
#endif
#if MCPET_NOINLINE_RARE
#define K_RARE __device__ __noinline__
#else
#define K_RARE __device__
#endif

namespace {

struct Rng {
  int x, y;
};

// Event counters, a compile-time policy of every device function below.
// FullCounters (instrumented instantiation): every event count, per lane in registers,
// flushed once per lane. LiteCounters (production): only decays, streams, tallied pairs
// and coincidences, as workgroup counters in LDS, so that the transport keeps no counter
// registers. Both instantiations perform the same arithmetic on the same random numbers.
struct FullCounters {
  unsigned c[KC_N];
  __device__ __forceinline__ void init(unsigned*)
  {
#pragma unroll
    for (int i = 0; i < KC_N; i++) c[i] = 0;
  }
  __device__ __forceinline__ void inc(int k) { c[k]++; }
  __device__ __forceinline__ void add(int k, unsigned v) { c[k] += v; }
  __device__ __forceinline__ void flush(const KParams& P, const unsigned*)
  {
#pragma unroll
    for (int i = 0; i < KC_N; i++)
      if (c[i]) gatomic_add(&P.counters[i], (unsigned long long)c[i]);
  }
};
struct LiteCounters {
  unsigned* s;  // KC_N workgroup counters in LDS
  __device__ __forceinline__ void init(unsigned* lds) { s = lds; }
  __device__ __forceinline__ void inc(int k)
  {
    if (k == KC_DECAYS || k == KC_STREAMS || k == KC_PAIRS || k == KC_COINC) atomicAdd(&s[k], 1u);
  }
  __device__ __forceinline__ void add(int, unsigned) {}
  __device__ __forceinline__ void flush(const KParams&, const unsigned*) {}
};

template <class CN>
__device__ __forceinline__ float rn(Rng& r, CN& C)
{
  C.inc(KC_RNDRAWS);
  return mcr_ranecu(&r.x, &r.y);
}
template <class CN>
__device__ __forceinline__ double rnd(Rng& r, CN& C)
{
  C.inc(KC_RNDRAWS);
  return mcr_ranecu_double(&r.x, &r.y);
}

// locate_voxel (K:1097-1124). NaN positions count as outside and indices are clamped
// to the grid (the reference would read out of bounds in both cases). Evaluated without
// branches: the six bound tests (a NaN coordinate fails them) combine into one select, and
// inside the box every floor(p * inverse size) lies in [0, n], so it converts directly after
// clamping to [0, n - 1]; the same index as the reference's early returns and
// __float2int_rd.
__device__ __forceinline__ int locate(const KParams& P, const float* p)
{
  const bool inside = (p[1] >= K_EPS) & (p[1] <= (P.by - K_EPS)) & (p[0] >= K_EPS) & (p[0] <= (P.bx - K_EPS)) &
                      (p[2] >= K_EPS) & (p[2] <= (P.bz - K_EPS));
  const int vx = (int)fminf(fmaxf(floorf(p[0] * P.ivx), 0.0f), (float)(P.nx - 1));
  const int vy = (int)fminf(fmaxf(floorf(p[1] * P.ivy), 0.0f), (float)(P.ny - 1));
  const int vz = (int)fminf(fmaxf(floorf(p[2] * P.ivz), 0.0f), (float)(P.nz - 1));
  return inside ? (vx + vy * (P.nx) + vz * (P.nx) * (P.ny)) : -100;
}

// rotate_double (K:1159-1200): PENELOPE DIRECT in double precision on a float vector.
__device__ __forceinline__ void rotate_double(float* d, double costh, double phi)
{
  double DXY, NORM, cosphi, sinphi, SDT;
  DXY = (double)(d[0] * d[0] + d[1] * d[1]);
  mcf_sincos(phi, &sinphi, &cosphi);
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

// The Rayleigh (GRAa) and Compton (GCOa) samplers below follow the PENELOPE 2006 algorithms
// as the reference states them (MCGPU-PET_kernel.cu:1216-1228, 1318-1330), which carry this
// notice:
//   PENELOPE/PENGEOM (version 2006)
//     Copyright (c) 2001-2006
//     Universitat de Barcelona
//   Permission to use, copy, modify, distribute and sell this software and its documentation
//   for any purpose is hereby granted without fee, provided that the above copyright notice
//   appears in all copies and that both that copyright notice and this permission notice
//   appear in all supporting documentation. The Universitat de Barcelona makes no
//   representations about the suitability of this software for any purpose. It is provided
//   "as is" without express or implied warranty.

// GRAa (K:1230-1296): Rayleigh angle by RITA inverse cumulative + Thomson rejection.
template <class CN>
K_RARE void graa(float energy, double* costh, int mat, float pmax_current, Rng& r, CN& C,
                     const mcpet_rayleigh_table* __restrict__ cgra_in)
{
  const DU_GLOBAL mcpet_rayleigh_table* cgra = (const DU_GLOBAL mcpet_rayleigh_table*)cgra_in;
  double xmax = ((double)energy) * 8.065535669099010e-5;
  double x2lim = (double)cgra->xco[(mat + 1) * MCPET_NP_RAYLEIGH - 1];
  double x2max = (xmax * xmax) < x2lim ? (xmax * xmax) : x2lim;
  if (xmax < 0.01) {
    int trials = 0;
    do {
      *costh = 1.0 - rnd(r, C) * 2.0;
    } while (rnd(r, C) > (((*costh) * (*costh) + 1.0) * 0.5) && ++trials < K_MAX_TRIALS);
    return;
  }
  *costh = 1.0;
  for (int trial = 0; trial < K_MAX_TRIALS; trial++) {
    double ru = rnd(r, C) * (double)pmax_current;
    int itn = (int)(ru * (MCPET_NP_RAYLEIGH - 1));
    itn = min(max(itn, 0), MCPET_NP_RAYLEIGH - 1);
    int i__ = (int)cgra->itlco[itn + mat * MCPET_NP_RAYLEIGH];
    int j = (int)cgra->ituco[itn + mat * MCPET_NP_RAYLEIGH];
    if ((j - i__) > 1) {
      do {
        int k = (i__ + j) >> 1;
        if (ru > cgra->pco[k - 1 + mat * MCPET_NP_RAYLEIGH]) i__ = k;
        else j = k;
      } while ((j - i__) > 1);
    }
    int index = i__ - 1 + mat * MCPET_NP_RAYLEIGH;
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
      if (rnd(r, C) < (((*costh) * (*costh) + 1.0) * 0.5)) break;
    }
  }
}

// Compton profile term of shell i at momentum transfer factor `cdtf` (the inner body of
// the incoherent-scattering-function loops, K:1407-1449); 0 for inaccessible shells.
__device__ __forceinline__ float shell_term(const mcpet_compton_table& cg, int mat, int i, float energy, float cdtf, bool use_cdt)
{
  float temp = cg.uico[mat + i * K_MAXMAT];
  if (!(temp < energy)) return 0.0f;
  float pzomc;
  if (!use_cdt) {
    float aux = energy * (energy - temp) * 2.f;
    pzomc = cg.fj0[mat + i * K_MAXMAT] * (aux - temp * 510998.918f) / (sqrtf(aux + aux + temp * temp) * 510998.918f);
    if (pzomc > 0.0f)
      temp = (0.707106781186545f + pzomc * 1.4142135623731f) * (0.707106781186545f + pzomc * 1.4142135623731f);
    else
      temp = (0.707106781186545f - pzomc * 1.4142135623731f) * (0.707106781186545f - pzomc * 1.4142135623731f);
    temp = 0.5f * mcf_expf(0.5f - temp);
    if (pzomc > 0.0f) temp = 1.0f - temp;
    return temp;
  }
  float aux = (energy) * (energy - temp) * cdtf;
  if ((aux > 1.0e-12f) || (temp > 1.0e-12f))
    pzomc = cg.fj0[mat + i * K_MAXMAT] * (aux - temp * 510998.918f) / (sqrtf(aux + aux + temp * temp) * 510998.918f);
  else
    pzomc = 0.002f;
  temp = pzomc * 1.4142135623731f;
  if (pzomc > 0.0f)
    temp = 0.5f - (temp + 0.70710678118654502f) * (temp + 0.70710678118654502f);
  else
    temp = 0.5f - (0.70710678118654502f - temp) * (0.70710678118654502f - temp);
  temp = 0.5f * mcf_expf(temp);
  if (pzomc > 0.0f) temp = 1.0f - temp;
  return temp;
}

// GCOa (K:1334-1563): Compton scattering, relativistic impulse approximation with
// analytical one-electron profiles, split into resumable pieces so that the persistent
// kernel can run one rejection trial per phase: gcoa_init (S(theta=pi) over the shells),
// gcoa_tau_trial (one trial of the tau/shell-profile loop, K:1389-1459), gcoa_pz_trial
// (one trial of the p_z loop, K:1463-1539) and gcoa_finish (the final energy, K:1543-1557).
// The per-shell values rn[i] of the accepted tau trial are kept for the first four shells
// and recomputed for later shells instead of a 30-entry array (same values).
struct ComptonState {
  float a1, s0, taumin;            // set by gcoa_init
  float tau, s, cdtf;              // accepted tau trial
  float rc0, rc1, rc2, rc3;        // its profile terms of shells 0..3
  float pzomc;
  double cdt1;
  int trials;
};

__device__ __forceinline__ void gcoa_init(float E, int mat, const mcpet_compton_table& cg, ComptonState& cs)
{
  const int my_noscco = cg.noscco[mat];
  const float ek = E * 1.956951306108245e-6f;
  const float ek2 = ek * 2.f + 1.f;
  cs.taumin = 1.f / ek2;
  cs.a1 = mcf_logf(ek2);
  float s0 = 0.0f;
  for (int i = 0; i < my_noscco; i++) {
    if (cg.uico[mat + i * K_MAXMAT] < E) s0 += cg.fco[mat + i * K_MAXMAT] * shell_term(cg, mat, i, E, 0.0f, false);
  }
  cs.s0 = s0;
  cs.trials = 0;
}

// One trial of the tau loop; true when the loop is over (accepted or trial cap).
template <class CN>
__device__ __forceinline__ bool gcoa_tau_trial(float E, int mat, Rng& r, CN& C, const mcpet_compton_table& cg,
                                               ComptonState& cs)
{
  const int my_noscco = cg.noscco[mat];
  const float ek = E * 1.956951306108245e-6f;
  const float ek2 = ek * 2.f + 1.f;
  const float ek3 = ek * ek;
  const float taumin = cs.taumin, a1 = cs.a1;
  float tau;
  if (rn(r, C) * (a1 + 2. * ek * (ek + 1.f) * taumin * taumin) < a1)
    tau = mcf_powf(taumin, rn(r, C));
  else
    tau = sqrtf(1.f + rn(r, C) * (taumin * taumin - 1.f));
  double cdt1 = (double)(1.f - tau) / (((double)tau) * ((double)E) * 1.956951306108245e-6);
  if (cdt1 > 2.0) cdt1 = 1.99999999;
  const float cdtf = (float)cdt1;
  float s = 0.0f;
  C.add(KC_SHELLEVAL, (unsigned)my_noscco);
  float rc0 = 0.0f, rc1 = 0.0f, rc2 = 0.0f, rc3 = 0.0f;
  for (int i = 0; i < my_noscco; i++) {
    if (cg.uico[mat + i * K_MAXMAT] < E) {
      const float t = shell_term(cg, mat, i, E, cdtf, true);
      s += cg.fco[mat + i * K_MAXMAT] * t;
      rc0 = (i == 0) ? t : rc0;
      rc1 = (i == 1) ? t : rc1;
      rc2 = (i == 2) ? t : rc2;
      rc3 = (i == 3) ? t : rc3;
    }
  }
  cs.tau = tau;
  cs.cdt1 = cdt1;
  cs.cdtf = cdtf;
  cs.s = s;
  cs.rc0 = rc0;
  cs.rc1 = rc1;
  cs.rc2 = rc2;
  cs.rc3 = rc3;
  const bool reject =
      (rn(r, C) * cs.s0) > (s * (1.0f + tau * ((ek3 - ek2 - 1.0f) + tau * (ek2 + tau * ek3))) / (ek3 * tau * (tau * tau + 1.0f)));
  if (!reject) return true;
  return ++cs.trials >= K_MAX_TRIALS;
}

__device__ __forceinline__ float gcoa_rnv(float E, int mat, const mcpet_compton_table& cg, const ComptonState& cs, int i)
{
  if (i < 4) return i == 0 ? cs.rc0 : i == 1 ? cs.rc1 : i == 2 ? cs.rc2 : cs.rc3;
  return shell_term(cg, mat, i, E, cs.cdtf, true);
}

// One trial of the p_z loop; true when the loop is over (accepted or trial cap).
template <class CN>
__device__ __forceinline__ bool gcoa_pz_trial(float E, int mat, Rng& r, CN& C, const mcpet_compton_table& cg,
                                              ComptonState& cs)
{
  const int my_noscco = cg.noscco[mat];
  const double costh = 1.0 - cs.cdt1;
  const float tau = cs.tau;
  float temp = cs.s * rn(r, C);
  float pac = 0.0f;
  int ishell = my_noscco - 1;
  for (int i = 0; i < (my_noscco - 1); i++) {
    pac += cg.fco[mat + i * K_MAXMAT] * gcoa_rnv(E, mat, cg, cs, i);
    if (pac > temp) {
      ishell = i;
      break;
    }
  }
  temp = rn(r, C) * gcoa_rnv(E, mat, cg, cs, ishell);
  const float fj = cg.fj0[mat + ishell * K_MAXMAT];
  float pzomc;
  if (temp < 0.5f)
    pzomc = (0.70710678118654502f - sqrtf(0.5f - mcf_logf(temp + temp))) / (fj * 1.4142135623731f);
  else
    pzomc = (sqrtf(0.5f - mcf_logf(2.0f - 2.0f * temp)) - 0.70710678118654502f) / (fj * 1.4142135623731f);
  cs.pzomc = pzomc;
  if (!(pzomc < -1.0f)) {
    float af;
    temp = (float)(tau * (tau - costh * 2.f) + 1.f);
    if (temp > 1.0e-20f)
      af = sqrtf(temp) * (tau * (tau - ((float)costh)) / temp + 1.f);
    else
      af = 0.00200f;
    if (af > 0.0f) temp = af * 0.2f + 1.f;
    else temp = 1.f - af * 0.2f;
    float pzc = pzomc < 0.2f ? pzomc : 0.2f;
    pzc = pzc > -0.2f ? pzc : -0.2f;
    if (rn(r, C) * temp < (af * pzc + 1.f)) return true;
  }
  return ++cs.trials >= K_MAX_TRIALS;
}

// Final energy (K:1543-1557); returns cos(theta) of the accepted trial.
__device__ __forceinline__ double gcoa_finish(float* energy, const ComptonState& cs)
{
  const double costh = 1.0 - cs.cdt1;
  const float tau = cs.tau, pzomc = cs.pzomc;
  float t, b1, b2, temp;
  t = pzomc * pzomc;
  b1 = 1.f - t * tau * tau;
  b2 = 1.f - t * tau * ((float)costh);
  temp = sqrtf(fabsf(b2 * b2 - b1 * (1.0f - t)));
  if (pzomc < 0.0f) temp *= -1.0f;
  t = (tau / b1) * (b2 + temp);
  if (t > 1.0f) t = 1.0f;
  (*energy) *= t;
  return costh;
}

struct Stream {
  Rng r;
  unsigned long long remaining;  // acquisition time left (ps)
  double inv_act;                // tpb / activity, decayed per emission
  int absvox;
  unsigned long long event;      // (reference thread id << 32) | number of the current decay
};

// Cylinder intersection of both photons (K:1645-1707 / K:536-577).
__device__ __forceinline__ float cyl_dist(const KParams& P, float x0, float y0, const float* dir)
{
  float A = dir[0] * dir[0] + dir[1] * dir[1];
  float B = 2.0f * (x0 * dir[0] + y0 * dir[1]);
  float Cc = x0 * x0 + y0 * y0 - P.radius * P.radius;
  if (A == 0.0f) return 99999999.9f;
  return (-B + sqrtf(B * B - 4.0f * A * Cc)) / (2.0f * A);
}

__device__ __forceinline__ void spectrum_add(const KParams& P, int* s_spec, int iE)
{
  if (P.lds_spectrum) __hip_atomic_fetch_add((K_LDS int*)s_spec + iE, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else gatomic_add(&P.spectrum[iE], 1);
}

// tally_Sinogram (K:1623-1836).
template <class CN>
K_RARE void tally_sinogram(const KParams& P, float energy, float energyFirst, const float* pos, const float* posf,
                               const float* dir, const float* dirf, int ss, int ssf, Stream& st, CN& C, int* s_spec)
{
  float x0 = pos[0] - P.cx, y0 = pos[1] - P.cy;
  float dist = cyl_dist(P, x0, y0, dir);
  float pxn = x0 + dist * dir[0];
  float pyn = y0 + dist * dir[1];
  float pzn = (pos[2] - P.cz) + dist * dir[2];
  if (!(fabsf(pzn) < 0.5f * P.height)) return;
  x0 = posf[0] - P.cx;
  y0 = posf[1] - P.cy;
  dist = cyl_dist(P, x0, y0, dirf);
  float pfxn = x0 + dist * dirf[0];
  float pfyn = y0 + dist * dirf[1];
  float pfzn = (posf[2] - P.cz) + dist * dirf[2];
  if (!(fabsf(pfzn) < 0.5f * P.height)) return;
  // the azimuths only matter once both hits are inside the detector height (K:1669,1700)
  const float phi = mcf_atan2f(pyn, pxn);
  const float phi1 = mcf_atan2f(pfyn, pfxn);
  int ibin = sino_bin(P, phi, pzn, phi1, pfzn);
  if (ibin < 0) return;
  float randno1 = rn(st.r, C);
  float randno2 = rn(st.r, C);
  float g1 = sqrtf((float)(-2.0 * (double)mcf_logf(randno1 + 1e-8f))) * mcf_cosf((float)(2.0f * K_PI * (double)randno2));
  float randno3 = rn(st.r, C);
  float randno4 = rn(st.r, C);
  float g2 = sqrtf((float)(-2.0 * (double)mcf_logf(randno3 + 1.0e-8f))) * mcf_cosf((float)(2.0f * K_PI * (double)randno4));
  float Energia1 = energy + ((P.e_resol) / 2.35f) * energy * g1;
  float Energia2 = energyFirst + ((P.e_resol) / 2.35f) * energyFirst * g2;
  int iE1 = f2i_rd(Energia1 / 1000.0f);
  int iE2 = f2i_rd(Energia2 / 1000.0f);
  if ((iE1 < P.ne) && (iE1 >= 0)) spectrum_add(P, s_spec, iE1);
  if ((iE2 < P.ne) && (iE2 >= 0)) spectrum_add(P, s_spec, iE2);
  C.inc(KC_PAIRS);
  if ((Energia1) > (P.e_low) && (Energia1) < (P.e_high) && (Energia2) > (P.e_low) && (Energia2) < (P.e_high)) {
    if (ss == 0 && ssf == 0) {
      gatomic_add(&P.trues[ibin], 1);
      gatomic_add(&P.img_t[st.absvox], 1);
      C.inc(KC_COINC);
    }
    if (ss > 0 || ssf > 0) {
      gatomic_add(&P.scatter[ibin], 1);
      gatomic_add(&P.img_sc[st.absvox], 1);
      C.inc(KC_COINC);
    }
    if (P.rb_trues) {   // the same event on the rebinning grid (sinogram, angle, radial bin of ibin)
      const int per = P.nangles * P.nrad;
      const int izm = ibin / per, rem = ibin - izm * per;
      const int ith = rem / P.nrad, ir = rem - ith * P.nrad;
      const int cb = (P.rb_plane[izm] * P.rb_H + ith / P.rb_gh) * P.rb_W + ir / P.rb_gw;
      gatomic_add((ss == 0 && ssf == 0) ? &P.rb_trues[cb] : &P.rb_scatter[cb], 1);
    }
  }
}

__device__ __forceinline__ void write_psf(DU_GLOBAL mcpet_psf_record* rec, unsigned long long t_ps, float travel_ps, int vox, float e,
                                          float z, float phi, const float* v, int idx1)
{
  rec->emission_time_ps = t_ps;
  rec->travel_time_ps = travel_ps;
  rec->emission_absvox = vox;
  rec->energy = e;
  rec->z = z;
  rec->phi = phi;
  rec->vx = v[0];
  rec->vy = v[1];
  rec->vz = v[2];
  rec->index1 = (int16_t)idx1;
  rec->index2 = 0;
}

// tally_PSF_coincidences (K:530-644).
template <class CN>
K_RARE void tally_psf(const KParams& P, float energy, float energyFirst, const float* pos_in, const float* posf_in,
                          const float* dir, const float* dirf, int ss, int ssf, float travel, float travelFirst, Stream& st,
                          CN& C)
{
  float position[3], positionFirst[3];
  float x0 = pos_in[0] - P.cx, y0 = pos_in[1] - P.cy;
  float dist = cyl_dist(P, x0, y0, dir);
  travel += dist;
  position[0] = x0 + dist * dir[0];
  position[1] = y0 + dist * dir[1];
  position[2] = (pos_in[2] - P.cz) + dist * dir[2];
  if (!(fabsf(position[2]) < 0.5f * P.height)) return;
  x0 = posf_in[0] - P.cx;
  y0 = posf_in[1] - P.cy;
  dist = cyl_dist(P, x0, y0, dirf);
  travelFirst += dist;
  positionFirst[0] = x0 + dist * dirf[0];
  positionFirst[1] = y0 + dist * dirf[1];
  positionFirst[2] = (posf_in[2] - P.cz) + dist * dirf[2];
  float randno1 = rn(st.r, C);
  float randno2 = rn(st.r, C);
  float g1 = sqrtf(-2.0f * mcf_logf(randno1 + 1e-8f)) * mcf_cosf((float)(2.0f * K_PI * (double)randno2));
  float randno3 = rn(st.r, C);
  float randno4 = rn(st.r, C);
  float g2 = sqrtf(-2.0f * mcf_logf(randno3 + 1.0e-8f)) * mcf_cosf((float)(2.0f * K_PI * (double)randno4));
  float Energia1 = energy + ((P.e_resol) * (1.0f / 2.35f)) * energy * g1;
  float Energia2 = energyFirst + ((P.e_resol) * (1.0f / 2.35f)) * energyFirst * g2;
  if (fabsf(positionFirst[2]) < 0.5f * P.height && (Energia1 > P.e_low) && (Energia1 < P.e_high) && (Energia2 > P.e_low) &&
      (Energia2 < P.e_high)) {
    int index = gatomic_add(P.psf_index, 2);
    if (index > 2000000000) gatomic_min(P.psf_index, 2000000000);
    if (index < P.psf_size && P.psf) {
      int i0, i1;
      if (travel < travelFirst) {
        i0 = index;
        i1 = index + 1;
      } else {
        i0 = index + 1;
        i1 = index;
      }
      unsigned long long t_ps = P.acq_ps - st.remaining;
      // azimuths of the two records (K:554,577), needed only for a stored pair
      const float phi = mcf_atan2f(position[1], position[0]);
      const float phi1 = mcf_atan2f(positionFirst[1], positionFirst[0]);
      if (i0 < P.psf_size)
        write_psf((DU_GLOBAL mcpet_psf_record*)(P.psf + i0), t_ps, (float)(travel * K_INV_C), st.absvox, Energia1, position[2], phi, dir, ss);
      if (i1 < P.psf_size)
        write_psf((DU_GLOBAL mcpet_psf_record*)(P.psf + i1), t_ps, (float)(travelFirst * K_INV_C), st.absvox, Energia2, positionFirst[2], phi1, dirf, ssf);
    }
  }
}

// Singles list (no reference counterpart; the reference tallies pairs only): a photon that
// leaves the phantom and hits the detector cylinder within its height gets the detector's
// energy blur (the PSF path's form, K:580-587, two draws), joins the singles spectrum and,
// inside the energy window, is appended with its detection time = emission time + flight
// time to the cylinder.
template <class CN>
__device__ __forceinline__ void record_single(const KParams& P, const float* pos, const float* dir, float energy, float travel,
                                              int ss, int kind, Stream& st, CN& C, int* s_spec)
{
  const float x0 = pos[0] - P.cx, y0 = pos[1] - P.cy;
  const float dist = cyl_dist(P, x0, y0, dir);
  const float z = (pos[2] - P.cz) + dist * dir[2];
  if (!(fabsf(z) < 0.5f * P.height)) return;
  const float randno1 = rn(st.r, C);
  const float randno2 = rn(st.r, C);
  const float g = sqrtf(-2.0f * mcf_logf(randno1 + 1e-8f)) * mcf_cosf((float)(2.0f * K_PI * (double)randno2));
  const float e = energy + ((P.e_resol) * (1.0f / 2.35f)) * energy * g;
  const int iE = f2i_rd(e / 1000.0f);
  if ((iE < P.ne) && (iE >= 0)) spectrum_add(P, s_spec, iE);
  if (!(e > P.e_low && e < P.e_high)) return;
  const float phi = mcf_atan2f(y0 + dist * dir[1], x0 + dist * dir[0]);
  const unsigned long long flight = d2ull_rn((double)(travel + dist) * K_INV_C);
  const unsigned long long t = (P.acq_ps - st.remaining) + (flight < (1ull << 40) ? flight : (1ull << 40) - 1ull);
  const unsigned long long k = gatomic_add(P.singles_n, 1ull);
  if ((long long)k < P.singles_cap) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    DU_GLOBAL u32x4* rec = (DU_GLOBAL u32x4*)(P.singles + k);
    rec[0] = u32x4{(unsigned)t, (unsigned)(t >> 32), (unsigned)st.event, (unsigned)(st.event >> 32)};
    rec[1] = u32x4{__float_as_uint(e), __float_as_uint(z), __float_as_uint(phi), ((unsigned)ss & 0xffffu) | ((unsigned)kind << 16)};
  }
}

__device__ __forceinline__ void tally_dose(const KParams& P, float Edep, int mat, const float* pos, unsigned long long* s_mat)
{
  if (P.material_dose) {
    atomicAdd(&s_mat[2 * mat], f2ull_rn(Edep * 1.0f));
    atomicAdd(&s_mat[2 * mat + 1], f2ull_rn(Edep * Edep));
  }
  if (P.voxel_dose) {
    short vx = (short)f2i_rd(pos[0] * P.ivx);
    short vy = (short)f2i_rd(pos[1] * P.ivy);
    short vz = (short)f2i_rd(pos[2] * P.ivz);
    if ((vx < P.roi[0]) || (vx > P.roi[1]) || (vy < P.roi[2]) || (vy > P.roi[3]) || (vz < P.roi[4]) || (vz > P.roi[5])) return;
    int DX = 1 + (int)(P.roi[1] - P.roi[0]);
    int num = (int)(vx - P.roi[0]) + ((int)(vy - P.roi[2])) * DX + ((int)(vz - P.roi[4])) * DX * (1 + (int)(P.roi[3] - P.roi[2]));
    gatomic_add(&P.voxel_edep[2 * (size_t)num], f2ull_rn(Edep * 1.0f));
    gatomic_add(&P.voxel_edep[2 * (size_t)num + 1], f2ull_rn(Edep * Edep));
  }
}

// ---------------------------------------------------------------------------------- photon transport
// Transport is split into three per-lane building blocks so that the same arithmetic
// serves the nested (plain) kernel and the state-machine (persistent) kernel:
// photon_start (energy bin, Woodcock mean free path), woodcock_step (one delta-tracking
// step, K:319-342) and real_interaction (K:348-413).

struct Photon {
  float pos[3], dir[3];
  float E, travel, mfpW, mfp_density, prob, randno;
  float2 tab;            // (a, b) of the total cross section at (index, material of mat_old)
  int index;
  int mat_old;           // voxel material word of the last step (< 0: reload the cross section)
  int mat0, ss;
};

__device__ __forceinline__ void photon_start(const KParams& P, Photon& ph, float energy = K_E511)
{
  ph.E = energy;
  ph.ss = 0;
  ph.travel = 0.0f;
  ph.index = f2i_rd((ph.E - P.e0) * P.ide);
  const float2 w = gld(P.woodcock, ph.index);
  ph.mfpW = w.x + ph.E * w.y;
  ph.mat_old = -1;
}

enum StepOutcome { STEP_ESCAPED = 0, STEP_VIRTUAL = 1, STEP_REAL = 2 };

// Palette of (density, material) words for u8-indexed phantoms, filled by each
// workgroup's prologue.
__shared__ float2 k_palette[256];
// The total-cross-section row (a, b) of every material slot at the annihilation energy's bin,
// copied to LDS per workgroup: unscattered photons (most Woodcock steps) reload (a, b) from
// here on a table change instead of from global memory. The same values as P.tot's row.
__shared__ float2 k_tot511[K_MAXMAT];
__shared__ int k_idx511;

// Fill k_tot511 / k_idx511 (all lanes of the workgroup; a barrier must follow).
__device__ __forceinline__ void load_tot511(const KParams& P)
{
  const int idx = f2i_rd((K_E511 - P.e0) * P.ide);
  if (threadIdx.x < K_MAXMAT) k_tot511[threadIdx.x] = gld(P.tot, idx * K_MAXMAT + threadIdx.x);
  if (threadIdx.x == 0) k_idx511 = P.lds_xs511 ? idx : -1;   // -1: no energy bin matches, rows from global memory
}

// The (density, material) word of voxel absvox in the device layout chosen at set-up.
__device__ __forceinline__ float2 voxel_word(const KParams& P, int absvox)
{
  if (P.vox_fmt == 1) {
    const unsigned i = gld(P.vox8, absvox);
    return P.lds_palette ? k_palette[i] : gld(P.palette, i);
  }
  if (P.vox_fmt == 2) return gld(P.palette, gld(P.vox16, absvox));
  return gld(P.vox, absvox);
}

template <class CN>
__device__ __forceinline__ int woodcock_step(const KParams& P, Photon& ph, Rng& r, CN& C)
{
  const float step = -(ph.mfpW)*mcf_logf(rn(r, C));
  C.inc(KC_WOODCOCK);
  ph.travel += step;
  ph.pos[0] += step * ph.dir[0];
  ph.pos[1] += step * ph.dir[1];
  ph.pos[2] += step * ph.dir[2];
  const int absvox = locate(P, ph.pos);
  if (absvox < 0) return STEP_ESCAPED;
  const float2 md = voxel_word(P, absvox);
  // material word: slot (bits 0-7) | id of its total cross-section table (bits 8-15); slots
  // whose tables are bitwise identical share an id, so a change between them (a material
  // change of the reference, K:331-336) keeps the (a, b) already loaded: the same values
  const int w = __float_as_int(md.y);
  ph.mat0 = w & 0xff;
  if (w != ph.mat_old) {
    if (ph.mat_old < 0 || (w >> 8) != (ph.mat_old >> 8))
      ph.tab = ph.index == k_idx511 ? k_tot511[ph.mat0] : gld(P.tot, ph.index * K_MAXMAT + ph.mat0);
    ph.mat_old = w;
    C.inc(KC_MATCHANGE);
  }
  ph.mfp_density = ph.mfpW * md.x;
  ph.prob = 1.0f - ph.mfp_density * (ph.tab.x + ph.E * ph.tab.y);
  ph.randno = rn(r, C);
  return (ph.randno < ph.prob) ? STEP_VIRTUAL : STEP_REAL;
}

// Real event at the photon's position (K:348-413), in two halves around GCOa's loops:
// interaction_begin picks the process; Rayleigh and photoelectric events complete there
// (with their dose), a Compton event starts GCOa and is completed by compton_end after
// its rejection loops. interaction_begin returns IA_ALIVE, IA_ABSORBED or IA_COMPTON.
enum InteractionOutcome { IA_ALIVE = 0, IA_ABSORBED = 1, IA_COMPTON = 2 };

template <class CN>
__device__ __forceinline__ int interaction_begin(const KParams& P, Photon& ph, Rng& r, CN& C,
                                                 const mcpet_compton_table& cg, ComptonState& cs, unsigned long long* s_mat)
{
  const float4 cr = gld(P.cr, ph.index * K_MAXMAT + ph.mat0);
  float prob = ph.prob + ph.mfp_density * (cr.x + ph.E * cr.y);
  const float randno = ph.randno;
  if (randno < prob) {
    C.inc(KC_COMPTON);
    gcoa_init(ph.E, ph.mat0, cg, cs);
    return IA_COMPTON;
  }
  prob += ph.mfp_density * (cr.z + ph.E * cr.w);
  if (randno < prob) {
    double costh_Rayleigh;
    C.inc(KC_RAYLEIGH);
    const float pmax_current = gld(P.ray->pmax, (ph.index + 1) * K_MAXMAT + ph.mat0);
    graa(ph.E, &costh_Rayleigh, ph.mat0, pmax_current, r, C, P.ray);
    rotate_double(ph.dir, costh_Rayleigh, 6.28318530717958647693 * rnd(r, C));
    ph.ss = (ph.ss == 0) ? 2 : 3;
    return IA_ALIVE;
  }
  C.inc(KC_PHOTO);
  const float randno_pe = -ph.E;
  ph.index = -11;
  if (randno_pe < -0.001f) tally_dose(P, -1.0f * randno_pe, ph.mat0, ph.pos, s_mat);
  return IA_ABSORBED;
}

// Completion of a Compton event once both GCOa loops are done: new energy and direction,
// energy bin, Woodcock mean free path, scatter state and dose. False if absorbed (energy
// below the tables).
template <class CN>
__device__ __forceinline__ bool compton_end(const KParams& P, Photon& ph, Rng& r, CN& C, const ComptonState& cs,
                                            unsigned long long* s_mat)
{
  const float e_before = ph.E;
  const double costh_Compton = gcoa_finish(&ph.E, cs);
  rotate_double(ph.dir, costh_Compton, 6.28318530717958647693 * rnd(r, C));
  const float randno = ph.E - e_before;
  ph.index = f2i_rd((ph.E - P.e0) * P.ide);
  if (ph.index > -1) {
    const float2 w = gld(P.woodcock, ph.index);
    ph.mfpW = w.x + ph.E * w.y;
    ph.mat_old = -2;
    ph.ss = (ph.ss == 0) ? 1 : 3;
  }
  if (randno < -0.001f) tally_dose(P, -1.0f * randno, ph.mat0, ph.pos, s_mat);
  return ph.index >= 0;
}

// Whole real event, nested form. False when the photon is absorbed.
template <class CN>
__device__ __forceinline__ bool real_interaction(const KParams& P, Photon& ph, Rng& r, CN& C, const mcpet_compton_table& cg,
                                                 unsigned long long* s_mat)
{
  ComptonState cs;
  const int o = interaction_begin(P, ph, r, C, cg, cs, s_mat);
  if (o != IA_COMPTON) return o == IA_ALIVE;
  while (!gcoa_tau_trial(ph.E, ph.mat0, r, C, cg, cs)) {
  }
  cs.trials = 0;
  while (!gcoa_pz_trial(ph.E, ph.mat0, r, C, cg, cs)) {
  }
  return compton_end(P, ph, r, C, cs, s_mat);
}

// Whole photon (K:297-415), nested loops: escape or absorption.
template <class CN>
__device__ __forceinline__ void transport_photon(const KParams& P, Photon& ph, Stream& st, CN& C,
                                                 const mcpet_compton_table& cg, unsigned long long* s_mat)
{
  for (;;) {
    if (locate(P, ph.pos) < 0) return;
    int o;
    do {
      o = woodcock_step(P, ph, st.r, C);
    } while (o == STEP_VIRTUAL);
    if (o == STEP_ESCAPED) return;
    if (!real_interaction(P, ph, st.r, C, cg, s_mat)) return;
  }
}

// Pair emission at a uniform point of the stream's voxel (K:665-773).
// q = n / d for 0 <= n < 2^31 and 1 <= d < 2^20 with inv = 1/d in double: the product's
// error is below 2^-21 < 1/d, so its truncation is q, or q - 1 when n / d is an integer and the
// product rounds below it; one comparison fixes that. Replaces a 32-bit integer division (tens
// of instructions). Grid dimensions are far below 2^20.
__device__ __forceinline__ int div_by_inv(int n, int d, double inv)
{
  int q = (int)((double)n * inv);
  if (n - q * d >= d) q++;
  return q;
}
// (x, y, z) of voxel index v (x fastest), the reference's v % nx, (v / nx) % ny, v / (nx ny).
__device__ __forceinline__ void voxel_coords(const KParams& P, int v, int& ix, int& iy, int& iz)
{
  const int q = div_by_inv(v, P.nx, P.inv_nx);
  iz = div_by_inv(q, P.ny, P.inv_ny);
  iy = q - iz * P.ny;
  ix = v - q * P.nx;
}

template <class CN>
__device__ __forceinline__ void source_first(const KParams& P, Stream& st, CN& C, float* pos, float* dir)
{
  dir[2] = 1.0f - 2.0f * rn(st.r, C);
  float phi_sampled = 6.28318530716f * rn(st.r, C);
  float sin_theta = sqrtf(1.0f - dir[2] * dir[2]);
  float sp, cp;
  mcf_sincosf(phi_sampled, &sp, &cp);
  dir[1] = sin_theta * sp;
  dir[0] = sin_theta * cp;
  float r1 = rn(st.r, C);
  float r2 = rn(st.r, C);
  float r3 = rn(st.r, C);
  r1 = (K_EPS > r1 - K_EPS) ? K_EPS : r1 - K_EPS;
  r2 = (K_EPS > r2 - K_EPS) ? K_EPS : r2 - K_EPS;
  r3 = (K_EPS > r3 - K_EPS) ? K_EPS : r3 - K_EPS;
  int ix, iy, iz;
  voxel_coords(P, st.absvox, ix, iy, iz);
  pos[0] = ((float)ix + r1) / P.ivx;
  pos[1] = ((float)iy + r2) / P.ivy;
  pos[2] = ((float)iz + r3) / P.ivz;
}

// Isotropic direction (K:711-720, the BG-mode second photon).
template <class CN>
__device__ __forceinline__ void isotropic_dir(Stream& st, CN& C, float* dir)
{
  dir[2] = 1.0f - 2.0f * rn(st.r, C);
  float phi_sampled = 6.28318530716f * rn(st.r, C);
  float sin_theta = sqrtf(1.0f - dir[2] * dir[2]);
  float sp, cp;
  mcf_sincosf(phi_sampled, &sp, &cp);
  dir[1] = sin_theta * sp;
  dir[0] = sin_theta * cp;
}

template <class CN>
__device__ __forceinline__ void source_second(const KParams& P, Stream& st, CN& C, const float* dir1, float* dir)
{
  if (P.bg_mode) {
    isotropic_dir(st, C, dir);
    return;
  }
  float phi_sampled2 = (float)((double)mcf_atan2f(dir1[1], dir1[0]) + K_PI);
  float theta_sampled2 = (float)(K_PI - (double)mcf_acosf(dir1[2]));
  float V1 = rn(st.r, C);
  float V2 = rn(st.r, C);
  float THETA_NC = (float)((2.0 * 0.21276 * K_PI / 180.0) * (double)sqrtf(-2.0f * mcf_logf(V2 + 1e-8f)) *
                           (double)mcf_cosf((float)((2.0 * K_PI) * (double)V1)));
  float PHI_NC = (float)((2.0 * K_PI) * (double)rn(st.r, C));
  float ST, CT, CP, SP, ST1, CT1, CP1, SP1;
  mcf_sincosf(THETA_NC, &ST, &CT);
  mcf_sincosf(PHI_NC, &SP, &CP);
  mcf_sincosf(theta_sampled2, &ST1, &CT1);
  mcf_sincosf(phi_sampled2, &SP1, &CP1);
  float GX = CP * ST;
  float GY = SP * ST;
  float GZ = CT;
  dir[0] = GX * CT1 * CP1 - GY * SP1 + GZ * ST1 * CP1;
  dir[1] = GX * CT1 * SP1 + GY * CP1 + GZ * ST1 * SP1;
  dir[2] = -GX * ST1 + GZ * CT1;
}

// Time to the next annihilation of the stream (K:261-292). False when the stream's
// acquisition time is exhausted.
template <class CN>
__device__ __forceinline__ bool decay_time(const KParams& P, Stream& st, CN& C)
{
  const double Dt = -mcf_log(rnd(st.r, C)) * st.inv_act;
  const unsigned long long Dt_ps = d2ull_rn(Dt * 1e12);
  if (!(Dt_ps < st.remaining)) return false;
  st.remaining -= Dt_ps;
  st.event++;
  C.inc(KC_DECAYS);              // 2 photon histories per decay (K:282)
  st.inv_act = st.inv_act * mcf_exp(Dt * P.inv_mean_life);
  return true;
}

// Pair state kept between the two photons of an annihilation (K:200-203, 418-441).
struct Pair {
  float p1[3], d1[3];              // emission point and photon-1 direction
  float EF, travelF;               // photon 1 at escape (EF = -10 if absorbed)
  float posF[3], dirF[3];
  int ssF;
};

__device__ __forceinline__ bool pair_selected(const KParams& P, int ss, int ssF)
{
  return P.type == 0 || (P.type == 1 && ss == 0 && ssF == 0) || (P.type == 2 && (ss != 0 || ssF != 0));
}

template <class CN>
__device__ __forceinline__ void tally_pair(const KParams& P, const Photon& ph, const Pair& pr, Stream& st, CN& C, int* s_spec)
{
  if (P.mode == 0 || P.mode == 2) tally_sinogram(P, ph.E, pr.EF, ph.pos, pr.posF, ph.dir, pr.dirF, ph.ss, pr.ssF, st, C, s_spec);
  if (P.mode == 0 || P.mode == 1)
    tally_psf(P, ph.E, pr.EF, ph.pos, pr.posF, ph.dir, pr.dirF, ph.ss, pr.ssF, ph.travel, pr.travelF, st, C);
}

__device__ __forceinline__ void keep_first(Pair& pr, const Photon& ph)
{
  if (ph.index > -1) {
    pr.EF = ph.E;
    pr.travelF = ph.travel;
    pr.ssF = ph.ss;
    for (int k = 0; k < 3; k++) {
      pr.posF[k] = ph.pos[k];
      pr.dirF[k] = ph.dir[k];
    }
  }
}

// One annihilation, both photons, nested loops (reference main loop body, K:259-449).
template <class CN>
__device__ bool decay(const KParams& P, Stream& st, CN& C, const mcpet_compton_table& cg, int* s_spec,
                      unsigned long long* s_mat)
{
  if (!decay_time(P, st, C)) return false;
  C.inc(KC_EMIT_PAIR);
  Pair pr;
  Photon ph;
  source_first(P, st, C, pr.p1, pr.d1);
  pr.EF = -10.0f;
  for (int k = 0; k < 3; k++) {
    ph.pos[k] = pr.p1[k];
    ph.dir[k] = pr.d1[k];
  }
  photon_start(P, ph);
  transport_photon(P, ph, st, C, cg, s_mat);
  keep_first(pr, ph);
  source_second(P, st, C, pr.d1, ph.dir);
  for (int k = 0; k < 3; k++) ph.pos[k] = pr.p1[k];
  photon_start(P, ph);
  transport_photon(P, ph, st, C, cg, s_mat);
  if (ph.index > -1 && pr.EF > -0.5f && pair_selected(P, ph.ss, pr.ssF)) tally_pair(P, ph, pr, st, C, s_spec);
  return true;
}

__device__ __forceinline__ void stream_init(const KParams& P, long long s, Stream& st)
{
  int k;
  long long tid, per_voxel;
  if (P.stream_start) {   // balanced layout: the voxel owning stream s (binary search)
    int lo = 0, hi = P.n_active - 1;
    if (P.stream_hint) {  // narrowed to the owners of the first streams of s's block and the next
      const long long b = s >> P.hint_shift;
      lo = gld(P.stream_hint, b);
      hi = gld(P.stream_hint, b + 1);
    }
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (gld(P.stream_start, mid) <= s) lo = mid;
      else hi = mid - 1;
    }
    k = lo;
    const long long s0 = gld(P.stream_start, k);
    per_voxel = gld(P.stream_start, k + 1) - s0;
    tid = gld(P.stream_id, k) + (s - s0);
  } else {
    k = (int)(s / P.tpb);
    per_voxel = P.tpb;
    tid = (long long)gld(P.active_vox, k) * P.tpb + (s - (long long)k * P.tpb);
  }
  st.absvox = gld(P.active_vox, k);
  const float act = gld(P.active_act, k);
  mcr_init_stream(tid, P.seed, P.j1, P.j2, &st.r.x, &st.r.y);
  st.event = (unsigned long long)tid << 32;
  st.remaining = P.acq_ps;
  st.inv_act = ((double)per_voxel / (double)act);
}

__device__ __forceinline__ void stream_finish(const KParams& P, const Stream& st)
{
  if ((long long)(st.event >> 32) == P.last_thread_id) *(DU_GLOBAL int*)P.final_seed = st.r.x;
}

// Workgroup prologue/epilogue shared by the kernels: Compton table, spectrum, material
// dose and workgroup counters in LDS; counters flushed once per lane / per workgroup.
#define K_PROLOGUE(CN)                                                                           \
  __shared__ mcpet_compton_table s_cmp;                                                          \
  __shared__ unsigned long long s_mat[2 * K_MAXMAT];                                             \
  __shared__ unsigned s_cnt[KC_N];                                                               \
  extern __shared__ int s_spec[];                                                                \
  {                                                                                              \
    const int* src = reinterpret_cast<const int*>(P.cmp);                                        \
    int* dst = reinterpret_cast<int*>(&s_cmp);                                                   \
    for (int i = threadIdx.x; i < (int)(sizeof(mcpet_compton_table) / 4); i += blockDim.x)       \
      dst[i] = gld(src, i);                                                                      \
    if (threadIdx.x < 2 * K_MAXMAT) s_mat[threadIdx.x] = 0ull;                                   \
    if (threadIdx.x < KC_N) s_cnt[threadIdx.x] = 0u;                                             \
    if (P.vox_fmt == 1 && P.lds_palette)                                                         \
      for (int i = threadIdx.x; i < P.n_palette; i += blockDim.x) k_palette[i] = gld(P.palette, i); \
    load_tot511(P);                                                                              \
    if (P.lds_spectrum)                                                                          \
      for (int i = threadIdx.x; i < P.ne; i += blockDim.x) s_spec[i] = 0;                        \
  }                                                                                              \
  __syncthreads();                                                                               \
  CN C;                                                                                          \
  C.init(s_cnt);

#define K_EPILOGUE                                                                               \
  C.flush(P, s_cnt);                                                                             \
  __syncthreads();                                                                               \
  if (threadIdx.x < KC_N && s_cnt[threadIdx.x])                                                  \
    gatomic_add(&P.counters[threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);               \
  if (P.lds_spectrum)                                                                            \
    for (int i = threadIdx.x; i < P.ne; i += blockDim.x)                                         \
      if (s_spec[i]) gatomic_add(&P.spectrum[i], s_spec[i]);                                     \
  if (threadIdx.x < 2 * K_MAXMAT && s_mat[threadIdx.x]) gatomic_add(&P.mat_edep[threadIdx.x], s_mat[threadIdx.x]);

// Plain kernel: lane i of the grid runs stream i to the end (nested loops, no queue).
__global__ __launch_bounds__(256) K_OCCUPANCY void k_transport_plain(KParams P)
{
  K_PROLOGUE(FullCounters)
  const long long sid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (sid < P.n_streams) {
    Stream st;
    stream_init(P, sid, st);
    C.inc(KC_STREAMS);
    while (decay(P, st, C, s_cmp, s_spec, s_mat)) {
    }
    stream_finish(P, st);
  }
  K_EPILOGUE
}

// Persistent kernel: every lane is a small state machine over the phases of a decay
//   DECAY    : (refill a stream from the device queue,) sample the decay time, emit photon 1
//   STEP     : one Woodcock step of the current photon
//   INTERACT : one real interaction (Compton / Rayleigh / photoelectric + dose)
//   SOURCE2  : emit photon 2 of the pair
//   TALLY    : detector, energy blur, sinogram / PSF tallies of the pair
//   CTAU     : one trial of GCOa's tau loop
//   CPZ      : one trial of GCOa's p_z loop (+ completion of the Compton event)
//   DETECT   : singles list: detector hit, energy blur and record of the photon just ended
//   SOURCE3  : emit the decay's prompt gamma (isotropic)
// and each iteration the wave executes the one phase held by the most lanes (ballot
// counts). Lanes in other phases wait, so expensive phases (Compton rejection trials,
// detector geometry) run with many lanes at once instead of one or two, and a lane that
// needs more rejection trials than its neighbours does not hold them back. The phases that
// follow the selected one along a photon's path (GCOa's trials after an interaction, Woodcock
// steps after a Compton event or an emission, the next decay after a pair tally) run in the
// same iteration for every lane that has reached them. A lane's own sequence of random
// numbers and operations is unchanged.
enum LaneState {
  LS_DECAY = 0, LS_SOURCE2 = 1, LS_STEP = 2, LS_INTERACT = 3, LS_TALLY = 4, LS_CTAU = 5, LS_CPZ = 6, LS_DETECT = 7,
  LS_SOURCE3 = 8, LS_N = 9, LS_DONE = 10
};

// Per-lane state that only some phases touch lives in LDS slots instead of registers
// (structure of arrays, slot k of lane i at [k][i]: consecutive lanes, consecutive banks):
// the pair kept between the two photons (15 words) and the Compton sampler's state
// between its rejection trials (14 words). Only the stream and the current photon stay
// in registers across iterations, which keeps the kernel at 3 waves per SIMD.
#define K_BLOCK 256
enum PairSlot { PS_P1 = 0, PS_D1 = 3, PS_EF = 6, PS_TRAVELF = 7, PS_POSF = 8, PS_DIRF = 11, PS_SSF = 14, PS_N = 15 };
enum ComptonSlot {
  CS_A1 = PS_N, CS_S0, CS_TAUMIN, CS_TAU, CS_S, CS_CDTF, CS_RC0, CS_RC1, CS_RC2, CS_RC3, CS_PZOMC, CS_CDT1_LO, CS_CDT1_HI,
  CS_TRIALS, LANE_SLOTS
};

struct LaneSlots {
  float* p;  // &slots[0][lane]
  __device__ __forceinline__ float& operator[](int k) const { return p[k * K_BLOCK]; }
};

__device__ __forceinline__ void pair_store_source(const LaneSlots& L, const Pair& pr)
{
  for (int k = 0; k < 3; k++) {
    L[PS_P1 + k] = pr.p1[k];
    L[PS_D1 + k] = pr.d1[k];
  }
  L[PS_EF] = pr.EF;
}
__device__ __forceinline__ void pair_store_first(const LaneSlots& L, const Photon& ph)
{
  Pair pr;
  pr.EF = L[PS_EF];
  keep_first(pr, ph);
  if (ph.index > -1) {
    L[PS_EF] = pr.EF;
    L[PS_TRAVELF] = pr.travelF;
    L[PS_SSF] = __int_as_float(pr.ssF);
    for (int k = 0; k < 3; k++) {
      L[PS_POSF + k] = pr.posF[k];
      L[PS_DIRF + k] = pr.dirF[k];
    }
  }
}
__device__ __forceinline__ void pair_load(const LaneSlots& L, Pair& pr)
{
  for (int k = 0; k < 3; k++) {
    pr.p1[k] = L[PS_P1 + k];
    pr.d1[k] = L[PS_D1 + k];
    pr.posF[k] = L[PS_POSF + k];
    pr.dirF[k] = L[PS_DIRF + k];
  }
  pr.EF = L[PS_EF];
  pr.travelF = L[PS_TRAVELF];
  pr.ssF = __float_as_int(L[PS_SSF]);
}
__device__ __forceinline__ void cs_store(const LaneSlots& L, const ComptonState& cs)
{
  L[CS_A1] = cs.a1;
  L[CS_S0] = cs.s0;
  L[CS_TAUMIN] = cs.taumin;
  L[CS_TAU] = cs.tau;
  L[CS_S] = cs.s;
  L[CS_CDTF] = cs.cdtf;
  L[CS_RC0] = cs.rc0;
  L[CS_RC1] = cs.rc1;
  L[CS_RC2] = cs.rc2;
  L[CS_RC3] = cs.rc3;
  L[CS_PZOMC] = cs.pzomc;
  const unsigned long long b = (unsigned long long)__double_as_longlong(cs.cdt1);
  L[CS_CDT1_LO] = __uint_as_float((unsigned)b);
  L[CS_CDT1_HI] = __uint_as_float((unsigned)(b >> 32));
  L[CS_TRIALS] = __int_as_float(cs.trials);
}
__device__ __forceinline__ void cs_load(const LaneSlots& L, ComptonState& cs)
{
  cs.a1 = L[CS_A1];
  cs.s0 = L[CS_S0];
  cs.taumin = L[CS_TAUMIN];
  cs.tau = L[CS_TAU];
  cs.s = L[CS_S];
  cs.cdtf = L[CS_CDTF];
  cs.rc0 = L[CS_RC0];
  cs.rc1 = L[CS_RC1];
  cs.rc2 = L[CS_RC2];
  cs.rc3 = L[CS_RC3];
  cs.pzomc = L[CS_PZOMC];
  const unsigned long long b = (unsigned long long)__float_as_uint(L[CS_CDT1_LO]) |
                               ((unsigned long long)__float_as_uint(L[CS_CDT1_HI]) << 32);
  cs.cdt1 = __longlong_as_double((long long)b);
  cs.trials = __float_as_int(L[CS_TRIALS]);
}

// Next state once the current photon has escaped or was absorbed (K:418-442). Photons are
// numbered 1, 2 (annihilation pair) and 3 (prompt gamma); the singles pipeline records every
// photon in DETECT instead of pairing them here.
__device__ __forceinline__ int after_photon(const KParams& P, int photon_no, bool prompt, const Photon& ph, const LaneSlots& L)
{
  if (P.mode == MCPET_TALLY_SINGLES) return LS_DETECT;
  if (photon_no == 1) {
    pair_store_first(L, ph);
    return LS_SOURCE2;
  }
  if (photon_no == 2 && ph.index > -1 && L[PS_EF] > -0.5f && pair_selected(P, ph.ss, __float_as_int(L[PS_SSF])))
    return LS_TALLY;
  return (photon_no == 2 && prompt) ? LS_SOURCE3 : LS_DECAY;
}

// The state machine of one wave on one phantom, until the phantom's stream queue is empty
// and every lane of the wave has finished its stream. kExt: the singles list or prompt
// gammas are in use (phases DETECT and SOURCE3 exist); otherwise the wave only ballots
// over the seven phases of the reference's pair tallies.
// Streams are claimed from the queue by the wave, K_CHUNK indices per atomic, and handed to
// the lanes that need one in lane order; the wave's unused range (next index wq[0], count
// wq[1]) lives in LDS
// (a single global counter updated once per stream was the limit of short runs). The range
// is read by every lane and written by the lowest one; volatile keeps each access an LDS
// operation in program order, which a wave executes in order.
#define K_CHUNK 64
// Index (from the queue's origin) of the next stream for every calling lane. Not inlined:
// inlined into the state machine it raised the register pressure of every phase (32 more
// spilled VGPRs); as a call it costs a few register saves at each claim.
__device__ __attribute__((noinline)) long long claim_stream(unsigned long long* queue, volatile K_LDS long long* wq)
{
  // the lanes here all need a stream: rank among them, claim a new chunk if short
  // (the range is wave-uniform: read into scalar registers)
  const unsigned long long need = __ballot(1);
  const int cnt = __popcll(need);
  const int rank = __popcll(need & ((1ull << __lane_id()) - 1ull));
  const long long q0 = wq[0];
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)q0);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)q0 >> 32));
  const long long qn = (long long)(((unsigned long long)hi << 32) | lo);
  const int avail = __builtin_amdgcn_readfirstlane((int)wq[1]);
  long long idx;
  if (avail >= cnt) {
    idx = qn + rank;
    if (rank == 0) {
      wq[0] = qn + cnt;
      wq[1] = avail - cnt;
    }
  } else {
    unsigned long long base = 0;
    if (rank == 0) base = gatomic_add(queue, (unsigned long long)K_CHUNK);
    const unsigned blo = __builtin_amdgcn_readfirstlane((unsigned)base);
    const unsigned bhi = __builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
    const long long b = (long long)(((unsigned long long)bhi << 32) | blo);
    idx = rank < avail ? qn + rank : b + (rank - avail);
    if (rank == 0) {
      wq[0] = b + (cnt - avail);
      wq[1] = K_CHUNK - (cnt - avail);
    }
  }
  return idx;
}

template <class CN, bool kExt>
__device__ __forceinline__ void sm_waves(const KParams& P, CN& C, const mcpet_compton_table& s_cmp, int* s_spec,
                                         unsigned long long* s_mat, const LaneSlots& L, volatile K_LDS long long* wq)
{
  constexpr int kPhases = kExt ? LS_N : LS_CPZ + 1;
  if (__lane_id() == 0) {
    wq[0] = 0;
    wq[1] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  Stream st;
  Photon ph;
  bool have = false;
  bool prompt = false;           // the current decay still has its prompt gamma to emit
  int photon_no = 1;
  int state = LS_DECAY;
  for (;;) {
    int best = -1;
    if (P.sm_policy == 0) {
      int bestc = 0;
#pragma unroll
      for (int sidx = 0; sidx < kPhases; sidx++) {
        const int c = __popcll(__ballot(state == sidx));
        if (c > bestc) {
          bestc = c;
          best = sidx;
        }
      }
    } else {
      int bestw = 0x7fffffff;
#pragma unroll
      for (int sidx = 0; sidx < kPhases; sidx++) {
        const int c = __popcll(__ballot(state == sidx));
        const int w = (64 - c) * P.sm_cost[sidx];
        if (c > 0 && w < bestw) {
          bestw = w;
          best = sidx;
        }
      }
    }
    best = __builtin_amdgcn_readfirstlane(best);
    if (best < 0) break;
#if MCPET_SM_STATS
    const long long t_phase = clock64();
    {
      const int lanes = __popcll(__ballot(state == best));
      if (__lane_id() == 0) {
        C.add(KC_PHASE_EXEC + best, 1u);
        C.add(KC_PHASE_LANES + best, (unsigned)lanes);
      }
    }
#endif
    // The phases as blocks, each run for the lanes in its state when the wave selected it. With
    // chaining (MCPET_CHAIN) a block also runs right after the block that feeds it, for every
    // lane then in its state (the lanes that just arrived and those already waiting): a
    // Compton event's trials and the photon's next Woodcock steps follow its interaction
    // without a phase selection in between. The block order (a chained block comes after its
    // feeder) is the one with the fewest spilled registers among those measured.
    const bool run_int = best == LS_INTERACT;
    if (run_int && state == LS_INTERACT) {
      ComptonState cs;
      ph.prob = L[CS_TAU];
      ph.mfp_density = L[CS_S];
      ph.randno = L[CS_CDTF];
      const int o = interaction_begin(P, ph, st.r, C, s_cmp, cs, s_mat);
      if (o == IA_COMPTON) {
        cs_store(L, cs);
        state = LS_CTAU;
      } else if (o == IA_ABSORBED || locate(P, ph.pos) < 0) {
        state = after_photon(P, photon_no, prompt, ph, L);
      } else {
        state = LS_STEP;
      }
    }
    const bool run_tau = best == LS_CTAU || ((MCPET_CHAIN & 1) && run_int);
    if (run_tau && state == LS_CTAU) {
      ComptonState cs;
      cs_load(L, cs);
      bool ok;
      int k = 0;
      do {
        ok = gcoa_tau_trial(ph.E, ph.mat0, st.r, C, s_cmp, cs);
      } while (!ok && ++k < MCPET_TRIALS_PER_PHASE);
      if (ok) {
        cs.trials = 0;
        state = LS_CPZ;
      }
      cs_store(L, cs);
    }
    const bool run_pz = best == LS_CPZ || ((MCPET_CHAIN & 2) && run_tau);
    if (run_pz && state == LS_CPZ) {
      ComptonState cs;
      cs_load(L, cs);
      bool ok;
      int k = 0;
      do {
        ok = gcoa_pz_trial(ph.E, ph.mat0, st.r, C, s_cmp, cs);
      } while (!ok && ++k < MCPET_TRIALS_PER_PHASE);
      if (ok) {
        if (!compton_end(P, ph, st.r, C, cs, s_mat) || locate(P, ph.pos) < 0) state = after_photon(P, photon_no, prompt, ph, L);
        else state = LS_STEP;
      } else {
        L[CS_TRIALS] = __int_as_float(cs.trials);
      }
    }
    const bool run_tally = best == LS_TALLY;
    if (run_tally && state == LS_TALLY) {
      Pair pr;
      pair_load(L, pr);
      tally_pair(P, ph, pr, st, C, s_spec);
      state = prompt ? LS_SOURCE3 : LS_DECAY;
    }
    const bool run_decay = best == LS_DECAY || ((MCPET_CHAIN & 16) && run_tally);
    if (run_decay && state == LS_DECAY) {
      for (;;) {
        if (!have) {
          const long long idx = claim_stream(P.queue, wq);
          const long long sid = P.stream_base + (P.stream_stride > 1 ? idx * P.stream_stride : idx);
          if (sid >= P.n_streams) {
            state = LS_DONE;
            break;
          }
          stream_init(P, sid, st);
          C.inc(KC_STREAMS);
          have = true;
        }
        if (decay_time(P, st, C)) break;
        stream_finish(P, st);
        have = false;
      }
      if (state != LS_DONE) {
        // decay branches (no draws for a pure positron emitter without prompt gamma)
        bool pair = true;
        if (P.positron_frac < 1.0f) pair = rn(st.r, C) < P.positron_frac;
        prompt = false;
        if (P.prompt_frac > 0.0f) prompt = rn(st.r, C) < P.prompt_frac;
        Pair pr;
        source_first(P, st, C, pr.p1, pr.d1);
        pr.EF = -10.0f;
        pair_store_source(L, pr);
        if (prompt) C.inc(KC_EMIT_PROMPT);
        if (pair) {
          C.inc(KC_EMIT_PAIR);
          for (int k = 0; k < 3; k++) {
            ph.pos[k] = pr.p1[k];
            ph.dir[k] = pr.d1[k];
          }
          photon_start(P, ph);
          photon_no = 1;
          state = (locate(P, ph.pos) < 0) ? after_photon(P, 1, prompt, ph, L) : LS_STEP;
        } else {
          state = prompt ? LS_SOURCE3 : LS_DECAY;
        }
      }
    }
    const bool run_src2 = best == LS_SOURCE2;
    if (run_src2 && state == LS_SOURCE2) {
      float p1[3], d1[3];
      for (int k = 0; k < 3; k++) {
        p1[k] = L[PS_P1 + k];
        d1[k] = L[PS_D1 + k];
      }
      source_second(P, st, C, d1, ph.dir);
      for (int k = 0; k < 3; k++) ph.pos[k] = p1[k];
      photon_start(P, ph);
      photon_no = 2;
      state = (locate(P, ph.pos) < 0) ? after_photon(P, 2, prompt, ph, L) : LS_STEP;
    }
    const bool run_det = kExt && best == LS_DETECT;
    if (run_det && state == LS_DETECT) {
      if (ph.index > -1) record_single(P, ph.pos, ph.dir, ph.E, ph.travel, ph.ss, photon_no - 1, st, C, s_spec);
      state = (photon_no == 1) ? LS_SOURCE2 : (photon_no == 2 && prompt) ? LS_SOURCE3 : LS_DECAY;
    }
    const bool run_src3 = kExt && best == LS_SOURCE3;
    if (run_src3 && state == LS_SOURCE3) {
      for (int k = 0; k < 3; k++) ph.pos[k] = L[PS_P1 + k];
      isotropic_dir(st, C, ph.dir);
      photon_start(P, ph, P.prompt_energy);
      photon_no = 3;
      prompt = false;
      state = (locate(P, ph.pos) < 0) ? after_photon(P, 3, false, ph, L) : LS_STEP;
    }
    const bool run_step = best == LS_STEP || ((MCPET_CHAIN & 4) && run_pz) ||
                          ((MCPET_CHAIN & 8) && (run_decay || run_src2 || run_src3));
    if (run_step && state == LS_STEP) {
      int o, k = 0;
      do {
        o = woodcock_step(P, ph, st.r, C);
      } while (o == STEP_VIRTUAL && ++k < MCPET_STEPS_PER_PHASE);
      if (o == STEP_ESCAPED) {
        state = after_photon(P, photon_no, prompt, ph, L);
      } else if (o == STEP_REAL) {
        // the interaction's inputs wait in LDS (Compton slots, free until GCOa starts)
        L[CS_TAU] = ph.prob;
        L[CS_S] = ph.mfp_density;
        L[CS_CDTF] = ph.randno;
        state = LS_INTERACT;
      }
    }
#if MCPET_SM_STATS
    if (__lane_id() == 0) C.add(KC_PHASE_CYCLES + best, (unsigned)(clock64() - t_phase));
#endif
  }
}

// Persistent kernel over one phantom.
template <class CN, bool kExt>
// Parameters are read through a pointer to device memory: only the fields a phase uses are
// loaded (as scalars), instead of a by-value block that would be kept in scalar registers
// (and spilled) or copied to private memory.
__global__ __launch_bounds__(K_BLOCK) K_OCCUPANCY void k_transport_sm(const KParams* __restrict__ Pd)
{
  const KParams& P = *Pd;
  K_PROLOGUE(CN)
  __shared__ float s_slots[LANE_SLOTS][K_BLOCK];
  __shared__ long long s_wq[K_BLOCK / 64][2];
  const LaneSlots L{&s_slots[0][threadIdx.x]};
  sm_waves<CN, kExt>(P, C, s_cmp, s_spec, s_mat, L, (volatile K_LDS long long*)s_wq[threadIdx.x >> 6]);
  K_EPILOGUE
}

// Batched persistent kernel over n phantoms that share their interaction tables and
// scanner (params Ps[0..n-1], one per phantom, each with its own stream queue, voxels and
// tallies). Every wave works on one phantom at a time, so all parameters stay wave-uniform:
// it pulls streams from that phantom's queue, and when the queue is empty and its lanes are
// done, it flushes its own counters and material dose (per-wave LDS slices) to that phantom
// and moves to the next one. Waves start at different phantoms and visit every phantom
// once, so the phantoms' queues drain concurrently and no phantom waits for a launch of its
// own. The spectrum goes straight to global memory and the palette is read from global
// memory (both are per phantom). Per phantom, streams and random numbers are those of a
// single-phantom run, so each phantom's tallies equal its own run's bit for bit.
template <class CN, bool kExt>
__global__ __launch_bounds__(K_BLOCK) K_OCCUPANCY void k_transport_batch(const KParams* __restrict__ Ps, int n)
{
  __shared__ mcpet_compton_table s_cmp;
  __shared__ unsigned long long s_matw[K_BLOCK / 64][2 * K_MAXMAT];
  __shared__ unsigned s_cntw[K_BLOCK / 64][KC_N];
  __shared__ float s_slots[LANE_SLOTS][K_BLOCK];
  __shared__ long long s_wq[K_BLOCK / 64][2];
  {
    const int* src = reinterpret_cast<const int*>(Ps[0].cmp);
    int* dst = reinterpret_cast<int*>(&s_cmp);
    for (int i = threadIdx.x; i < (int)(sizeof(mcpet_compton_table) / 4); i += blockDim.x) dst[i] = gld(src, i);
    for (int i = threadIdx.x; i < (K_BLOCK / 64) * 2 * K_MAXMAT; i += blockDim.x) (&s_matw[0][0])[i] = 0ull;
    for (int i = threadIdx.x; i < (K_BLOCK / 64) * KC_N; i += blockDim.x) (&s_cntw[0][0])[i] = 0u;
    load_tot511(Ps[0]);   // the phantoms of a batch share their interaction tables
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const LaneSlots L{&s_slots[0][threadIdx.x]};
  const int first = (int)(((long long)blockIdx.x * (blockDim.x >> 6) + wave) % n);
  for (int j = 0; j < n; j++) {
    const int k = __builtin_amdgcn_readfirstlane((first + j) % n);
    const KParams& P = Ps[k];
    CN C;
    C.init(s_cntw[wave]);
    sm_waves<CN, kExt>(P, C, s_cmp, nullptr, s_matw[wave], L, (volatile K_LDS long long*)s_wq[wave]);
    C.flush(P, s_cntw[wave]);
    __builtin_amdgcn_wave_barrier();
    if (lane < KC_N && s_cntw[wave][lane]) {
      gatomic_add(&P.counters[lane], (unsigned long long)s_cntw[wave][lane]);
      s_cntw[wave][lane] = 0u;
    }
    if (lane < 2 * K_MAXMAT && s_matw[wave][lane]) {
      gatomic_add(&P.mat_edep[lane], s_matw[wave][lane]);
      s_matw[wave][lane] = 0ull;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// ---------------------------------------------------------------------------------- self tests

__global__ void k_selftest_ranecu(int seed, const long long* ids, int nstreams, int n, float* out, double* out_d, int as_double,
                                  const uint32_t* j1, const uint32_t* j2)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nstreams) return;
  int sx, sy;
  mcr_init_stream(ids[i], seed, j1, j2, &sx, &sy);
  for (int k = 0; k < n; k++) {
    if (as_double) out_d[(size_t)i * n + k] = mcr_ranecu_double(&sx, &sy);
    else out[(size_t)i * n + k] = mcr_ranecu(&sx, &sy);
  }
}

__global__ void k_selftest_fmath(const double* x, const double* y, int n, double* od, float* of)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = x[i], b = y[i];
  double s, c;
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

// Clear n words (a memset that a captured graph may hold).
__global__ void k_fill_u32(unsigned* __restrict__ p, size_t n)
{
  const size_t base = (size_t)blockIdx.x * 256 * 16 + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 16; k++)
    if (base + (size_t)k * 256 < n) p[base + (size_t)k * 256] = 0u;
}

// Zero up to K_FILL_MAX regions of 4-byte words in one launch: block row y clears region y,
// its blocks grid-striding over the region's words (16-byte stores where the region allows).
struct FillRegions {
  unsigned* p[K_FILL_MAX];
  unsigned long long n[K_FILL_MAX];   // words
};
__global__ __launch_bounds__(256) void k_fill_regions(FillRegions r)
{
  unsigned* p = r.p[blockIdx.y];
  const size_t n = r.n[blockIdx.y];
  const size_t stride = (size_t)gridDim.x * 256;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (((uintptr_t)p & 15) == 0) {
    uint4* q = (uint4*)p;
    const size_t n4 = n / 4;
    for (size_t k = i; k < n4; k += stride) q[k] = uint4{0u, 0u, 0u, 0u};
    for (size_t k = n4 * 4 + i; k < n; k += stride) p[k] = 0u;
  } else {
    for (size_t k = i; k < n; k += stride) p[k] = 0u;
  }
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

hipError_t mcpet_launch_transport(const KParams& P, const KParams* d_P, int kind, int count_events, int blocks, int block_size,
                                  hipStream_t stream)
{
  if (block_size > K_BLOCK || block_size % 64 != 0) return hipErrorInvalidValue;
  size_t lds = P.lds_spectrum ? (size_t)P.ne * sizeof(int) : 0;
  const bool ext = P.mode == MCPET_TALLY_SINGLES || P.prompt_frac > 0.0f;
  if (kind == MCPET_KERNEL_PERSISTENT && count_events && ext)
    hipLaunchKernelGGL((k_transport_sm<FullCounters, true>), dim3(blocks), dim3(block_size), lds, stream, d_P);
  else if (kind == MCPET_KERNEL_PERSISTENT && count_events)
    hipLaunchKernelGGL((k_transport_sm<FullCounters, false>), dim3(blocks), dim3(block_size), lds, stream, d_P);
  else if (kind == MCPET_KERNEL_PERSISTENT && ext)
    hipLaunchKernelGGL((k_transport_sm<LiteCounters, true>), dim3(blocks), dim3(block_size), lds, stream, d_P);
  else if (kind == MCPET_KERNEL_PERSISTENT)
    hipLaunchKernelGGL((k_transport_sm<LiteCounters, false>), dim3(blocks), dim3(block_size), lds, stream, d_P);
  else
    hipLaunchKernelGGL(k_transport_plain, dim3(blocks), dim3(block_size), lds, stream, P);
  return hipGetLastError();
}

hipError_t mcpet_launch_transport_batch(const KParams* d_params, int n, bool ext, int count_events, int blocks, int block_size,
                                        hipStream_t stream)
{
  if (block_size > K_BLOCK || block_size % 64 != 0 || n < 1) return hipErrorInvalidValue;
  if (count_events && ext)
    hipLaunchKernelGGL((k_transport_batch<FullCounters, true>), dim3(blocks), dim3(block_size), 0, stream, d_params, n);
  else if (count_events)
    hipLaunchKernelGGL((k_transport_batch<FullCounters, false>), dim3(blocks), dim3(block_size), 0, stream, d_params, n);
  else if (ext)
    hipLaunchKernelGGL((k_transport_batch<LiteCounters, true>), dim3(blocks), dim3(block_size), 0, stream, d_params, n);
  else
    hipLaunchKernelGGL((k_transport_batch<LiteCounters, false>), dim3(blocks), dim3(block_size), 0, stream, d_params, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------- projections

// Dense per-voxel map for the projections: .y = mu(511 keV) = density x (a + E b) of the
// voxel's material total-cross-section line (every voxel), .x = 0 (activity added next).
__global__ void k_dense_mu(const KParams* __restrict__ Pd, float2* __restrict__ out, long long nvox)
{
  const KParams& P = *Pd;
  const int idx = f2i_rd((K_E511 - P.e0) * P.ide);
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long long)gridDim.x * blockDim.x) {
    // the palette lives in global memory here (LDS holds it only inside the transport)
    const float2 md = P.vox_fmt == 1 ? P.palette[P.vox8[v]] : P.vox_fmt == 2 ? P.palette[P.vox16[v]] : P.vox[v];
    const float2 tb = P.tot[idx * K_MAXMAT + (__float_as_int(md.y) & 0xff)];
    out[v] = make_float2(0.0f, md.x * (tb.x + K_E511 * tb.y));
  }
}

// Activity of the emitting-voxel list into the map's .x.
__global__ void k_dense_activity(const int* __restrict__ vox, const float* __restrict__ act, long long n, float2* __restrict__ out)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[vox[i]].x = act[i];
}

// Line integrals through the phantom along the lines of response of a rebinned sinogram grid
// (the scatter surrogate's phantom inputs). Workgroup (h, w) owns the crystal pairs
// lors[off[h W + w] .. off[h W + w + 1]) of its transaxial bins; item i = d nl + l of the group
// is pair l in plane group d (plane-major: a wave's lanes share a plane, whose lines all lie in
// or all outside the phantom's axial extent), whose line runs in the plane z of the group's central ring sum
// between the two crystal centres (angle 2 pi (k + 1/2) / N - pi at radius R around the
// scanner axis) and is sampled at t0 + (j + 1/2) step < t1 inside the phantom's box [t0, t1]:
// act_out = mean over the pairs of the activity integral (Bq cm), att_out = mean of
// exp(-integral of mu(511 keV)) (map: the (activity, mu) of every voxel). One lane per item;
// the per-plane sums run over the pairs in order (deterministic).
__global__ __launch_bounds__(256) void k_project(const KParams* __restrict__ Pd, const float2* __restrict__ map,
                                                 const int2* __restrict__ lors, const int* __restrict__ off, int D, int H,
                                                 int W, int gz, float step, float* __restrict__ act_out,
                                                 float* __restrict__ att_out)
{
  const KParams& P = *Pd;
  __shared__ float s_act[256], s_att[256];
  const int g = blockIdx.x, h = g / W, w = g - h * W;
  const int l0 = off[g], nl = off[g + 1] - l0, n_items = nl * D;
  const float twopi_n = 2.0f * (float)K_PI / (float)P.ncrystals;
  float acc_a = 0.0f, acc_t = 0.0f;   // plane group threadIdx.x (< D)
  for (int base = 0; base < n_items; base += 256) {
    const int i = base + (int)threadIdx.x;
    float ia = 0.0f, im = 0.0f;
    if (i < n_items) {
      const int d = i / nl, l = i - d * nl;
      const float sc = fminf((float)(d * gz) + 0.5f * (float)(gz - 1), (float)(P.nzs - 1));
      const float z = P.cz + ((0.5f * sc + 0.5f) / (float)P.nrows - 0.5f) * P.fovz;
      const int2 pr = lors[l0 + l];
      const float f1 = twopi_n * ((float)pr.x + 0.5f) - (float)K_PI, f2 = twopi_n * ((float)pr.y + 0.5f) - (float)K_PI;
      const float ax = P.cx + P.radius * cosf(f1), ay = P.cy + P.radius * sinf(f1);
      float dx = P.cx + P.radius * cosf(f2) - ax, dy = P.cy + P.radius * sinf(f2) - ay;
      const float len = sqrtf(dx * dx + dy * dy);
      if (len > 0.0f && z > 0.0f && z < P.bz) {
        dx /= len;
        dy /= len;
        // clip to the box [0, bx] x [0, by]
        float t0 = 0.0f, t1 = len;
        const float o[2] = {ax, ay}, v[2] = {dx, dy}, hi[2] = {P.bx, P.by};
        for (int k = 0; k < 2; k++) {
          if (fabsf(v[k]) < 1e-12f) {
            if (o[k] <= 0.0f || o[k] >= hi[k]) t1 = -1.0f;
          } else {
            const float ta = (0.0f - o[k]) / v[k], tb = (hi[k] - o[k]) / v[k];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
          }
        }
        // four samples per pass, their loads issued together
        for (int j = 0; t0 + ((float)j + 0.5f) * step < t1; j += 4) {
          float2 m[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const float t = t0 + ((float)(j + u) + 0.5f) * step;
            const float p[3] = {ax + t * dx, ay + t * dy, z};
            const int v0 = t < t1 ? locate(P, p) : -1;
            m[u] = v0 >= 0 ? map[v0] : make_float2(0.0f, 0.0f);
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            ia += m[u].x;
            im += m[u].y;
          }
        }
      }
    }
    s_act[threadIdx.x] = ia * step;
    s_att[threadIdx.x] = expf(-im * step);
    __syncthreads();
    if ((int)threadIdx.x < D) {
      // items of plane d in this chunk: i = d nl + l, in increasing l
      const int d = threadIdx.x, lo = max(base, d * nl), hi = min(base + 256, (d + 1) * nl);
      for (int i = lo; i < hi; i++) {
        acc_a += s_act[i - base];
        acc_t += s_att[i - base];
      }
    }
    __syncthreads();
  }
  if ((int)threadIdx.x < D) {
    const size_t oi = ((size_t)threadIdx.x * H + h) * W + w;
    act_out[oi] = nl > 0 ? acc_a / (float)nl : 0.0f;
    att_out[oi] = nl > 0 ? acc_t / (float)nl : 1.0f;
  }
}

// a[0..n) times float32(1 / max(a)) (a >= 0; unchanged when the maximum is 0), in two launches:
// k_max_bits folds the block maxima into *mx (float bits, which order as unsigned integers for
// a >= 0; *mx starts at 0), k_scale_by reads it. The maximum is order-independent, so the
// result is exact and reproducible.
__global__ __launch_bounds__(256) void k_max_bits(const float* __restrict__ a, int n, unsigned* __restrict__ mx)
{
  __shared__ float s[256];
  float m = 0.0f;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = fmaxf(m, a[i]);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(mx, __float_as_uint(s[0]));
}

__global__ __launch_bounds__(256) void k_scale_by(float* __restrict__ a, int n, const unsigned* __restrict__ mx)
{
  const float m = __uint_as_float(*mx);
  if (!(m > 0.0f)) return;
  const float sc = (float)(1.0 / (double)m);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a[i] = a[i] * sc;
}

hipError_t mcpet_launch_project(const KParams* d_params, const int* active_vox, const float* active_act, long long n_active,
                                float2* map, size_t nvox, const int2* lors, const int* off, int D, int H, int W, int gz,
                                float step, float* act_out, float* att_out, hipStream_t stream)
{
  if (D < 1 || D > 256 || H < 1 || W < 1) return hipErrorInvalidValue;
  const size_t nb = (nvox + 255) / 256;
  const unsigned mb = nb < 1 ? 1u : nb > 8192 ? 8192u : (unsigned)nb;
  hipLaunchKernelGGL(k_dense_mu, dim3(mb), dim3(256), 0, stream, d_params, map, (long long)nvox);
  if (n_active > 0)
    hipLaunchKernelGGL(k_dense_activity, dim3((unsigned)((n_active + 255) / 256)), dim3(256), 0, stream, active_vox, active_act,
                       n_active, map);
  hipLaunchKernelGGL(k_project, dim3(H * W), dim3(256), 0, stream, d_params, map, lors, off, D, H, W, gz, step, act_out,
                     att_out);
  return hipGetLastError();
}

hipError_t mcpet_launch_scale_by_max(float* a, int n, unsigned* mx, hipStream_t stream)
{
  hipError_t e = hipMemsetAsync(mx, 0, sizeof(unsigned), stream);
  if (e != hipSuccess) return e;
  const int nb = n < 1 ? 1 : (n + 255) / 256 > 256 ? 256 : (n + 255) / 256;
  hipLaunchKernelGGL(k_max_bits, dim3(nb), dim3(256), 0, stream, a, n, mx);
  hipLaunchKernelGGL(k_scale_by, dim3(nb), dim3(256), 0, stream, a, n, mx);
  return hipGetLastError();
}

hipError_t mcpet_launch_fill_regions(void* const* p, const size_t* n_words, int n, hipStream_t stream)
{
  for (int i0 = 0; i0 < n; i0 += K_FILL_MAX) {
    FillRegions r{};
    const int m = std::min(K_FILL_MAX, n - i0);
    size_t most = 0;
    for (int k = 0; k < m; k++) {
      r.p[k] = (unsigned*)p[i0 + k];
      r.n[k] = n_words[i0 + k];
      most = std::max(most, n_words[i0 + k]);
    }
    if (most == 0) continue;
    const unsigned bx = (unsigned)std::min<size_t>(2048, (most / 4 + 255) / 256 + 1);
    hipLaunchKernelGGL(k_fill_regions, dim3(bx, m), dim3(256), 0, stream, r);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t mcpet_launch_fill_u32(void* p, size_t n_words, hipStream_t stream)
{
  if (n_words == 0) return hipSuccess;
  const size_t per = 256 * 16;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n_words + per - 1) / per)), dim3(256), 0, stream, (unsigned*)p, n_words);
  return hipGetLastError();
}

hipError_t mcpet_launch_selftest_ranecu(int seed, const long long* d_ids, int nstreams, int n, float* d_out, double* d_out_d,
                                        int as_double, const uint32_t* j1, const uint32_t* j2, hipStream_t stream)
{
  int blocks = (nstreams + 255) / 256;
  hipLaunchKernelGGL(k_selftest_ranecu, dim3(blocks), dim3(256), 0, stream, seed, d_ids, nstreams, n, d_out, d_out_d, as_double,
                     j1, j2);
  return hipGetLastError();
}

hipError_t mcpet_launch_selftest_fmath(const double* d_x, const double* d_y, int n, double* d_od, float* d_of, hipStream_t stream)
{
  hipLaunchKernelGGL(k_selftest_fmath, dim3((n + 255) / 256), dim3(256), 0, stream, d_x, d_y, n, d_od, d_of);
  return hipGetLastError();
}

hipError_t mcpet_launch_copy(const uint4* src, uint4* dst, size_t n16, hipStream_t stream)
{
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, stream, src, dst, n16);
  return hipGetLastError();
}
