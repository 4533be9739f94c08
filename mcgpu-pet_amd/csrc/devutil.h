// devutil.h — device helpers shared by the transport and coincidence kernels.
#ifndef MCPET_DEVUTIL_H
#define MCPET_DEVUTIL_H

#include <hip/hip_runtime.h>
#include <climits>

#define DU_PI 3.14159265358979323846

// Accesses through pointers that address global memory, as global (not flat) instructions.
// Pointers read from a parameter block in memory are generic to the compiler, which then emits
// flat loads and atomics; a flat access counts against both the vector-memory and the LDS wait
// counters, so every later wait for an LDS result would also wait for it.
#define DU_GLOBAL __attribute__((address_space(1)))
template <class T>
static __device__ __forceinline__ T gld(const T* p, size_t i)
{
  return ((const DU_GLOBAL T*)p)[i];
}
// HIP's float2/float4 are classes whose members take generic references: load them as
// builtin vectors.
static __device__ __forceinline__ float2 gld(const float2* p, size_t i)
{
  typedef float v2 __attribute__((ext_vector_type(2)));
  const v2 v = ((const DU_GLOBAL v2*)p)[i];
  return make_float2(v.x, v.y);
}
static __device__ __forceinline__ float4 gld(const float4* p, size_t i)
{
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = ((const DU_GLOBAL v4*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
template <class T>
static __device__ __forceinline__ T gatomic_add(T* p, T v)
{
  return __hip_atomic_fetch_add((DU_GLOBAL T*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
static __device__ __forceinline__ T gatomic_min(T* p, T v)
{
  return __hip_atomic_fetch_min((DU_GLOBAL T*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// __float2int_rd with NaN -> INT_MIN and saturation (matches the host checker), without
// branches: floor, clamp into int range (a NaN clamps to -2^31), convert, and select INT_MAX
// for values from 2^31 up (2147483520 is the largest float below 2^31).
static __device__ __forceinline__ int f2i_rd(float x)
{
  const float f = floorf(x);
  const int r = (int)fminf(fmaxf(f, -2147483648.0f), 2147483520.0f);
  return (f >= 2147483648.0f) ? INT_MAX : r;
}
static __device__ __forceinline__ unsigned long long d2ull_rn(double x)
{
  double r = rint(x);
  if (!(r > 0.0)) return 0ull;
  if (r >= 18446744073709551616.0) return ~0ull;
  return (unsigned long long)r;
}
static __device__ __forceinline__ unsigned long long f2ull_rn(float x)
{
  float r = rintf(x);
  if (!(r > 0.0f)) return 0ull;
  if (r >= 18446744073709551616.0f) return ~0ull;
  return (unsigned long long)r;
}

// Michelogram bin of a pair of detector hits (azimuth phi, axial z relative to the detector
// centre), for any parameter block G with the sinogram fields (ncrystals, nrows, fovz, mrd,
// span, nzs, nangles, nrad, nsinos): the reference's binning (K:1749-1788); -1 if rejected.
// ith >= nangles is rejected as well (only possible when NANGLES < NCRYSTALS/2, where the
// reference writes out of bounds).
template <class G>
static __device__ int sino_bin(const G& P, float phi, float zn, float phi1, float zfn)
{
  int ix1 = f2i_rd((float)(((double)phi / (2.0 * DU_PI) + 0.5) * (double)P.ncrystals));
  int ix2 = f2i_rd((float)(((double)phi1 / (2.0 * DU_PI) + 0.5) * (double)P.ncrystals));
  int iz1 = f2i_rd((float)(((double)(zn / P.fovz) + 0.5) * (double)P.nrows));
  int iz2 = f2i_rd((float)(((double)(zfn / P.fovz) + 0.5) * (double)P.nrows));
  if (iz1 < 0 || iz2 < 0 || iz1 >= P.nrows || iz2 >= P.nrows) return -1;
  if (abs(iz2 - iz1) > P.mrd) return -1;
  float delta = (float)abs(iz2 - iz1);
  int incl = f2i_rd((delta - 1 + ((P.span + 1) / 2)) / (float)P.span);
  int iseg = 2 * incl;
  if (iz2 < iz1 && iseg > 0) iseg = iseg - 1;
  int ofseg = iseg * P.nzs;
  for (int kka = 2; kka <= iseg; kka++) {
    ofseg -= (P.span + 1);
    if (kka > 3) ofseg -= f2i_rd((float)((kka - 2) * 0.5)) * 2 * P.span;
  }
  int ofz = 0;
  if (incl > 0) ofz = ((P.span + 1) / 2) + (incl - 1) * P.span;
  int izm = ofseg + (iz1 + iz2) - ofz;
  if (ix2 < ix1) {
    int t = ix2;
    ix2 = ix1;
    ix1 = t;
  }
  int ith = ((ix1 + ix2 + P.nangles + 1) % P.ncrystals) / 2;
  int ir = abs(ix2 - ix1 - P.nangles);
  if ((ix1 < ith) || (ix2 >= (ith + P.nangles))) ir = -ir;
  ir = ir + P.nrad / 2;
  if (izm < 0 || izm >= P.nsinos || ir < 0 || ir >= P.nrad || ith < 0 || ith >= P.nangles) return -1;
  return izm * P.nangles * P.nrad + ith * P.nrad + ir;
}

#endif
