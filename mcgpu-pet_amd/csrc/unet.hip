// unet.hip — learned scatter surrogate: a small 3D U-Net on the rebinned scatter sinogram,
// trained on simulator output (no reference counterpart; the reference only runs Monte Carlo),
// and the device rebinning of sinograms onto its grid.
//
// Every convolution product (forward, input gradient, weight gradient) is an implicit GEMM
//   C[m][n] (+)= sum_k A(m, k) * B(k, n)
// on BF16 MFMA tiles (v_mfma_f32_16x16x32_bf16: 16x16 outputs, K = 32 per instruction, FP32
// accumulation); operands are rounded to BF16, parameters, activations and gradients are kept
// in FP32 (im2col is never materialised). k_conv holds the three products; k_gemm is a plain
// GEMM on the same tiles, used by the self test of the MFMA operand layout.
// MFMA lane layout (16x16x32 BF16): lane l holds A row (l % 16), k = 8 (l / 16) .. +7; B column
// (l % 16), the same k; the accumulator's 4 values are rows 4 (l / 16) .. +3 of column (l % 16).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.h"
#include "mcpet.h"
#include "runtime_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }   // round to nearest even

// ------------------------------------------------------------------ the plain GEMM (self test)

struct Dense {      // row-major matrix, element (r, c) at p[r * ld + c]
  const float* p;
  int rows, cols, ld;
  __device__ __forceinline__ float operator()(int r, int c) const { return (r < rows && c < cols) ? p[(size_t)r * ld + c] : 0.0f; }
};

// Block: 4 waves; block tile 32 (m) x 64 (n); wave w owns columns [16 w, 16 w + 16) and both
// 16-row halves. Grid: (ceil(N/64), ceil(M/32), split). Split > 1 accumulates partial sums
// over K ranges with atomics.
template <class AOp, class BOp>
__global__ __launch_bounds__(256) void k_gemm(AOp A, BOp B, float* __restrict__ C, int M, int N, int K, int ldc,
                                              const float* __restrict__ bias, int k_per_split, int accumulate)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kg = lane >> 4;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 64 + wave * 16;
  const int kb = blockIdx.z * k_per_split;
  const int ke = min(K, kb + k_per_split);
  floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    bf16x8 a0, a1, b;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = k0 + kg * 8 + j;
      const bool in = k < ke;
      a0[j] = to_bf16(in ? A(m0 + r, k) : 0.0f);
      a1[j] = to_bf16(in ? A(m0 + 16 + r, k) : 0.0f);
      b[j] = to_bf16(in ? B(k, n0 + r) : 0.0f);
    }
    acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, acc1, 0, 0, 0);
  }
  const int col = n0 + r;
  if (col >= N) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int row0 = m0 + kg * 4 + i, row1 = row0 + 16;
    if (row0 < M) {
      float v = acc0[i];
      if (accumulate) atomicAdd(&C[(size_t)row0 * ldc + col], v);
      else C[(size_t)row0 * ldc + col] = v + (bias ? bias[row0] : 0.0f);
    }
    if (row1 < M) {
      float v = acc1[i];
      if (accumulate) atomicAdd(&C[(size_t)row1 * ldc + col], v);
      else C[(size_t)row1 * ldc + col] = v + (bias ? bias[row1] : 0.0f);
    }
  }
}

// accumulate: add into C with atomics (C holds earlier partial sums); else store (+ bias).
template <class AOp, class BOp>
hipError_t gemm(AOp A, BOp B, float* C, int M, int N, int K, int ldc, const float* bias, int split, bool accumulate,
                hipStream_t st)
{
  if (M <= 0 || N <= 0) return hipSuccess;
  if (split < 1) split = 1;
  if (split > 1) accumulate = true;
  int kps = (K + split - 1) / split;
  kps = (kps + 31) / 32 * 32;
  split = (K + kps - 1) / kps;
  if (split < 1) split = 1;
  dim3 grid((N + 63) / 64, (M + 31) / 32, split);
  hipLaunchKernelGGL((k_gemm<AOp, BOp>), grid, dim3(256), 0, st, A, B, C, M, N, K, ldc, bias, kps, accumulate ? 1 : 0);
  return hipGetLastError();
}

// ------------------------------------------------------------------ convolution kernels
//
// Each convolution product is one implicit GEMM with a tap-major reduction index
// k = t * Cp + c (t: kernel tap, 27 for 3x3x3 or 1 for 1x1x1; c: channel, padded to
// Cp = roundup(C, 8)), so a lane's 8 consecutive k share one tap: one address and bounds
// computation, then 8 loads at channel stride.
//   FWD   : Y[co][p]        = act(sum_k Wf[co][k] X[c][s p + d_t] + b[co])    m = co, n = output voxel
//   BDATA : dX[ci][p]      (+)= mask(sum_k Wb[ci][k] dY[c][q], s q + d_t = p)  m = ci, n = input voxel
//   BWGT  : G[co][ci T + t] += sum_q dY[co][q] X[ci][s q + d_t]              m = co, n = t Cin + ci, k = q
// Wf, Wb: BF16 copies of the FP32 weights in those layouts (zero padded). FWD and BDATA read
// their B tensor from a channels-last BF16 copy ([voxel][Cp], padding zero), so a lane's 8
// channels are one 16-byte load; their epilogues write such copies of what they produce (for
// the next product) beside the FP32 channel-first tensor (which BWGT, the bias gradients and
// the ReLU masks read). A workgroup is four waves on one 16 MT x 16 output tile that split the
// k-steps between them and reduce through LDS; BWGT also splits its (voxel) reduction over
// workgroups and adds with atomics.

enum { FWD = 0, BDATA = 1, BWGT = 2 };

// Where an epilogue writes a channels-last BF16 copy of its output: element (m, n) at
// p[n * ld + off + m] (p null: no copy).
struct ClOut {
  __bf16* p;
  int ld, off;
};

struct ConvK {
  const __bf16* A;     // FWD, BDATA: BF16 weights [rows][Kp]
  const float* Ad;     // BWGT: dY [M][K]
  const float* X;      // BWGT: input activations [Cin][D][H][W]
  const __bf16* Xl;    // FWD: input activations [D][H][W][Cbp]; BDATA: dY [oD][oH][oW][Cbp] (channels-last BF16)
  ClOut o1, o2;        // FWD, BDATA: channels-last copies of the output
  float* C;            // FWD: Y [M][N]; BDATA: dX [M][N]; BWGT: weight gradient [Cout][Cin][T]
  const float* bias;   // FWD
  const float* mask;   // BDATA: the layer input (gradient zero where it is <= 0: the ReLU before it); may be null
  int relu, beta;      // FWD: ReLU on the output; BDATA: add to C
  int M, N, K, Kp;     // GEMM sizes; Kp: row length of A
  int Cb, Cbp;         // channels of the B tensor (FWD, BWGT: Cin; BDATA: Cout) and Cb rounded up to 8
  int D, H, W, oD, oH, oW, ks, stride, T;
  int k_per_split;     // BWGT: voxels per workgroup
  int vec;             // BWGT: K % 4 == 0 (dY rows are float4-aligned)
};

__device__ __forceinline__ void tap3(int t, int ks, int& dz, int& dy, int& dx)
{
  if (ks == 1) {
    dz = dy = dx = 0;
  } else {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
  }
}

// q with s q + d = p (stride 1 or 2); false if none
__device__ __forceinline__ bool src_of(int p, int d, int s, int n, int& q)
{
  const int v = p - d;
  if (v < 0) return false;
  if (s == 2) {
    if (v & 1) return false;
    q = v >> 1;
  } else {
    q = v;
  }
  return q < n;
}

template <int MODE, int MT>
__global__ __launch_bounds__(256) void k_conv(ConvK a)
{
  __shared__ floatx4 red[4][MT][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kg = lane >> 4;
  const int m0 = blockIdx.y * (16 * MT);
  const int n = blockIdx.x * 16 + r;                 // this lane's B column
  const bool nvalid = n < a.N;
  const size_t vin = (size_t)a.D * a.H * a.W, vout = (size_t)a.oD * a.oH * a.oW;
  floatx4 acc[MT];
#pragma unroll
  for (int i = 0; i < MT; i++) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};

  // per-lane column coordinates
  int cz = 0, cy = 0, cx = 0, bt = 0, bci = 0;
  if (MODE == FWD) {
    cx = n % a.oW;
    cy = (n / a.oW) % a.oH;
    cz = n / (a.oW * a.oH);
  } else if (MODE == BDATA) {
    cx = n % a.W;
    cy = (n / a.W) % a.H;
    cz = n / (a.W * a.H);
  } else {
    bt = n / a.Cb;
    bci = n - bt * a.Cb;
  }
  int kb = 0, ke = a.K;
  if (MODE == BWGT) {
    kb = blockIdx.z * a.k_per_split;
    ke = min(a.K, kb + a.k_per_split);
  }
  if constexpr (MODE == FWD || MODE == BDATA) {
    // The lane's k = k0 + 8 kg advances by 128 per step: its tap t and channel c0 (k = t Cbp +
    // c0) are carried forward (one division before the loop), the tap's offsets come from two
    // multiplications (exact for t < 27: t / 9 = (7282 t) >> 16, t / 3 = (21846 t) >> 16), the
    // voxel index is 32-bit and the weight rows' addresses are fixed.
    const int Cbp = a.Cbp, dt = 128 / Cbp, dc = 128 - dt * Cbp;
    int k = wave * 32 + kg * 8;
    int t = k / Cbp, c0 = k - t * Cbp;
    const __bf16* arow[MT];
#pragma unroll
    for (int i = 0; i < MT; i++) arow[i] = a.A + (size_t)(m0 + 16 * i + r) * a.Kp;
    bf16x8 zero;
#pragma unroll
    for (int e = 0; e < 8; e++) zero[e] = to_bf16(0.0f);
    for (int k0 = wave * 32; k0 < ke; k0 += 128, k += 128) {
      int dz = -1, dy = -1, dx = -1;   // kernel size 1: tap 0, no offset
      if (a.ks == 3) {
        const int q9 = (t * 7282) >> 16, q3 = (t * 21846) >> 16;
        dz += q9;
        dy += q3 - 3 * q9;
        dx += t - 3 * q3;
      } else {
        dz = dy = dx = 0;
      }
      bool in = nvalid && k < ke;
      int vox = 0;
      if (MODE == FWD) {
        const int z = cz * a.stride + dz, y = cy * a.stride + dy, x = cx * a.stride + dx;
        in = in && z >= 0 && y >= 0 && x >= 0 && z < a.D && y < a.H && x < a.W;
        vox = (z * a.H + y) * a.W + x;
      } else {
        int qz = 0, qy = 0, qx = 0;
        in = in && src_of(cz, dz, a.stride, a.oD, qz) && src_of(cy, dy, a.stride, a.oH, qy) && src_of(cx, dx, a.stride, a.oW, qx);
        vox = (qz * a.oH + qy) * a.oW + qx;
      }
      bf16x8 b = zero;
      if (in) b = *(const bf16x8*)(a.Xl + (size_t)vox * Cbp + c0);
#pragma unroll
      for (int i = 0; i < MT; i++) {
        const bf16x8 av = *(const bf16x8*)(arow[i] + k);
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b, acc[i], 0, 0, 0);
      }
      c0 += dc;
      t += dt;
      if (c0 >= Cbp) {
        c0 -= Cbp;
        t++;
      }
    }
  } else
  for (int k0 = kb + wave * 32; k0 < ke; k0 += 128) {
    const int k = k0 + kg * 8;
    bf16x8 b;
    {
      // B: X[ci][s q + d_t] for q = k .. k + 7
      int dz, dy, dx;
      tap3(bt, a.ks, dz, dy, dx);
      int ox = k % a.oW, oy = (k / a.oW) % a.oH, oz = k / (a.oW * a.oH);
      const float* xc = a.X + (size_t)bci * vin;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int z = oz * a.stride + dz, y = oy * a.stride + dy, x = ox * a.stride + dx;
        const bool in = nvalid && k + e < ke && z >= 0 && y >= 0 && x >= 0 && z < a.D && y < a.H && x < a.W;
        b[e] = to_bf16(in ? xc[((size_t)z * a.H + y) * a.W + x] : 0.0f);
        if (++ox == a.oW) {
          ox = 0;
          if (++oy == a.oH) {
            oy = 0;
            ++oz;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < MT; i++) {
        const int m = m0 + 16 * i + r;
        bf16x8 av;
        if (m < a.M && a.vec && k + 8 <= ke) {
          const float4* p4 = (const float4*)(a.Ad + (size_t)m * a.K + k);
          const float4 u = p4[0], w = p4[1];
          av[0] = to_bf16(u.x); av[1] = to_bf16(u.y); av[2] = to_bf16(u.z); av[3] = to_bf16(u.w);
          av[4] = to_bf16(w.x); av[5] = to_bf16(w.y); av[6] = to_bf16(w.z); av[7] = to_bf16(w.w);
        } else {
#pragma unroll
          for (int e = 0; e < 8; e++) av[e] = to_bf16((m < a.M && k + e < ke) ? a.Ad[(size_t)m * a.K + k + e] : 0.0f);
        }
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b, acc[i], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < MT; i++) red[wave][i][lane] = acc[i];
  __syncthreads();
  // epilogue: wave w finishes the row blocks i = w, w + 4, ...
  for (int i = wave; i < MT; i += 4) {
    const floatx4 s = red[0][i][lane] + red[1][i][lane] + red[2][i][lane] + red[3][i][lane];
    const int row0 = m0 + 16 * i + kg * 4;   // this lane's four rows row0 .. row0 + 3
    if (!nvalid) continue;
    if (MODE == FWD || MODE == BDATA) {
      float ov[4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = row0 + v;
        ov[v] = 0.0f;
        if (row >= a.M) continue;
        const size_t idx = (size_t)row * a.N + n;
        float o;
        if (MODE == FWD) {
          o = s[v] + a.bias[row];
          if (a.relu) o = fmaxf(o, 0.0f);
          if (a.C) a.C[idx] = o;   // null: no FP32 copy (inference)
        } else {
          o = s[v] + (a.beta ? a.C[idx] : 0.0f);
          if (a.mask && !(a.mask[idx] > 0.0f)) o = 0.0f;
          a.C[idx] = o;
        }
        ov[v] = o;
      }
      // channels-last copies: the four channels in one 8-byte store when they all exist
      const ClOut* outs[2] = {&a.o1, &a.o2};
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const ClOut& o = *outs[j];
        if (!o.p) continue;
        __bf16* dst = o.p + (size_t)n * o.ld + o.off + row0;
        if (row0 + 3 < a.M && ((o.ld | o.off) & 3) == 0) {
          bf16x4 q;
#pragma unroll
          for (int v = 0; v < 4; v++) q[v] = to_bf16(ov[v]);
          *(bf16x4*)dst = q;
        } else {
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (row0 + v < a.M) dst[v] = to_bf16(ov[v]);
        }
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (row0 + v < a.M) atomicAdd(&a.C[((size_t)(row0 + v) * a.Cb + bci) * a.T + bt], s[v]);
    }
  }
}

// Forward and input-gradient products as k_conv's, with NC 16-voxel column tiles per workgroup:
// each wave's weight fragments (MT per k-step) serve the NC tiles, so the weights stream from
// L2 once per 16 NC voxels; the four waves still split the k-steps and reduce through LDS.
#ifndef MCPET_CONV_COLS
#define MCPET_CONV_COLS 2
#endif
template <int MODE, int MT, int NC>
__global__ __launch_bounds__(256) void k_conv_cols(ConvK a)
{
  static_assert(MODE == FWD || MODE == BDATA, "forward and input-gradient products only");
  __shared__ floatx4 red[4][NC * MT][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kg = lane >> 4;
  const int m0 = blockIdx.y * (16 * MT);
  int nn[NC], cz[NC], cy[NC], cx[NC];
  bool nvalid[NC];
#pragma unroll
  for (int j = 0; j < NC; j++) {
    const int n = (blockIdx.x * NC + j) * 16 + r;   // this lane's column in tile j
    nn[j] = n;
    nvalid[j] = n < a.N;
    const int w = MODE == FWD ? a.oW : a.W, h = MODE == FWD ? a.oH : a.H;
    cx[j] = n % w;
    cy[j] = (n / w) % h;
    cz[j] = n / (w * h);
  }
  floatx4 acc[NC][MT];
#pragma unroll
  for (int j = 0; j < NC; j++)
#pragma unroll
    for (int i = 0; i < MT; i++) acc[j][i] = floatx4{0.f, 0.f, 0.f, 0.f};
  // tap and channel carried across k-steps as in k_conv
  const int Cbp = a.Cbp, dt = 128 / Cbp, dc = 128 - dt * Cbp, ke = a.K;
  int k = wave * 32 + kg * 8;
  int t = k / Cbp, c0 = k - t * Cbp;
  const __bf16* arow[MT];
#pragma unroll
  for (int i = 0; i < MT; i++) arow[i] = a.A + (size_t)(m0 + 16 * i + r) * a.Kp;
  bf16x8 zero;
#pragma unroll
  for (int e = 0; e < 8; e++) zero[e] = to_bf16(0.0f);
  for (int k0 = wave * 32; k0 < ke; k0 += 128, k += 128) {
    int dz = 0, dy = 0, dx = 0;
    if (a.ks == 3) {
      const int q9 = (t * 7282) >> 16, q3 = (t * 21846) >> 16;
      dz = q9 - 1;
      dy = q3 - 3 * q9 - 1;
      dx = t - 3 * q3 - 1;
    }
    bf16x8 av[MT];
#pragma unroll
    for (int i = 0; i < MT; i++) av[i] = *(const bf16x8*)(arow[i] + k);
#pragma unroll
    for (int j = 0; j < NC; j++) {
      bool in = nvalid[j] && k < ke;
      int vox = 0;
      if (MODE == FWD) {
        const int z = cz[j] * a.stride + dz, y = cy[j] * a.stride + dy, x = cx[j] * a.stride + dx;
        in = in && z >= 0 && y >= 0 && x >= 0 && z < a.D && y < a.H && x < a.W;
        vox = (z * a.H + y) * a.W + x;
      } else {
        int qz = 0, qy = 0, qx = 0;
        in = in && src_of(cz[j], dz, a.stride, a.oD, qz) && src_of(cy[j], dy, a.stride, a.oH, qy) &&
             src_of(cx[j], dx, a.stride, a.oW, qx);
        vox = (qz * a.oH + qy) * a.oW + qx;
      }
      bf16x8 b = zero;
      if (in) b = *(const bf16x8*)(a.Xl + (size_t)vox * Cbp + c0);
#pragma unroll
      for (int i = 0; i < MT; i++) acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], b, acc[j][i], 0, 0, 0);
    }
    c0 += dc;
    t += dt;
    if (c0 >= Cbp) {
      c0 -= Cbp;
      t++;
    }
  }
#pragma unroll
  for (int j = 0; j < NC; j++)
#pragma unroll
    for (int i = 0; i < MT; i++) red[wave][j * MT + i][lane] = acc[j][i];
  __syncthreads();
  // epilogue (as k_conv's): wave w finishes the (tile, row block) pairs w, w + 4, ...
  for (int q = wave; q < NC * MT; q += 4) {
    const int j = q / MT, i = q - (q / MT) * MT;
    const floatx4 s = red[0][q][lane] + red[1][q][lane] + red[2][q][lane] + red[3][q][lane];
    const int row0 = m0 + 16 * i + kg * 4;
    const int n = nn[j];
    if (!nvalid[j]) continue;
    float ov[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = row0 + v;
      ov[v] = 0.0f;
      if (row >= a.M) continue;
      const size_t idx = (size_t)row * a.N + n;
      float o;
      if (MODE == FWD) {
        o = s[v] + a.bias[row];
        if (a.relu) o = fmaxf(o, 0.0f);
        if (a.C) a.C[idx] = o;
      } else {
        o = s[v] + (a.beta ? a.C[idx] : 0.0f);
        if (a.mask && !(a.mask[idx] > 0.0f)) o = 0.0f;
        a.C[idx] = o;
      }
      ov[v] = o;
    }
    const ClOut* outs[2] = {&a.o1, &a.o2};
#pragma unroll
    for (int jj = 0; jj < 2; jj++) {
      const ClOut& o = *outs[jj];
      if (!o.p) continue;
      __bf16* dst = o.p + (size_t)n * o.ld + o.off + row0;
      if (row0 + 3 < a.M && ((o.ld | o.off) & 3) == 0) {
        bf16x4 qv;
#pragma unroll
        for (int v = 0; v < 4; v++) qv[v] = to_bf16(ov[v]);
        *(bf16x4*)dst = qv;
      } else {
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (row0 + v < a.M) dst[v] = to_bf16(ov[v]);
      }
    }
  }
}

// Forward and input-gradient products of a stride-1 3x3x3 convolution whose B tensor (the
// input, or the output gradient) has a multiple of 32 (padded) channels, on a tile of 4 FT_R
// rows x 16 voxels of one z-plane per workgroup (wave w: rows y0 + FT_R w .. + FT_R - 1, the 16
// voxels of each as its MFMA columns). The B tile with its one-voxel halo (3 x (4 FT_R + 2) x 18
// voxels) is staged in LDS 32 channels at a time, zero outside the volume, so a k-step is one
// LDS read per row at a compile-time offset per tap (no tap decode, no bounds test, no global
// gather) and MT MFMAs per row; each weight fragment, read as in k_conv, serves the wave's FT_R
// rows (the weights stream from L2 once per FT_R rows). FWD reads voxel p + d_t for tap t, BDATA
// (stride 1: q = p - d_t) voxel p - d_t; the epilogues are k_conv's.
#ifndef MCPET_TILED_ROWS
#define MCPET_TILED_ROWS 2
#endif
#define FT_R MCPET_TILED_ROWS
#define FT_Y (4 * FT_R)
#define FT_X 16
#define FT_HY (FT_Y + 2)
#define FT_HX (FT_X + 2)
#define FT_HALO (3 * FT_HY * FT_HX)

template <int MODE, int MT>
__global__ __launch_bounds__(256) void k_conv_tiled(ConvK a)
{
  __shared__ bf16x8 s_halo[FT_HALO * 4];              // [halo voxel][channel group of 8]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kg = lane >> 4;
  const int tiles_x = (a.W + FT_X - 1) / FT_X, tiles_y = (a.H + FT_Y - 1) / FT_Y;
  int b = blockIdx.x;
  const int x0 = (b % tiles_x) * FT_X;
  b /= tiles_x;
  const int y0 = (b % tiles_y) * FT_Y;
  const int z = b / tiles_y;
  const int m0 = blockIdx.y * (16 * MT);
  floatx4 acc[FT_R][MT];
#pragma unroll
  for (int j = 0; j < FT_R; j++)
#pragma unroll
    for (int i = 0; i < MT; i++) acc[j][i] = floatx4{0.f, 0.f, 0.f, 0.f};
  const __bf16* arow[MT];
#pragma unroll
  for (int i = 0; i < MT; i++) arow[i] = a.A + (size_t)(m0 + 16 * i + r) * a.Kp + kg * 8;
  bf16x8 zero;
#pragma unroll
  for (int e = 0; e < 8; e++) zero[e] = to_bf16(0.0f);
  for (int c0 = 0; c0 < a.Cbp; c0 += 32) {
    for (int i = threadIdx.x; i < FT_HALO * 4; i += 256) {
      const int hv = i >> 2, q = i & 3;
      const int hx = hv % FT_HX, hy = (hv / FT_HX) % FT_HY, hz = hv / (FT_HX * FT_HY);
      const int gz = z + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
      bf16x8 v = zero;
      if (gz >= 0 && gz < a.D && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
        v = *(const bf16x8*)(a.Xl + (((size_t)gz * a.H + gy) * a.W + gx) * a.Cbp + c0 + q * 8);
      s_halo[i] = v;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 27; t++) {
      const int dz = MODE == FWD ? t / 9 : 2 - t / 9;            // halo coordinates of p + d_t or p - d_t
      const int dy = MODE == FWD ? (t / 3) % 3 : 2 - (t / 3) % 3;
      const int dx = MODE == FWD ? t % 3 : 2 - t % 3;
      const int k = t * a.Cbp + c0;
      bf16x8 av[MT];
#pragma unroll
      for (int i = 0; i < MT; i++) av[i] = *(const bf16x8*)(arow[i] + k);
#pragma unroll
      for (int j = 0; j < FT_R; j++) {
        const bf16x8 bv = s_halo[((dz * FT_HY + FT_R * wave + j + dy) * FT_HX + r + dx) * 4 + kg];
#pragma unroll
        for (int i = 0; i < MT; i++) acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv, acc[j][i], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // epilogue (as k_conv's): this lane's columns are voxels (z, y0 + FT_R wave + j, x0 + r)
  const int x = x0 + r;
#pragma unroll
  for (int j = 0; j < FT_R; j++) {
    const int y = y0 + FT_R * wave + j;
    if (x >= a.W || y >= a.H) continue;
    const int n = (z * a.H + y) * a.W + x;
#pragma unroll
    for (int i = 0; i < MT; i++) {
      const int row0 = m0 + 16 * i + kg * 4;
      float ov[4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = row0 + v;
        ov[v] = 0.0f;
        if (row >= a.M) continue;
        const size_t idx = (size_t)row * a.N + n;
        float o;
        if (MODE == FWD) {
          o = acc[j][i][v] + a.bias[row];
          if (a.relu) o = fmaxf(o, 0.0f);
          if (a.C) a.C[idx] = o;
        } else {
          o = acc[j][i][v] + (a.beta ? a.C[idx] : 0.0f);
          if (a.mask && !(a.mask[idx] > 0.0f)) o = 0.0f;
          a.C[idx] = o;
        }
        ov[v] = o;
      }
      const ClOut* outs[2] = {&a.o1, &a.o2};
#pragma unroll
      for (int jj = 0; jj < 2; jj++) {
        const ClOut& o = *outs[jj];
        if (!o.p) continue;
        __bf16* dst = o.p + (size_t)n * o.ld + o.off + row0;
        if (row0 + 3 < a.M && ((o.ld | o.off) & 3) == 0) {
          bf16x4 q;
#pragma unroll
          for (int v = 0; v < 4; v++) q[v] = to_bf16(ov[v]);
          *(bf16x4*)dst = q;
        } else {
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (row0 + v < a.M) dst[v] = to_bf16(ov[v]);
        }
      }
    }
  }
}

inline int mt_for(int M)
{
  const int mt = (M + 15) / 16;
  if (mt <= 4) return mt;
  for (int c = 4; c >= 2; c--)
    if (mt % c == 0) return c;
  return 4;
}
inline int rows_for(int M)
{
  const int t = 16 * mt_for(M);
  return (M + t - 1) / t * t;
}
inline int rup(int x, int m) { return (x + m - 1) / m * m; }

template <int MODE>
hipError_t launch_conv(const ConvK& a, hipStream_t st)
{
  const int mt = mt_for(a.M);
  if constexpr (MODE != BWGT && MCPET_CONV_COLS > 1) {
    constexpr int NC = MCPET_CONV_COLS;
    dim3 g((a.N + 16 * NC - 1) / (16 * NC), (a.M + 16 * mt - 1) / (16 * mt), 1);
    switch (mt) {
      case 1: hipLaunchKernelGGL((k_conv_cols<MODE, 1, NC>), g, dim3(256), 0, st, a); break;
      case 2: hipLaunchKernelGGL((k_conv_cols<MODE, 2, NC>), g, dim3(256), 0, st, a); break;
      case 3: hipLaunchKernelGGL((k_conv_cols<MODE, 3, NC>), g, dim3(256), 0, st, a); break;
      default: hipLaunchKernelGGL((k_conv_cols<MODE, 4, NC>), g, dim3(256), 0, st, a); break;
    }
    return hipGetLastError();
  }
  dim3 grid((a.N + 15) / 16, (a.M + 16 * mt - 1) / (16 * mt), 1);
  ConvK b = a;
  if (MODE == BWGT) {
    const long long xy = (long long)grid.x * grid.y;
    int split = (int)std::max<long long>(1, std::min<long long>(2048 / std::max<long long>(1, xy), (a.K + 255) / 256));
    int kps = (a.K + split - 1) / split;
    kps = rup(kps, 128);
    split = (a.K + kps - 1) / kps;
    grid.z = split;
    b.k_per_split = kps;
    b.vec = (a.K % 4 == 0) ? 1 : 0;
  }
  switch (mt) {
    case 1: hipLaunchKernelGGL((k_conv<MODE, 1>), grid, dim3(256), 0, st, b); break;
    case 2: hipLaunchKernelGGL((k_conv<MODE, 2>), grid, dim3(256), 0, st, b); break;
    case 3: hipLaunchKernelGGL((k_conv<MODE, 3>), grid, dim3(256), 0, st, b); break;
    default: hipLaunchKernelGGL((k_conv<MODE, 4>), grid, dim3(256), 0, st, b); break;
  }
  return hipGetLastError();
}

// Stride-1 3x3x3 FWD / BDATA products with Cbp % 32 == 0 (k_conv_tiled); the row blocks as
// k_conv's, so the A rows read stay inside the padded weight copy. Counted in 4-row tiles.
inline long long tiled_blocks(const ConvK& a)
{
  return (long long)((a.W + FT_X - 1) / FT_X) * ((a.H + 3) / 4) * a.D * ((a.M + 16 * mt_for(a.M) - 1) / (16 * mt_for(a.M)));
}
// The tiled products on volumes with at least 1024 four-row tiles (the full-resolution level);
// smaller volumes (the coarser U-Net levels) keep k_conv's 16-voxel tiles and in-workgroup k split.
#ifndef MCPET_TILED_MIN
#define MCPET_TILED_MIN 1024
#endif
inline bool use_tiled(const ConvK& a)
{
  return a.ks == 3 && a.stride == 1 && a.Cbp % 32 == 0 && tiled_blocks(a) >= MCPET_TILED_MIN;
}

template <int MODE>
hipError_t launch_conv_tiled(const ConvK& a, hipStream_t st)
{
  const int mt = mt_for(a.M);
  const long long tiles = (long long)((a.W + FT_X - 1) / FT_X) * ((a.H + FT_Y - 1) / FT_Y) * a.D;
  if (a.ks != 3 || a.stride != 1 || a.D != a.oD || a.H != a.oH || a.W != a.oW || a.Cbp % 32 != 0 || a.Kp < 27 * a.Cbp ||
      tiles > 0x7fffffffLL)
    return hipErrorInvalidValue;
  dim3 grid((unsigned)tiles, (a.M + 16 * mt - 1) / (16 * mt), 1);
  switch (mt) {
    case 1: hipLaunchKernelGGL((k_conv_tiled<MODE, 1>), grid, dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_conv_tiled<MODE, 2>), grid, dim3(256), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_conv_tiled<MODE, 3>), grid, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((k_conv_tiled<MODE, 4>), grid, dim3(256), 0, st, a); break;
  }
  return hipGetLastError();
}

// Weight gradient of a 3x3x3 convolution of stride S (1 or 2) from channels-last BF16 copies of
// its input X and of its output gradient dY:
//   G[co][ci][t] = sum_q dY[co][q] X[ci][S q + d_t]      (MFMA m = co, n = ci, k = q)
// A workgroup owns one block of 32 output and 32 input channels and a slab of output-voxel tiles
// (4 rows x 16 voxels of one z-plane, 64 voxels = two 32-voxel k-steps); six waves split the
// accumulators by tap plane dz (3) and input-channel half (2), each holding the 3 x 3 (dy, dx)
// taps x 2 output-channel halves = 18 MFMA tiles. Per tile the workgroup stages in LDS, both
// transposed to voxel-contiguous rows: dY of the tile's 64 voxels (32 x 64) and X of the input
// region the tile's taps read (32 channels x 3 planes x 6 rows x 18 voxels for S = 1; 3 planes x
// 9 rows x 33 voxels for S = 2, each row split into its even and odd voxels), so that each 8-voxel
// k-group of a fragment is one aligned 16-byte LDS read: S = 1 voxel x at index x - x0 + 8, the
// dx = -1 / +1 taps that read shifted by one voxel with v_alignbit across the neighbouring words;
// S = 2 the even voxels 2 (x0 + j) at E[j] (dx = 0), the odd voxels 2 (x0 + j) - 1 at O[j] (dx = -1,
// and shifted by one, dx = +1). Each workgroup writes its slab's sums to part[slab][t][co][ci];
// k_bwgt_reduce adds the slabs in a fixed order into the gradient (deterministic, no atomics).
#define BW_Y 4
#define BW_X 16
#define BW_THREADS 384

struct BwK {
  const __bf16* Xl;   // input, channels-last [vox][ldx] (channels >= ldx: zero)
  const __bf16* Gl;   // output gradient, channels-last [ovox][ldg]
  float* part;        // [n_slabs][27][32 cob][32 cib]
  int ldx, ldg, Cin, Cout, D, H, W, oD, oH, oW;
  int tiles_x, tiles_y, n_tiles, per_slab, cob, cib;
};

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// words of v shifted by one BF16 element towards lower indices, the last one from `next`
__device__ __forceinline__ u32x4 shift_down1(u32x4 v, unsigned next)
{
  u32x4 o;
  o[0] = __builtin_amdgcn_alignbit(v[1], v[0], 16);
  o[1] = __builtin_amdgcn_alignbit(v[2], v[1], 16);
  o[2] = __builtin_amdgcn_alignbit(v[3], v[2], 16);
  o[3] = __builtin_amdgcn_alignbit(next, v[3], 16);
  return o;
}

template <int S>
struct BwLds;
template <>
struct BwLds<1> {
  __bf16 x[32][3][6][32];   // 36 KB: [channel][plane][row][x - x0 + 8]
};
template <>
struct BwLds<2> {
  __bf16 e[32][3][9][16];   // even input voxels 2 (x0 + j), j = 0..15
  __bf16 o[32][3][9][24];   // odd input voxels 2 (x0 + j) - 1, j = 0..16
};

template <int S>
__global__ __launch_bounds__(BW_THREADS) void k_bwgt_tiled(BwK a)
{
  constexpr int HY = S == 1 ? 6 : 9, HX = S == 1 ? 18 : 33;   // input region: 3 x HY x HX voxels
  __shared__ __attribute__((aligned(16))) BwLds<S> L;
  __shared__ __attribute__((aligned(16))) __bf16 gs[32][64];   // 4 KB
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, g = lane >> 4;
  const int dz = wave % 3, nj = wave / 3;   // tap plane and input-channel half of this wave
  const int slab = blockIdx.x, co0 = blockIdx.y * 32, ci0 = blockIdx.z * 32;
  floatx4 acc[3][3][2];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int m = 0; m < 2; m++) acc[i][j][m] = floatx4{0.f, 0.f, 0.f, 0.f};
  bf16x8 zero;
#pragma unroll
  for (int e = 0; e < 8; e++) zero[e] = to_bf16(0.0f);
  const int t_end = min(a.n_tiles, (slab + 1) * a.per_slab);
  for (int tile = slab * a.per_slab; tile < t_end; tile++) {
    int b = tile;
    const int x0 = (b % a.tiles_x) * BW_X;
    b /= a.tiles_x;
    const int y0 = (b % a.tiles_y) * BW_Y;
    const int z0 = b / a.tiles_y;
    // input region: voxel (S z0 - 1 + hz, S y0 - 1 + hy, S x0 - 1 + hx), 4 groups of 8 channels
    for (int i = threadIdx.x; i < 3 * HY * HX * 4; i += BW_THREADS) {
      const int q = i & 3, hv = i >> 2;
      const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
      const int gz = S * z0 + hz - 1, gy = S * y0 + hy - 1, gx = S * x0 + hx - 1;
      const int c = ci0 + q * 8;
      bf16x8 v = zero;
      if (gz >= 0 && gz < a.D && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W && c < a.ldx)
        v = *(const bf16x8*)(a.Xl + (((size_t)gz * a.H + gy) * a.W + gx) * a.ldx + c);
      if constexpr (S == 1) {
#pragma unroll
        for (int e = 0; e < 8; e++) L.x[q * 8 + e][hz][hy][hx + 7] = v[e];
      } else if (hx & 1) {
#pragma unroll
        for (int e = 0; e < 8; e++) L.e[q * 8 + e][hz][hy][hx >> 1] = v[e];
      } else {
#pragma unroll
        for (int e = 0; e < 8; e++) L.o[q * 8 + e][hz][hy][hx >> 1] = v[e];
      }
    }
    // dY of the tile's 64 output voxels, transposed into gs[c][q] (q = row * 16 + x)
    for (int i = threadIdx.x; i < 64 * 4; i += BW_THREADS) {
      const int q4 = i & 3, qv = i >> 2;
      const int gy = y0 + (qv >> 4), gx = x0 + (qv & 15);
      const int c = co0 + q4 * 8;
      bf16x8 v = zero;
      if (gy < a.oH && gx < a.oW && c < a.ldg)
        v = *(const bf16x8*)(a.Gl + (((size_t)z0 * a.oH + gy) * a.oW + gx) * a.ldg + c);
#pragma unroll
      for (int e = 0; e < 8; e++) gs[q4 * 8 + e][qv] = v[e];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      const int r = 2 * ks + (g >> 1), xh = (g & 1) * 8;   // this lane's 8 voxels: row r, x = xh .. xh + 7
      bf16x8 av[2];
#pragma unroll
      for (int m = 0; m < 2; m++) av[m] = *(const bf16x8*)&gs[16 * m + lr][r * 16 + xh];
#pragma unroll
      for (int dy = 0; dy < 3; dy++) {
        bf16x8 bx[3];   // the dx = -1, 0, +1 fragments
        if constexpr (S == 1) {
          const __bf16* row = &L.x[16 * nj + lr][dz][r + dy][0];
          const u32x4 bm = *(const u32x4*)(row + xh + 8);          // x = xh .. xh + 7
          const unsigned lo = *(const unsigned*)(row + xh + 6);    // x = xh - 2, xh - 1
          const unsigned hi = *(const unsigned*)(row + xh + 16);   // x = xh + 8, xh + 9
          u32x4 bl;
          bl[0] = __builtin_amdgcn_alignbit(bm[0], lo, 16);
          bl[1] = __builtin_amdgcn_alignbit(bm[1], bm[0], 16);
          bl[2] = __builtin_amdgcn_alignbit(bm[2], bm[1], 16);
          bl[3] = __builtin_amdgcn_alignbit(bm[3], bm[2], 16);
          bx[0] = as_bf16x8(bl);
          bx[1] = as_bf16x8(bm);
          bx[2] = as_bf16x8(shift_down1(bm, hi));
        } else {
          const int yy = 2 * r + dy;
          const __bf16* orow = &L.o[16 * nj + lr][dz][yy][0];
          const u32x4 ov = *(const u32x4*)(orow + xh);             // O[xh .. xh + 7]: x = 2 j - 1
          const unsigned hi = *(const unsigned*)(orow + xh + 8);   // O[xh + 8], O[xh + 9]
          bx[0] = as_bf16x8(ov);
          bx[1] = *(const bf16x8*)&L.e[16 * nj + lr][dz][yy][xh];  // x = 2 j
          bx[2] = as_bf16x8(shift_down1(ov, hi));                  // x = 2 j + 1
        }
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
#pragma unroll
          for (int m = 0; m < 2; m++)
            acc[dy][dx][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bx[dx], acc[dy][dx][m], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // this lane: columns ci = ci0 + 16 nj + lr, rows co = co0 + 16 m + 4 g + v of each tap's tile
  const int CoutR = 32 * a.cob, CinR = 32 * a.cib;
#pragma unroll
  for (int dy = 0; dy < 3; dy++)
#pragma unroll
    for (int dx = 0; dx < 3; dx++) {
      const int t = dz * 9 + dy * 3 + dx;
      float* p = a.part + ((size_t)slab * 27 + t) * CoutR * CinR + (ci0 + 16 * nj + lr);
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int v = 0; v < 4; v++) p[(size_t)(co0 + 16 * m + 4 * g + v) * CinR] = acc[dy][dx][m][v];
    }
}

// G[co][ci][t] += sum over slabs of part[slab][t][co][ci]. A workgroup takes 64 consecutive
// outputs in the partial sums' (t, co, ci) order (each slab's reads contiguous across them) and
// four slab phases: thread (j, ph) adds slabs ph, ph + 4, ... of output j, and the four phase
// sums are added in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_bwgt_reduce(const float* __restrict__ part, float* __restrict__ G, int Cout, int Cin,
                                                     int CoutR, int CinR, int n_slabs)
{
  __shared__ float ps[4][64];
  const long long n = (long long)Cout * Cin * 27;
  const int j = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const long long i = (long long)blockIdx.x * 64 + j;
  float s = 0.0f;
  int ci = 0, co = 0, t = 0;
  if (i < n) {
    ci = (int)(i % Cin);
    co = (int)((i / Cin) % Cout);
    t = (int)(i / ((long long)Cin * Cout));
    const size_t stride = (size_t)27 * CoutR * CinR;
    const float* p = part + ((size_t)t * CoutR + co) * CinR + ci;
#pragma unroll 8
    for (int k = ph; k < n_slabs; k += 4) s += p[(size_t)k * stride];
  }
  ps[ph][j] = s;
  __syncthreads();
  if (ph == 0 && i < n) G[((size_t)co * Cin + ci) * 27 + t] += ((ps[0][j] + ps[1][j]) + ps[2][j]) + ps[3][j];
}

// Workgroups of the tiled weight gradient: about BW_WGS in total, the output-voxel tiles dealt
// to slabs of equal length; the partial sums then fit in BW_WGS x 27 x 1024 floats.
#ifndef BW_WGS
#define BW_WGS 512
#endif
inline void bwgt_shape(int oD, int oH, int oW, int Cin, int Cout, BwK& b, int& n_slabs)
{
  b.tiles_x = (oW + BW_X - 1) / BW_X;
  b.tiles_y = (oH + BW_Y - 1) / BW_Y;
  b.n_tiles = b.tiles_x * b.tiles_y * oD;
  b.cob = (Cout + 31) / 32;
  b.cib = (Cin + 31) / 32;
  n_slabs = std::max(1, std::min(b.n_tiles, BW_WGS / (b.cob * b.cib)));
  b.per_slab = (b.n_tiles + n_slabs - 1) / n_slabs;
  n_slabs = (b.n_tiles + b.per_slab - 1) / b.per_slab;
}
constexpr size_t kBwPartFloats = (size_t)BW_WGS * 27 * 1024;

// true when the tiled weight gradient takes this layer (its slab sums fit the buffer)
inline bool bwgt_fits(int oD, int oH, int oW, int Cin, int Cout)
{
  BwK b{};
  int n_slabs = 0;
  bwgt_shape(oD, oH, oW, Cin, Cout, b, n_slabs);
  return (size_t)n_slabs * 27 * (32 * b.cob) * (32 * b.cib) <= kBwPartFloats;
}

// stride 1 (input and output volumes equal) or 2 (output (D/2, H/2, W/2) rounded down or up)
hipError_t launch_bwgt_tiled(int stride, const __bf16* xl, int ldx, const __bf16* gyl, int ldg, int Cin, int Cout, int D, int H,
                             int W, int oD, int oH, int oW, float* part, float* G, hipStream_t st)
{
  BwK b{};
  int n_slabs = 0;
  bwgt_shape(oD, oH, oW, Cin, Cout, b, n_slabs);
  if ((size_t)n_slabs * 27 * (32 * b.cob) * (32 * b.cib) > kBwPartFloats || ldx % 8 != 0 || ldg % 8 != 0 ||
      (stride != 1 && stride != 2))
    return hipErrorInvalidValue;
  b.Xl = xl;
  b.Gl = gyl;
  b.part = part;
  b.ldx = ldx;
  b.ldg = ldg;
  b.Cin = Cin;
  b.Cout = Cout;
  b.D = D;
  b.H = H;
  b.W = W;
  b.oD = oD;
  b.oH = oH;
  b.oW = oW;
  if (stride == 1) hipLaunchKernelGGL(k_bwgt_tiled<1>, dim3(n_slabs, b.cob, b.cib), dim3(BW_THREADS), 0, st, b);
  else hipLaunchKernelGGL(k_bwgt_tiled<2>, dim3(n_slabs, b.cob, b.cib), dim3(BW_THREADS), 0, st, b);
  const long long n = (long long)Cout * Cin * 27;
  hipLaunchKernelGGL(k_bwgt_reduce, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, (const float*)part, G, Cout, Cin,
                     32 * b.cob, 32 * b.cib, n_slabs);
  return hipGetLastError();
}

// BF16 weight copies: out[m][t * Cp + c] = W[co][ci][t] with (co, ci) = (m, c) (forward) or
// (c, m) (transposed, for the input gradient); zero outside the real channels.
__global__ void k_derive(const float* __restrict__ w, __bf16* __restrict__ out, int rows, int Kp, int Cout, int Cin, int T, int Cp,
                         int transpose)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * Kp) return;
  const int m = (int)(i / Kp), k = (int)(i % Kp);
  const int t = k / Cp, c = k - t * Cp;
  const int co = transpose ? c : m, ci = transpose ? m : c;
  const bool ok = t < T && co < Cout && ci < Cin;
  out[i] = to_bf16(ok ? w[((size_t)co * Cin + ci) * T + t] : 0.0f);
}

// ------------------------------------------------------------------ elementwise kernels

// nearest x2 upsampling of [C][d][h][w] into [C][2d][2h][2w] (and its channels-last copy o)
__global__ void k_upsample(const float* __restrict__ x, float* __restrict__ y, int C, int d, int h, int w, ClOut o)
{
  const size_t n = (size_t)C * 8 * d * h * w;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int W2 = 2 * w, H2 = 2 * h, D2 = 2 * d;
  const size_t vox2 = (size_t)W2 * H2 * D2;
  const int xx = (int)(i % W2), yy = (int)((i / W2) % H2), zz = (int)((i / ((size_t)W2 * H2)) % D2);
  const int c = (int)(i / vox2);
  const float v = x[(((size_t)c * d + zz / 2) * h + yy / 2) * w + xx / 2];
  if (y) y[i] = v;   // null: no FP32 copy (inference)
  if (o.p) o.p[(i % vox2) * o.ld + o.off + c] = to_bf16(v);
}
// k_upsample with each thread on 8 channels of one output voxel: the FP32 stores stay
// coalesced across threads (consecutive voxels), and the channels-last copy of the 8 channels
// is one 16-byte store (C % 8 == 0, o.ld and o.off multiples of 8).
__global__ void k_upsample8(const float* __restrict__ x, float* __restrict__ y, int C, int d, int h, int w, ClOut o)
{
  const int W2 = 2 * w, H2 = 2 * h, D2 = 2 * d;
  const size_t vox2 = (size_t)W2 * H2 * D2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vox2 * (size_t)(C / 8)) return;
  const size_t v = i % vox2;
  const int c0 = (int)(i / vox2) * 8;
  const int xx = (int)(v % W2), yy = (int)((v / W2) % H2), zz = (int)(v / ((size_t)W2 * H2));
  const size_t src = ((size_t)(zz / 2) * h + yy / 2) * w + xx / 2, vin = (size_t)d * h * w;
  bf16x8 q;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const float val = x[(size_t)(c0 + e) * vin + src];
    if (y) y[(size_t)(c0 + e) * vox2 + v] = val;   // null: no FP32 copy (inference)
    q[e] = to_bf16(val);
  }
  if (o.p) *(bf16x8*)(o.p + v * o.ld + o.off + c0) = q;
}
// channel-first FP32 [C][vox] -> channels-last BF16 copy
__global__ void k_to_cl(const float* __restrict__ x, int C, size_t vox, ClOut o)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)C * vox) return;
  o.p[(i % vox) * o.ld + o.off + i / vox] = to_bf16(x[i]);
}
__global__ void k_upsample_bwd(const float* __restrict__ gy, float* __restrict__ gx, int C, int d, int h, int w, ClOut o)
{
  const size_t n = (size_t)C * d * h * w;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int xx = (int)(i % w), yy = (int)((i / w) % h), zz = (int)((i / ((size_t)w * h)) % d);
  const int c = (int)(i / ((size_t)w * h * d));
  const int W2 = 2 * w, H2 = 2 * h, D2 = 2 * d;
  float s = 0.0f;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < 2; e++) s += gy[(((size_t)c * D2 + 2 * zz + a) * H2 + 2 * yy + b) * W2 + 2 * xx + e];
  gx[i] = s;
  if (o.p) o.p[(i % ((size_t)w * h * d)) * o.ld + o.off + c] = to_bf16(s);
}
// per-channel sum (bias gradient): workgroups (channel, chunk of `per` voxels), added atomically
__global__ void k_channel_sum(const float* __restrict__ g, float* __restrict__ out, int vox, int per)
{
  __shared__ float sh[256];
  const float* p = g + (size_t)blockIdx.x * vox;
  const int b = blockIdx.y * per, e = min(vox, b + per);
  float s = 0.0f;
  for (int i = b + threadIdx.x; i < e; i += blockDim.x) s += p[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&out[blockIdx.x], sh[0]);
}
// out = x0 + r (residual output) and, with a target, the loss gradient and squared error
__global__ void k_output(const float* __restrict__ x0, const float* __restrict__ r, float* __restrict__ out,
                         const float* __restrict__ target, float* __restrict__ g, float scale, float* __restrict__ sse, int n,
                         ClOut gl)
{
  __shared__ float sh[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (i < n) {
    const float o = x0[i] + r[i];
    out[i] = o;
    if (target) {
      const float d = o - target[i];
      g[i] = 2.0f * d * scale;
      if (gl.p) gl.p[(size_t)i * gl.ld] = to_bf16(g[i]);
      e2 = d * d;
    }
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && target) atomicAdd(sse, sh[0]);
}
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, size_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
__global__ void k_zero(float* __restrict__ p, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// Total of a float array in double, in a fixed order: 256 workgroups sum contiguous chunks,
// then one workgroup sums their results. Integer-valued inputs (counts) give the exact total.
__global__ void k_sum_partial(const float* __restrict__ x, size_t n, double* __restrict__ partial)
{
  __shared__ double sh[256];
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x, b = blockIdx.x * chunk, e = min(n, b + chunk);
  double s = 0.0;
  for (size_t i = b + threadIdx.x; i < e; i += blockDim.x) s += (double)x[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ void k_sum_final(const double* __restrict__ partial, int m, double* __restrict__ out)
{
  __shared__ double sh[256];
  sh[threadIdx.x] = (int)threadIdx.x < m ? partial[threadIdx.x] : 0.0;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}
// network input: scatter and trues counts times float(voxels / total trues)
__global__ void k_normalize(const float* __restrict__ s, const float* __restrict__ t, float* __restrict__ x0, size_t n,
                            const double* __restrict__ total)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tot = total[0];
  const float scale = tot > 0.0 ? (float)((double)n / tot) : 0.0f;
  x0[i] = s[i] * scale;
  x0[n + i] = t[i] * scale;
}
// scatter per decay: max(output, 0) times float(total trues / (voxels * decays))
__global__ void k_estimate(const float* __restrict__ out, float* __restrict__ est, size_t n, const double* __restrict__ total,
                           const unsigned long long* __restrict__ decay_counter)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tot = total[0], decays = (double)(long long)*decay_counter;
  const float f = (tot > 0.0 && decays > 0.0) ? (float)(tot / ((double)n * decays)) : 0.0f;
  est[i] = fmaxf(out[i], 0.0f) * f;
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

// ------------------------------------------------------------------ the network

// One convolution layer: W [Cout][Cin][T] and b [Cout] at offsets of the flat parameter
// vector; BF16 copies at offsets of the BF16 buffer.
struct Conv {
  int cin, cout, ks, stride;
  size_t w, b;                 // offsets into the parameter vector
  size_t wf, wb;               // offsets of the forward / transposed BF16 copies
  int rows_f, kp_f, rows_b, kp_b;
};

struct Vol {
  int C, D, H, W;
  size_t vox() const { return (size_t)D * H * W; }
  size_t size() const { return (size_t)C * vox(); }
};

}  // namespace

// 3D U-Net: enc1 (2 convs, width c) -> stride-2 conv (2c) -> conv -> stride-2 conv (4c) -> conv
// -> upsample, concat with the enc2 skip -> conv (2c) -> upsample, concat with enc1 -> conv (c)
// -> 1x1x1 conv -> residual added to input channel 0. ReLU after every 3x3x3 convolution.
struct mcpet_surrogate {
  int D, H, W, cin, c;
  Conv L[9];
  size_t n_params = 0, n_bf16 = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  __bf16* wbf = nullptr;
  // activations (post-ReLU) and their gradients; concat buffers hold the skip tensors
  Vol vx, va1, vcat1, vd1, vcat2, vd2, vb3, ve2, ve1, vr;
  float *x0 = nullptr, *a1 = nullptr, *cat1 = nullptr, *d1 = nullptr, *cat2 = nullptr, *d2 = nullptr, *b3 = nullptr,
        *e2 = nullptr, *e1 = nullptr, *r = nullptr, *out = nullptr, *tgt = nullptr;
  float *ga1 = nullptr, *gcat1 = nullptr, *gd1 = nullptr, *gcat2 = nullptr, *gd2 = nullptr, *gb3 = nullptr, *ge2 = nullptr,
        *ge1 = nullptr, *gr = nullptr, *sse = nullptr;
  // channels-last BF16 copies read as B operands (activations, then gradients)
  __bf16 *x0l = nullptr, *a1l = nullptr, *b1l = nullptr, *cat1l = nullptr, *d1l = nullptr, *b2l = nullptr, *cat2l = nullptr,
         *d2l = nullptr, *e1l = nullptr;
  __bf16 *grl = nullptr, *ge1l = nullptr, *ge2l = nullptr, *gb3l = nullptr, *gd2l = nullptr, *gb2l = nullptr, *gd1l = nullptr,
         *gb1l = nullptr;
  float *raw_s = nullptr, *raw_t = nullptr, *est = nullptr;   // the fused estimate's buffers
  double* sums = nullptr;                                       // 256 partial totals + the total
  float* bw_part = nullptr;                                     // slab sums of the tiled weight gradient
  long long t_adam = 0;
  int device = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev_f0 = nullptr, ev_f1 = nullptr;   // around the last forward of mcpet_surrogate_forward
  float* h_stage = nullptr;   // pinned host staging of the fused estimate's two outputs (2 x voxels)
  unsigned long long* h_decays = nullptr;   // pinned: the estimated run's decays
  bool fwd_timed = false;
  std::vector<void*> allocs;
  std::string err;
  double flops_fwd = 0.0;
};

namespace {

mcpet_status serr(mcpet_surrogate* s, mcpet_status st, const std::string& m)
{
  if (s) s->err = m;
  mcpet_host_set_error(m.c_str());
  return st;
}

#define SHIP(s, call)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) return serr(s, MCPET_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc_t(mcpet_surrogate* s, T** p, size_t n)
{
  hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(1, n));
  if (e == hipSuccess) {
    s->allocs.push_back(*p);
    e = hipMemset(*p, 0, sizeof(T) * std::max<size_t>(1, n));
  }
  return e;
}

// refresh the BF16 weight copies from the FP32 parameters
hipError_t derive(mcpet_surrogate* s)
{
  for (int i = 0; i < 9; i++) {
    const Conv& l = s->L[i];
    const int T = l.ks * l.ks * l.ks;
    size_t n = (size_t)l.rows_f * l.kp_f;
    hipLaunchKernelGGL(k_derive, dim3(blocks_for(n)), dim3(256), 0, s->st, s->p + l.w, s->wbf + l.wf, l.rows_f, l.kp_f, l.cout,
                       l.cin, T, rup(l.cin, 8), 0);
    n = (size_t)l.rows_b * l.kp_b;
    hipLaunchKernelGGL(k_derive, dim3(blocks_for(n)), dim3(256), 0, s->st, s->p + l.w, s->wbf + l.wb, l.rows_b, l.kp_b, l.cout,
                       l.cin, T, rup(l.cout, 8), 1);
  }
  return hipGetLastError();
}

ConvK conv_args(const Conv& l, const Vol& vin, const Vol& vout)
{
  ConvK a{};
  a.D = vin.D;
  a.H = vin.H;
  a.W = vin.W;
  a.oD = vout.D;
  a.oH = vout.H;
  a.oW = vout.W;
  a.ks = l.ks;
  a.stride = l.stride;
  a.T = l.ks * l.ks * l.ks;
  return a;
}

const ClOut kNoCl{nullptr, 0, 0};

// y = act(conv(x) + b); x read from its channels-last copy xl; o1, o2: copies of y to write
hipError_t conv_fwd(mcpet_surrogate* s, const Conv& l, const Vol& vin, const __bf16* xl, const Vol& vout, float* y, bool relu,
                    ClOut o1, ClOut o2 = kNoCl)
{
  ConvK a = conv_args(l, vin, vout);
  a.A = s->wbf + l.wf;
  a.Xl = xl;
  a.C = y;
  a.o1 = o1;
  a.o2 = o2;
  a.bias = s->p + l.b;
  a.relu = relu ? 1 : 0;
  a.M = l.cout;
  a.N = (int)vout.vox();
  a.Cb = l.cin;
  a.Cbp = rup(l.cin, 8);
  a.K = a.T * a.Cbp;
  a.Kp = l.kp_f;
  s->flops_fwd += 2.0 * l.cout * (double)l.cin * a.T * a.N;
  if (use_tiled(a)) return launch_conv_tiled<FWD>(a, s->st);
  return launch_conv<FWD>(a, s->st);
}
// gx (+)= dL/dx of the layer input, zeroed where mask <= 0; gy read from its channels-last copy
hipError_t conv_bwd_data(mcpet_surrogate* s, const Conv& l, const Vol& vin, const Vol& vout, const __bf16* gyl, float* gx,
                         const float* mask, bool add, ClOut o1)
{
  ConvK a = conv_args(l, vin, vout);
  a.A = s->wbf + l.wb;
  a.Xl = gyl;
  a.C = gx;
  a.o1 = o1;
  a.o2 = kNoCl;
  a.mask = mask;
  a.beta = add ? 1 : 0;
  a.M = l.cin;
  a.N = (int)vin.vox();
  a.Cb = l.cout;
  a.Cbp = rup(l.cout, 8);
  a.K = a.T * a.Cbp;
  a.Kp = l.kp_b;
  if (use_tiled(a)) return launch_conv_tiled<BDATA>(a, s->st);
  return launch_conv<BDATA>(a, s->st);
}
// weight and bias gradients accumulated from gy (dL/d pre-activation) and the layer input x;
// xl, gyl: their channels-last BF16 copies (ClOut ld = row length) or null. 3x3x3 layers with
// both copies take the LDS-tiled product, the others k_conv's.
hipError_t conv_bwd_weights(mcpet_surrogate* s, const Conv& l, const Vol& vin, const float* x, const Vol& vout, const float* gy,
                            ClOut xl = ClOut{nullptr, 0, 0}, ClOut gyl = ClOut{nullptr, 0, 0})
{
  const int vox = (int)vout.vox(), per = 4096;
  if (l.ks == 3 && (l.stride == 1 || l.stride == 2) && xl.p && gyl.p && xl.off == 0 && gyl.off == 0 && xl.ld % 8 == 0 &&
      gyl.ld % 8 == 0 && bwgt_fits(vout.D, vout.H, vout.W, l.cin, l.cout)) {
    hipError_t e = launch_bwgt_tiled(l.stride, xl.p, xl.ld, gyl.p, gyl.ld, l.cin, l.cout, vin.D, vin.H, vin.W, vout.D, vout.H,
                                     vout.W, s->bw_part, s->g + l.w, s->st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_channel_sum, dim3(l.cout, (vox + per - 1) / per), dim3(256), 0, s->st, gy, s->g + l.b, vox, per);
    return hipGetLastError();
  }
  ConvK a = conv_args(l, vin, vout);
  a.Ad = gy;
  a.X = x;
  a.C = s->g + l.w;
  a.M = l.cout;
  a.N = a.T * l.cin;
  a.K = (int)vout.vox();
  a.Cb = l.cin;
  a.Cbp = l.cin;
  hipError_t e = launch_conv<BWGT>(a, s->st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_channel_sum, dim3(l.cout, (vox + per - 1) / per), dim3(256), 0, s->st, gy, s->g + l.b, vox, per);
  return hipGetLastError();
}
hipError_t upsample(mcpet_surrogate* s, const Vol& v, const float* x, float* y, ClOut o)
{
  if (v.C % 8 == 0 && (!o.p || (o.ld % 8 == 0 && o.off % 8 == 0)))
    hipLaunchKernelGGL(k_upsample8, dim3(blocks_for(v.size())), dim3(256), 0, s->st, x, y, v.C, v.D, v.H, v.W, o);
  else
    hipLaunchKernelGGL(k_upsample, dim3(blocks_for(v.size() * 8)), dim3(256), 0, s->st, x, y, v.C, v.D, v.H, v.W, o);
  return hipGetLastError();
}
hipError_t upsample_bwd(mcpet_surrogate* s, const Vol& v, const float* gy, float* gx, ClOut o)
{
  hipLaunchKernelGGL(k_upsample_bwd, dim3(blocks_for(v.size())), dim3(256), 0, s->st, gy, gx, v.C, v.D, v.H, v.W, o);
  return hipGetLastError();
}
ClOut cl(__bf16* p, int channels, int off = 0) { return ClOut{p, rup(channels, 8), off}; }

// Forward of one sample already in s->x0 (r = the residual). keep: write the FP32 copy of every
// activation (training reads them); without it only the tensors the forward itself reads in
// FP32 (the upsampling inputs b3 and e2, the residual r) are written, the layers passing their
// outputs to the next one through the channels-last BF16 copies alone.
hipError_t forward(mcpet_surrogate* s, bool keep)
{
  hipError_t e = hipSuccess;
  const int c = s->c;
  const Vol vb1{c, s->D, s->H, s->W}, vb2{2 * c, s->D / 2, s->H / 2, s->W / 2};
  float* b1 = s->cat1 + (size_t)2 * c * vb1.vox();   // enc1 output lives in the concat buffer
  float* b2 = s->cat2 + (size_t)4 * c * vb2.vox();
  s->flops_fwd = 0.0;
  hipLaunchKernelGGL(k_to_cl, dim3(blocks_for(s->vx.size())), dim3(256), 0, s->st, s->x0, s->cin, s->vx.vox(), cl(s->x0l, s->cin));
  e = hipGetLastError();
#define STEP(x) \
  if (e == hipSuccess) e = (x)
  auto k = [keep](float* p) { return keep ? p : nullptr; };
  STEP(conv_fwd(s, s->L[0], s->vx, s->x0l, s->va1, k(s->a1), true, cl(s->a1l, c)));
  STEP(conv_fwd(s, s->L[1], s->va1, s->a1l, vb1, k(b1), true, cl(s->b1l, c), cl(s->cat1l, 3 * c, 2 * c)));
  STEP(conv_fwd(s, s->L[2], vb1, s->b1l, s->vd1, k(s->d1), true, cl(s->d1l, 2 * c)));
  STEP(conv_fwd(s, s->L[3], s->vd1, s->d1l, vb2, k(b2), true, cl(s->b2l, 2 * c), cl(s->cat2l, 6 * c, 4 * c)));
  STEP(conv_fwd(s, s->L[4], vb2, s->b2l, s->vd2, k(s->d2), true, cl(s->d2l, 4 * c)));
  STEP(conv_fwd(s, s->L[5], s->vd2, s->d2l, s->vb3, s->b3, true, kNoCl));
  STEP(upsample(s, s->vb3, s->b3, k(s->cat2), cl(s->cat2l, 6 * c)));    // first 4c channels of cat2
  STEP(conv_fwd(s, s->L[6], s->vcat2, s->cat2l, s->ve2, s->e2, true, kNoCl));
  STEP(upsample(s, s->ve2, s->e2, k(s->cat1), cl(s->cat1l, 3 * c)));    // first 2c channels of cat1
  STEP(conv_fwd(s, s->L[7], s->vcat1, s->cat1l, s->ve1, k(s->e1), true, cl(s->e1l, c)));
  STEP(conv_fwd(s, s->L[8], s->ve1, s->e1l, s->vr, s->r, false, kNoCl));
#undef STEP
  return e;
}

// Backward of one sample from gr (= dL/dr, with its channels-last copy); accumulates the
// parameter gradients. The ReLU derivative is applied where each gradient is produced (mask =
// the post-ReLU activation; the upsampled and concatenated copies carry the same sign as their
// sources). A gradient's channels-last copy is written by the product that completes it.
hipError_t backward(mcpet_surrogate* s)
{
  hipError_t e = hipSuccess;
  const int c = s->c;
  const Vol vb1{c, s->D, s->H, s->W}, vb2{2 * c, s->D / 2, s->H / 2, s->W / 2};
  const float* b1 = s->cat1 + (size_t)2 * c * vb1.vox();
  const float* b2 = s->cat2 + (size_t)4 * c * vb2.vox();
  float* gb1 = s->gcat1 + (size_t)2 * c * vb1.vox();
  float* gb2 = s->gcat2 + (size_t)4 * c * vb2.vox();
#define STEP(x) \
  if (e == hipSuccess) e = (x)
  STEP(conv_bwd_weights(s, s->L[8], s->ve1, s->e1, s->vr, s->gr));
  STEP(conv_bwd_data(s, s->L[8], s->ve1, s->vr, s->grl, s->ge1, s->e1, false, cl(s->ge1l, c)));
  STEP(conv_bwd_weights(s, s->L[7], s->vcat1, s->cat1, s->ve1, s->ge1, cl(s->cat1l, 3 * c), cl(s->ge1l, c)));
  STEP(conv_bwd_data(s, s->L[7], s->vcat1, s->ve1, s->ge1l, s->gcat1, s->cat1, false, kNoCl));   // g(up1), g(b1 skip)
  STEP(upsample_bwd(s, s->ve2, s->gcat1, s->ge2, cl(s->ge2l, 2 * c)));
  STEP(conv_bwd_weights(s, s->L[6], s->vcat2, s->cat2, s->ve2, s->ge2, cl(s->cat2l, 6 * c), cl(s->ge2l, 2 * c)));
  STEP(conv_bwd_data(s, s->L[6], s->vcat2, s->ve2, s->ge2l, s->gcat2, s->cat2, false, kNoCl));   // g(up2), g(b2 skip)
  STEP(upsample_bwd(s, s->vb3, s->gcat2, s->gb3, cl(s->gb3l, 4 * c)));
  STEP(conv_bwd_weights(s, s->L[5], s->vd2, s->d2, s->vb3, s->gb3, cl(s->d2l, 4 * c), cl(s->gb3l, 4 * c)));
  STEP(conv_bwd_data(s, s->L[5], s->vd2, s->vb3, s->gb3l, s->gd2, s->d2, false, cl(s->gd2l, 4 * c)));
  STEP(conv_bwd_weights(s, s->L[4], vb2, b2, s->vd2, s->gd2, cl(s->b2l, 2 * c), cl(s->gd2l, 4 * c)));
  STEP(conv_bwd_data(s, s->L[4], vb2, s->vd2, s->gd2l, gb2, b2, true, cl(s->gb2l, 2 * c)));   // + the skip's gradient
  STEP(conv_bwd_weights(s, s->L[3], s->vd1, s->d1, vb2, gb2, cl(s->d1l, 2 * c), cl(s->gb2l, 2 * c)));
  STEP(conv_bwd_data(s, s->L[3], s->vd1, vb2, s->gb2l, s->gd1, s->d1, false, cl(s->gd1l, 2 * c)));
  STEP(conv_bwd_weights(s, s->L[2], vb1, b1, s->vd1, s->gd1, cl(s->b1l, c), cl(s->gd1l, 2 * c)));
  STEP(conv_bwd_data(s, s->L[2], vb1, s->vd1, s->gd1l, gb1, b1, true, cl(s->gb1l, c)));       // + the skip's gradient
  STEP(conv_bwd_weights(s, s->L[1], s->va1, s->a1, vb1, gb1, cl(s->a1l, c), cl(s->gb1l, c)));
  STEP(conv_bwd_data(s, s->L[1], s->va1, vb1, s->gb1l, s->ga1, s->a1, false, kNoCl));
  STEP(conv_bwd_weights(s, s->L[0], s->vx, s->x0, s->va1, s->ga1));
#undef STEP
  return e;
}

}  // namespace

extern "C" mcpet_status mcpet_surrogate_create(int32_t D, int32_t H, int32_t W, int32_t in_channels, int32_t width, uint64_t seed,
                                               mcpet_surrogate** out)
{
  if (!out) return serr(nullptr, MCPET_E_ARG, "mcpet_surrogate_create: null output pointer");
  *out = nullptr;
  if (D < 4 || H < 4 || W < 4 || D % 4 || H % 4 || W % 4 || in_channels < 1 || width < 1 || in_channels > 64 || width > 256 ||
      (long long)D * H * W > (1LL << 28))
    return serr(nullptr, MCPET_E_ARG, "mcpet_surrogate_create: sizes must be multiples of 4 (at most 2^28 voxels), "
                                      "1 <= channels <= 64, 1 <= width <= 256");
  mcpet_surrogate* s = new mcpet_surrogate();
  s->D = D;
  s->H = H;
  s->W = W;
  s->cin = in_channels;
  s->c = width;
  const int c = width;
  const int spec[9][4] = {{in_channels, c, 3, 1}, {c, c, 3, 1},         {c, 2 * c, 3, 2},     {2 * c, 2 * c, 3, 1}, {2 * c, 4 * c, 3, 2},
                          {4 * c, 4 * c, 3, 1},    {6 * c, 2 * c, 3, 1}, {3 * c, c, 3, 1},     {c, 1, 1, 1}};
  size_t off = 0, offb = 0;
  for (int i = 0; i < 9; i++) {
    Conv& l = s->L[i];
    l.cin = spec[i][0];
    l.cout = spec[i][1];
    l.ks = spec[i][2];
    l.stride = spec[i][3];
    const int T = l.ks * l.ks * l.ks;
    l.w = off;
    off += (size_t)l.cout * l.cin * T;
    l.b = off;
    off += (size_t)l.cout;
    l.rows_f = rows_for(l.cout);
    l.kp_f = rup(T * rup(l.cin, 8), 32);
    l.rows_b = rows_for(l.cin);
    l.kp_b = rup(T * rup(l.cout, 8), 32);
    l.wf = offb;
    offb += (size_t)l.rows_f * l.kp_f;
    l.wb = offb;
    offb += (size_t)l.rows_b * l.kp_b;
  }
  s->n_params = off;
  s->n_bf16 = offb;
  s->vx = {in_channels, D, H, W};
  s->va1 = {c, D, H, W};
  s->vcat1 = {3 * c, D, H, W};
  s->vd1 = {2 * c, D / 2, H / 2, W / 2};
  s->vcat2 = {6 * c, D / 2, H / 2, W / 2};
  s->vd2 = {4 * c, D / 4, H / 4, W / 4};
  s->vb3 = {4 * c, D / 4, H / 4, W / 4};
  s->ve2 = {2 * c, D / 2, H / 2, W / 2};
  s->ve1 = {c, D, H, W};
  s->vr = {1, D, H, W};
  hipError_t e = hipGetDevice(&s->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&s->ev_f0);
  if (e == hipSuccess) e = hipEventCreate(&s->ev_f1);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_stage, 2 * sizeof(float) * (size_t)D * H * W, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_decays, sizeof(unsigned long long), hipHostMallocDefault);
  struct A {
    float** p;
    size_t n;
  } al[] = {{&s->p, off},          {&s->g, off},          {&s->m, off},           {&s->v, off},
            {&s->x0, s->vx.size()}, {&s->a1, s->va1.size()}, {&s->cat1, s->vcat1.size()}, {&s->d1, s->vd1.size()},
            {&s->cat2, s->vcat2.size()}, {&s->d2, s->vd2.size()}, {&s->b3, s->vb3.size()}, {&s->e2, s->ve2.size()},
            {&s->e1, s->ve1.size()}, {&s->r, s->vr.size()}, {&s->out, s->vr.size()}, {&s->tgt, s->vr.size()},
            {&s->ga1, s->va1.size()}, {&s->gcat1, s->vcat1.size()}, {&s->gd1, s->vd1.size()}, {&s->gcat2, s->vcat2.size()},
            {&s->gd2, s->vd2.size()}, {&s->gb3, s->vb3.size()}, {&s->ge2, s->ve2.size()}, {&s->ge1, s->ve1.size()},
            {&s->gr, s->vr.size()}, {&s->sse, 1}, {&s->raw_s, s->vr.size()}, {&s->raw_t, s->vr.size()},
            {&s->est, s->vr.size()}};
  for (auto& a : al)
    if (e == hipSuccess) e = dalloc_t(s, a.p, a.n);
  if (e == hipSuccess) e = dalloc_t(s, &s->wbf, s->n_bf16);
  if (e == hipSuccess) e = dalloc_t(s, &s->sums, 257);
  if (e == hipSuccess) e = dalloc_t(s, &s->bw_part, kBwPartFloats);
  {
    const size_t v0 = s->vx.vox(), v1 = s->vd1.vox(), v2 = s->vd2.vox();
    struct B {
      __bf16** p;
      size_t n;
    } bl[] = {{&s->x0l, v0 * rup(in_channels, 8)}, {&s->a1l, v0 * rup(c, 8)},      {&s->b1l, v0 * rup(c, 8)},
              {&s->cat1l, v0 * rup(3 * c, 8)},      {&s->d1l, v1 * rup(2 * c, 8)}, {&s->b2l, v1 * rup(2 * c, 8)},
              {&s->cat2l, v1 * rup(6 * c, 8)},      {&s->d2l, v2 * rup(4 * c, 8)}, {&s->e1l, v0 * rup(c, 8)},
              {&s->grl, v0 * 8},                    {&s->ge1l, v0 * rup(c, 8)},    {&s->ge2l, v1 * rup(2 * c, 8)},
              {&s->gb3l, v2 * rup(4 * c, 8)},       {&s->gd2l, v2 * rup(4 * c, 8)}, {&s->gb2l, v1 * rup(2 * c, 8)},
              {&s->gd1l, v1 * rup(2 * c, 8)},       {&s->gb1l, v0 * rup(c, 8)}};
    for (auto& x : bl)
      if (e == hipSuccess) e = dalloc_t(s, x.p, x.n);
  }
  if (e != hipSuccess) {
    std::string m = std::string("mcpet_surrogate_create: ") + hipGetErrorString(e);
    mcpet_surrogate_destroy(s);
    return serr(nullptr, MCPET_E_NOMEM, m);
  }
  // He-normal weights from a seeded generator (splitmix64 + Box-Muller), zero biases; the 1x1x1
  // output layer is zero, so the untrained network returns its input channel 0
  std::vector<float> hp(off, 0.0f);
  uint64_t z = seed ? seed : 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)((x >> 11) + 1) * 0x1.0p-53;
  };
  for (int i = 0; i < 9; i++) {
    const Conv& l = s->L[i];
    const int T = l.ks * l.ks * l.ks;
    const double sd = i == 8 ? 0.0 : std::sqrt(2.0 / (l.cin * T));   // the output layer starts at zero: identity
    for (size_t k = 0; k < (size_t)l.cout * l.cin * T; k++)
      hp[l.w + k] = (float)(sd * std::sqrt(-2.0 * std::log(next())) * std::cos(6.283185307179586 * next()));
  }
  e = hipMemcpy(s->p, hp.data(), sizeof(float) * off, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = derive(s);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  if (e != hipSuccess) {
    mcpet_surrogate_destroy(s);
    return serr(nullptr, MCPET_E_HIP, "mcpet_surrogate_create: upload");
  }
  *out = s;
  return MCPET_OK;
}

extern "C" int64_t mcpet_surrogate_param_count(const mcpet_surrogate* s) { return s ? (int64_t)s->n_params : 0; }

extern "C" mcpet_status mcpet_surrogate_get_params(mcpet_surrogate* s, float* out)
{
  if (!s || !out) return serr(s, MCPET_E_ARG, "mcpet_surrogate_get_params: null argument");
  SHIP(s, hipStreamSynchronize(s->st));
  SHIP(s, hipMemcpy(out, s->p, sizeof(float) * s->n_params, hipMemcpyDeviceToHost));
  return MCPET_OK;
}
extern "C" mcpet_status mcpet_surrogate_set_params(mcpet_surrogate* s, const float* in)
{
  if (!s || !in) return serr(s, MCPET_E_ARG, "mcpet_surrogate_set_params: null argument");
  SHIP(s, hipStreamSynchronize(s->st));
  SHIP(s, hipMemcpy(s->p, in, sizeof(float) * s->n_params, hipMemcpyHostToDevice));
  SHIP(s, derive(s));
  SHIP(s, hipStreamSynchronize(s->st));
  return MCPET_OK;
}

// x: batch x cin x D x H x W, y (out): batch x D x H x W; returns the forward GFLOP per sample.
extern "C" mcpet_status mcpet_surrogate_forward(mcpet_surrogate* s, const float* x, int32_t batch, float* y, double* gflop)
{
  if (!s || !x || !y || batch < 1) return serr(s, MCPET_E_ARG, "mcpet_surrogate_forward: bad arguments");
  const size_t nx = s->vx.size(), nv = s->vr.size();
  for (int b = 0; b < batch; b++) {
    SHIP(s, hipMemcpyAsync(s->x0, x + b * nx, sizeof(float) * nx, hipMemcpyHostToDevice, s->st));
    const bool last = b == batch - 1;
    if (last) SHIP(s, hipEventRecord(s->ev_f0, s->st));
    SHIP(s, forward(s, false));
    hipLaunchKernelGGL(k_output, dim3(blocks_for(nv)), dim3(256), 0, s->st, s->x0, s->r, s->out, (const float*)nullptr,
                       (float*)nullptr, 0.0f, (float*)nullptr, (int)nv, kNoCl);
    SHIP(s, hipGetLastError());
    if (last) SHIP(s, hipEventRecord(s->ev_f1, s->st));
    SHIP(s, hipMemcpyAsync(y + b * nv, s->out, sizeof(float) * nv, hipMemcpyDeviceToHost, s->st));
  }
  SHIP(s, hipStreamSynchronize(s->st));
  s->fwd_timed = true;
  if (gflop) *gflop = s->flops_fwd * 1e-9;
  return MCPET_OK;
}

// Device time of the network in the last mcpet_surrogate_forward (its last sample: every kernel
// from the input conversion to the output, HIP events on the network's stream).
extern "C" mcpet_status mcpet_surrogate_forward_ms(mcpet_surrogate* s, double* ms)
{
  if (!s || !ms) return serr(s, MCPET_E_ARG, "mcpet_surrogate_forward_ms: null argument");
  if (!s->fwd_timed) return serr(s, MCPET_E_STATE, "mcpet_surrogate_forward_ms: no forward yet");
  float t = 0.0f;
  SHIP(s, hipEventElapsedTime(&t, s->ev_f0, s->ev_f1));
  *ms = t;
  return MCPET_OK;
}

// One estimate from the last run of ctx, all on the device: rebinning of its scatter and
// trues onto the network's grid, normalisation (counts times float(voxels / total trues), the
// total exact in double), the network, and scatter per decay = max(output, 0) times
// float(total trues / (voxels * decays)).
void mcpet_surrogate_dims(const mcpet_surrogate* s, int* D, int* H, int* W, int* cin)
{
  *D = s->D;
  *H = s->H;
  *W = s->W;
  *cin = s->cin;
}

extern "C" mcpet_status mcpet_surrogate_estimate(mcpet_surrogate* s, mcpet_ctx* ctx, float* estimate, float* scatter_counts,
                                                 int64_t* decays)
{
  if (!s || !ctx || !estimate) return serr(s, MCPET_E_ARG, "mcpet_surrogate_estimate: null argument");
  if (s->cin != 2 && s->cin != 4)
    return serr(s, MCPET_E_ARG, "mcpet_surrogate_estimate: the network must take 2 or 4 input channels");
  if (mcpet_ctx_device(ctx) != s->device) return serr(s, MCPET_E_ARG, "mcpet_surrogate_estimate: context on another device");
  SHIP(s, hipSetDevice(s->device));
  const size_t nv = s->vr.size();
  // everything below is enqueued behind the context's run on the device: the host waits once,
  // for the outputs
  mcpet_status st = mcpet_ctx_rebin2_device(ctx, 1, 0, s->D, s->H, s->W, s->raw_s, s->raw_t, s->st);
  const unsigned long long* d_decays = mcpet_ctx_decay_counter(ctx);
  const float* maps = nullptr;
  if (st == MCPET_OK && s->cin == 4) st = mcpet_ctx_project_device(ctx, s->D, s->H, s->W, &maps);
  if (st != MCPET_OK) return serr(s, st, std::string("mcpet_surrogate_estimate: ") + mcpet_last_error(ctx));
  if (maps)   // the scaled activity integrals and the attenuation factors
    SHIP(s, hipMemcpyAsync(s->x0 + 2 * nv, maps + nv, 2 * sizeof(float) * nv, hipMemcpyDeviceToDevice, s->st));
  hipLaunchKernelGGL(k_sum_partial, dim3(256), dim3(256), 0, s->st, s->raw_t, nv, s->sums);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s->st, s->sums, 256, s->sums + 256);
  hipLaunchKernelGGL(k_normalize, dim3(blocks_for(nv)), dim3(256), 0, s->st, s->raw_s, s->raw_t, s->x0, nv, s->sums + 256);
  SHIP(s, hipGetLastError());
  SHIP(s, forward(s, false));
  hipLaunchKernelGGL(k_output, dim3(blocks_for(nv)), dim3(256), 0, s->st, s->x0, s->r, s->out, (const float*)nullptr,
                     (float*)nullptr, 0.0f, (float*)nullptr, (int)nv, kNoCl);
  hipLaunchKernelGGL(k_estimate, dim3(blocks_for(nv)), dim3(256), 0, s->st, s->out, s->est, nv, s->sums + 256, d_decays);
  SHIP(s, hipGetLastError());
  // through pinned staging (one DMA per output), then host copies into the caller's buffers
  SHIP(s, hipMemcpyAsync(s->h_stage, s->est, sizeof(float) * nv, hipMemcpyDeviceToHost, s->st));
  if (scatter_counts) SHIP(s, hipMemcpyAsync(s->h_stage + nv, s->raw_s, sizeof(float) * nv, hipMemcpyDeviceToHost, s->st));
  SHIP(s, hipMemcpyAsync(s->h_decays, d_decays, sizeof(unsigned long long), hipMemcpyDeviceToHost, s->st));
  SHIP(s, hipStreamSynchronize(s->st));
  std::memcpy(estimate, s->h_stage, sizeof(float) * nv);
  if (scatter_counts) std::memcpy(scatter_counts, s->h_stage + nv, sizeof(float) * nv);
  if (decays) *decays = (int64_t)*s->h_decays;
  return MCPET_OK;
}

// Mean-squared-error gradients of the batch (loss = mean over batch and voxels); leaves them
// in the device buffer for an optimiser step.
static mcpet_status batch_gradients(mcpet_surrogate* s, const float* x, const float* target, int32_t batch, float* loss)
{
  const size_t nx = s->vx.size(), nv = s->vr.size();
  hipLaunchKernelGGL(k_zero, dim3(blocks_for(s->n_params)), dim3(256), 0, s->st, s->g, s->n_params);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s->st, s->sse, (size_t)1);
  SHIP(s, hipGetLastError());
  const float scale = 1.0f / ((float)nv * (float)batch);
  for (int b = 0; b < batch; b++) {
    SHIP(s, hipMemcpyAsync(s->x0, x + b * nx, sizeof(float) * nx, hipMemcpyHostToDevice, s->st));
    SHIP(s, hipMemcpyAsync(s->tgt, target + b * nv, sizeof(float) * nv, hipMemcpyHostToDevice, s->st));
    SHIP(s, forward(s, true));
    hipLaunchKernelGGL(k_output, dim3(blocks_for(nv)), dim3(256), 0, s->st, s->x0, s->r, s->out, s->tgt, s->gr, scale, s->sse,
                       (int)nv, ClOut{s->grl, 8, 0});
    SHIP(s, hipGetLastError());
    SHIP(s, backward(s));
  }
  float sse = 0.0f;
  SHIP(s, hipMemcpyAsync(&sse, s->sse, sizeof(float), hipMemcpyDeviceToHost, s->st));
  SHIP(s, hipStreamSynchronize(s->st));
  if (loss) *loss = sse * scale;
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_surrogate_gradients(mcpet_surrogate* s, const float* x, const float* target, int32_t batch,
                                                  float* loss, float* grad)
{
  if (!s || !x || !target || batch < 1) return serr(s, MCPET_E_ARG, "mcpet_surrogate_gradients: bad arguments");
  mcpet_status st = batch_gradients(s, x, target, batch, loss);
  if (st != MCPET_OK || !grad) return st;
  SHIP(s, hipMemcpy(grad, s->g, sizeof(float) * s->n_params, hipMemcpyDeviceToHost));
  return MCPET_OK;
}

// One Adam step (beta 0.9, 0.999, eps 1e-8) on the batch's MSE.
extern "C" mcpet_status mcpet_surrogate_train_step(mcpet_surrogate* s, const float* x, const float* target, int32_t batch,
                                                   float lr, float* loss)
{
  if (!s || !x || !target || batch < 1 || !(lr > 0.0f)) return serr(s, MCPET_E_ARG, "mcpet_surrogate_train_step: bad arguments");
  mcpet_status st = batch_gradients(s, x, target, batch, loss);
  if (st != MCPET_OK) return st;
  s->t_adam++;
  const float c1 = 1.0f - std::pow(0.9f, (float)s->t_adam), c2 = 1.0f - std::pow(0.999f, (float)s->t_adam);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(s->n_params)), dim3(256), 0, s->st, s->p, s->g, s->m, s->v, s->n_params, lr, 0.9f,
                     0.999f, 1e-8f, c1, c2);
  SHIP(s, hipGetLastError());
  SHIP(s, derive(s));
  SHIP(s, hipStreamSynchronize(s->st));
  return MCPET_OK;
}

extern "C" void mcpet_surrogate_destroy(mcpet_surrogate* s)
{
  if (!s) return;
  if (s->st) (void)hipStreamSynchronize(s->st);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
  if (s->h_decays) (void)hipHostFree(s->h_decays);
  if (s->ev_f0) (void)hipEventDestroy(s->ev_f0);
  if (s->ev_f1) (void)hipEventDestroy(s->ev_f1);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

extern "C" const char* mcpet_surrogate_last_error(const mcpet_surrogate* s)
{
  if (s && !s->err.empty()) return s->err.c_str();
  return mcpet_last_error(nullptr);
}

namespace {

// ------------------------------------------------------------------ sinogram rebinning

// One workgroup per (angle group h, plane group d): the sinograms of the plane group (CSR list)
// times the group's angle rows are read as whole radial rows (coalesced), each count added to
// the LDS sum of its radial group. out[(d H + h) W + w] = float(sum).
__global__ __launch_bounds__(256) void k_rebin(const int* __restrict__ sino, const int* __restrict__ grp_off,
                                               const int* __restrict__ grp_sino, int nangles, int nrad, int H, int W, int gh,
                                               int gw, float* __restrict__ out)
{
  __shared__ unsigned long long acc[256];
  const int h = blockIdx.x, d = blockIdx.y;
  for (int i = threadIdx.x; i < W; i += blockDim.x) acc[i] = 0ull;
  __syncthreads();
  const int a0 = h * gh, rows = max(0, min(gh, nangles - a0));
  const int j0 = grp_off[d], nj = grp_off[d + 1] - j0;
  const long long items = (long long)nj * rows * nrad;
  for (long long it = threadIdx.x; it < items; it += blockDim.x) {
    const int r = (int)(it % nrad);
    const long long q = it / nrad;
    const int a = (int)(q % rows), j = (int)(q / rows);
    const int v = sino[((size_t)grp_sino[j0 + j] * nangles + a0 + a) * nrad + r];
    if (v) atomicAdd(&acc[r / gw], (unsigned long long)(unsigned)v);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < W; i += blockDim.x) out[((size_t)d * H + h) * W + i] = (float)acc[i];
}

// The same sums for NRAD <= 256 without atomics: thread r owns radial bin r and adds it up
// over the group's sinogram rows (each row read whole, coalesced), then the radial groups are
// summed from LDS in order.
__global__ __launch_bounds__(256) void k_rebin_rows(const int* __restrict__ sino, const int* __restrict__ grp_off,
                                                    const int* __restrict__ grp_sino, int nangles, int nrad, int H, int W,
                                                    int gh, int gw, float* __restrict__ out)
{
  __shared__ unsigned long long bin[256];
  const int h = blockIdx.x, d = blockIdx.y, r = threadIdx.x;
  const int a0 = h * gh, rows = max(0, min(gh, nangles - a0));
  const int j0 = grp_off[d], nj = grp_off[d + 1] - j0;
  unsigned long long acc = 0ull;
  if (r < nrad)
    for (int j = 0; j < nj; j++) {
      const int* row = sino + ((size_t)grp_sino[j0 + j] * nangles + a0) * nrad + r;
      for (int a = 0; a < rows; a++) acc += (unsigned)row[(size_t)a * nrad];
    }
  bin[r] = acc;
  __syncthreads();
  if (r < W) {
    unsigned long long t = 0ull;
    for (int i = r * gw; i < min(nrad, r * gw + gw); i++) t += bin[i];
    out[((size_t)d * H + h) * W + r] = (float)t;
  }
}

// k_rebin_rows for two tallies of the same geometry in one pass, with the group's sinogram rows
// dealt to RB_SLICES slices of RB_LANES threads (thread r of slice s adds rows j = s, s + RB_SLICES,
// ... of radial bin r), so five times as many rows are read at once; the slices' integer sums
// are added in LDS in slice order. The same integer sums as two k_rebin_rows launches.
#define RB_LANES 192
#define RB_SLICES 5
__global__ __launch_bounds__(RB_LANES * RB_SLICES) void k_rebin_rows2(const int* __restrict__ sa, const int* __restrict__ sb,
                                                                      const int* __restrict__ grp_off,
                                                                      const int* __restrict__ grp_sino, int nangles, int nrad,
                                                                      int H, int W, int gh, int gw, float* __restrict__ oa,
                                                                      float* __restrict__ ob)
{
  __shared__ unsigned long long bin[RB_SLICES][2][RB_LANES];
  const int h = blockIdx.x, d = blockIdx.y;
  const int r = threadIdx.x % RB_LANES, sl = threadIdx.x / RB_LANES;
  const int a0 = h * gh, rows = max(0, min(gh, nangles - a0));
  const int j0 = grp_off[d], nj = grp_off[d + 1] - j0;
  unsigned long long acc_a = 0ull, acc_b = 0ull;
  if (r < nrad && rows > 0) {
    // rows (j, a) of the group, a fastest, dealt round-robin to the slices; the loop is unrolled
    // so that several rows' loads are in flight per thread
    const int items = nj * rows;
#pragma unroll 4
    for (int it = sl; it < items; it += RB_SLICES) {
      const int j = it / rows, a = it - j * rows;
      const size_t off = ((size_t)grp_sino[j0 + j] * nangles + a0 + a) * nrad + r;
      acc_a += (unsigned)sa[off];
      acc_b += (unsigned)sb[off];
    }
  }
  bin[sl][0][r] = acc_a;
  bin[sl][1][r] = acc_b;
  __syncthreads();
  if (sl == 0 && r < W) {
    unsigned long long ta = 0ull, tb = 0ull;
    for (int i = r * gw; i < min(nrad, r * gw + gw); i++)
      for (int k = 0; k < RB_SLICES; k++) {
        ta += bin[k][0][i];
        tb += bin[k][1][i];
      }
    oa[((size_t)d * H + h) * W + r] = (float)ta;
    ob[((size_t)d * H + h) * W + r] = (float)tb;
  }
}

__global__ void k_i2f2(const int* __restrict__ a, const int* __restrict__ b, float* __restrict__ oa, float* __restrict__ ob,
                       size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  oa[i] = (float)a[i];
  ob[i] = (float)b[i];
}

}  // namespace

hipError_t mcpet_launch_i2f2(const int* a, const int* b, float* oa, float* ob, size_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_i2f2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, oa, ob, n);
  return hipGetLastError();
}

hipError_t mcpet_launch_rebin2(const int* sa, const int* sb, const int* grp_off, const int* grp_sino, int nangles, int nrad, int D,
                               int H, int W, int gh, int gw, float* oa, float* ob, hipStream_t st)
{
  if (W > RB_LANES || nrad > RB_LANES) {
    hipError_t e = mcpet_launch_rebin(sa, grp_off, grp_sino, nangles, nrad, D, H, W, gh, gw, oa, st);
    return e == hipSuccess ? mcpet_launch_rebin(sb, grp_off, grp_sino, nangles, nrad, D, H, W, gh, gw, ob, st) : e;
  }
  hipLaunchKernelGGL(k_rebin_rows2, dim3(H, D), dim3(RB_LANES * RB_SLICES), 0, st, sa, sb, grp_off, grp_sino, nangles, nrad, H,
                     W, gh, gw, oa, ob);
  return hipGetLastError();
}

hipError_t mcpet_launch_rebin(const int* sino, const int* grp_off, const int* grp_sino, int nangles, int nrad, int D, int H, int W,
                              int gh, int gw, float* out, hipStream_t st)
{
  if (W > 256) return hipErrorInvalidValue;
  if (nrad <= 256)
    hipLaunchKernelGGL(k_rebin_rows, dim3(H, D), dim3(256), 0, st, sino, grp_off, grp_sino, nangles, nrad, H, W, gh, gw, out);
  else
    hipLaunchKernelGGL(k_rebin, dim3(H, D), dim3(256), 0, st, sino, grp_off, grp_sino, nangles, nrad, H, W, gh, gw, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ self test of the GEMM

extern "C" mcpet_status mcpet_selftest_gemm_bf16(const float* a, const float* b, int32_t M, int32_t N, int32_t K, int32_t split,
                                                 float* c)
{
  if (!a || !b || !c || M < 1 || N < 1 || K < 1 || split < 1) {
    mcpet_host_set_error("selftest_gemm_bf16: bad args");
    return MCPET_E_ARG;
  }
  float *da = nullptr, *db = nullptr, *dc = nullptr;
  hipError_t e = hipMalloc((void**)&da, 4 * (size_t)M * K);
  if (e == hipSuccess) e = hipMalloc((void**)&db, 4 * (size_t)K * N);
  if (e == hipSuccess) e = hipMalloc((void**)&dc, 4 * (size_t)M * N);
  if (e == hipSuccess) e = hipMemcpy(da, a, 4 * (size_t)M * K, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b, 4 * (size_t)K * N, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dc, 0, 4 * (size_t)M * N);
  if (e == hipSuccess) e = gemm(Dense{da, M, K, K}, Dense{db, K, N, N}, dc, M, N, K, N, nullptr, split, false, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(c, dc, 4 * (size_t)M * N, hipMemcpyDeviceToHost);
  if (da) (void)hipFree(da);
  if (db) (void)hipFree(db);
  if (dc) (void)hipFree(dc);
  if (e != hipSuccess) {
    mcpet_host_set_error((std::string("selftest_gemm_bf16: ") + hipGetErrorString(e)).c_str());
    return MCPET_E_HIP;
  }
  return MCPET_OK;
}
