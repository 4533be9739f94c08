// kparams.h — argument block of the transport kernels (host and device view).
#ifndef MCPET_KPARAMS_H
#define MCPET_KPARAMS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mcpet.h"
#include "ranecu.h"

// Event counters accumulated by the kernels (order of mcpet_stats after the two times).
enum KCounter {
  KC_DECAYS = 0,
  KC_STREAMS,
  KC_WOODCOCK,
  KC_MATCHANGE,
  KC_COMPTON,
  KC_SHELLEVAL,
  KC_RAYLEIGH,
  KC_PHOTO,
  KC_PAIRS,
  KC_COINC,
  KC_RNDRAWS,
  KC_PHASE_CYCLES = 16, // per-phase shader clock cycles, summed over waves (MCPET_SM_STATS builds)
  KC_PHASE_EXEC = 24,   // per-phase executions of the persistent kernel (MCPET_SM_STATS builds)
  KC_PHASE_LANES = 32,  // per-phase sum of active lanes (MCPET_SM_STATS builds)
  KC_N = 40
};

struct KParams {
  // phantom: packed voxel words (density, material index 0..14 as int bits), x fastest
  const float2* vox;
  int nx, ny, nz;
  float ivx, ivy, ivz;         // inverse voxel size
  float bx, by, bz;            // bounding box size
  // decay streams: stream s -> emitting voxel active_vox[s / tpb], lane s % tpb
  const int* active_vox;
  const float* active_act;
  long long n_streams;
  int tpb;
  unsigned long long* queue;   // persistent kernel: next stream index
  // source
  int seed;
  int bg_mode;
  unsigned long long acq_ps;
  double inv_mean_life;
  uint32_t j1[MCR_JUMP_BITS];
  uint32_t j2[MCR_JUMP_BITS];
  // interaction tables
  const float2* woodcock;      // [N]          (a, b)
  const float2* tot;           // [N][15]      (a_total, b_total)
  const float4* cr;            // [N][15]      (a_compton, b_compton, a_rayleigh, b_rayleigh)
  const mcpet_rayleigh_table* ray;
  const mcpet_compton_table* cmp;
  int num_values;
  float e0, ide;
  // detector
  float cx, cy, cz, radius, height;
  int psf_size, mode, type;
  // sinogram
  float e_resol, e_low, e_high, fovz;
  int nrows, ncrystals, nangles, nrad, nsinos, nzs, ne, mrd, span;
  // dose
  int material_dose;
  int voxel_dose;
  short roi[6];
  // tallies
  int* trues;
  int* scatter;
  int* img_t;
  int* img_sc;
  int* spectrum;
  mcpet_psf_record* psf;
  int* psf_index;
  unsigned long long* voxel_edep;
  unsigned long long* mat_edep;
  unsigned long long* counters;  // KC_N
  int* final_seed;
  long long last_thread_id;      // reference thread id whose final seed is stored (K:454-455)
  int lds_spectrum;              // 1: privatise the spectrum in LDS
  // Phase selection of the persistent kernel: 0 = phase held by most lanes;
  // 1 = phase with the fewest wasted lane-slots, (64 - lanes) * phase_cost.
  int sm_policy;
  int sm_cost[8];                // relative cost per phase (DECAY, SOURCE2, STEP, INTERACT, TALLY, CTAU, CPZ)
};

// Launchers implemented in kernels.hip.
hipError_t mcpet_launch_transport(const KParams& P, int kind, int count_events, int blocks, int block_size, hipStream_t stream);
hipError_t mcpet_launch_selftest_ranecu(int seed, const long long* d_ids, int nstreams, int n, float* d_out, double* d_out_d,
                                        int as_double, const uint32_t* j1, const uint32_t* j2, hipStream_t stream);
hipError_t mcpet_launch_selftest_fmath(const double* d_x, const double* d_y, int n, double* d_od, float* d_of, hipStream_t stream);
hipError_t mcpet_launch_copy(const uint4* src, uint4* dst, size_t n16, hipStream_t stream);
int mcpet_transport_regs_hint();

#endif
