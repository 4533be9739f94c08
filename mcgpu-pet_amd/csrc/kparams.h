// kparams.h — argument block of the transport kernels (host and device view).
#ifndef MCPET_KPARAMS_H
#define MCPET_KPARAMS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

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
  KC_EMIT_PAIR,         // annihilation pairs emitted
  KC_EMIT_PROMPT,       // prompt gammas emitted
  KC_PHASE_CYCLES = 16, // per-phase shader clock cycles, summed over waves (MCPET_SM_STATS builds)
  KC_PHASE_EXEC = 16 + MCPET_NPHASES,       // per-phase executions (MCPET_SM_STATS builds)
  KC_PHASE_LANES = 16 + 2 * MCPET_NPHASES,  // per-phase sum of active lanes (MCPET_SM_STATS builds)
  KC_N = 16 + 3 * MCPET_NPHASES
};

struct KParams {
  // phantom, x fastest. vox_fmt 0: one packed word (density, material index 0..14 as int
  // bits) per voxel; 1: u8 index into a palette of such words (palette copied to LDS);
  // 2: u16 index into the palette (read from global memory).
  const float2* vox;
  const uint8_t* vox8;
  const uint16_t* vox16;
  const float2* palette;
  int vox_fmt, n_palette;
  int nx, ny, nz;
  double inv_nx, inv_ny;       // 1 / nx, 1 / ny: voxel coordinates by multiplication (src_coords)
  float ivx, ivy, ivz;         // inverse voxel size
  float bx, by, bz;            // bounding box size
  // decay streams: stream s -> emitting voxel active_vox[s / tpb], lane s % tpb; with a
  // balanced layout (stream_start != null) voxel k owns streams [stream_start[k],
  // stream_start[k + 1]) whose random-number stream ids start at stream_id[k]
  const int* active_vox;
  const float* active_act;
  const long long* stream_start;
  const long long* stream_id;
  // balanced layout: stream_hint[b] = the voxel owning stream b << hint_shift (b = 0 .. last
  // block; one more entry = n_active - 1), so the owner of stream s lies between
  // stream_hint[s >> hint_shift] and the next entry (null: search all voxels)
  const int* stream_hint;
  int hint_shift;
  int n_active;
  long long n_streams;           // streams [stream_base, n_streams) are simulated by a launch
  long long stream_base;
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
  // rebinned tallies (null: off): trues and scatter also added on the D x H x W grid of the
  // sinogram rebinning, bin (rb_plane[sinogram] H + angle / rb_gh) W + radial / rb_gw
  int* rb_trues;
  int* rb_scatter;
  const int* rb_plane;
  int rb_H, rb_W, rb_gh, rb_gw;
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
  int lds_palette;               // 1: u8 palette copied to LDS (single-phantom kernels)
  int lds_xs511;                 // 1: the 511 keV total-cross-section row read from LDS
  long long stream_stride;       // > 1: the k-th claimed stream is stream_base + k * stride (sampling)
  // Decay branching (drawn per decay only when active) and the singles list.
  float positron_frac;           // < 1: a decay emits the annihilation pair with this probability
  float prompt_frac;             // > 0: a decay emits a prompt gamma with this probability
  float prompt_energy;           // eV
  mcpet_single* singles;
  unsigned long long* singles_n; // singles detected (may exceed the capacity)
  long long singles_cap;
  // coincidence pairing of the sorted singles
  unsigned long long window_ps;
  int* randoms;
  int* spurious;
  unsigned long long* coinc_counts;  // binned trues, scatter, randoms, spurious; multiples
  // Phase selection of the persistent kernel: 0 = phase held by most lanes;
  // 1 = phase with the fewest wasted lane-slots, (64 - lanes) * phase_cost.
  int sm_policy;
  int sm_cost[MCPET_NPHASES];    // relative cost per phase (DECAY, SOURCE2, STEP, INTERACT, TALLY, CTAU, CPZ, DETECT, SOURCE3)
};

// Launchers implemented in kernels.hip.
hipError_t mcpet_launch_transport(const KParams& P, const KParams* d_P, int kind, int count_events, int blocks, int block_size,
                                  hipStream_t stream);
hipError_t mcpet_launch_transport_batch(const KParams* d_params, int n, bool ext, int count_events, int blocks, int block_size,
                                        hipStream_t stream);
hipError_t mcpet_launch_fill_u32(void* p, size_t n_words, hipStream_t stream);
// zero n regions of n_words[i] 4-byte words at p[i], K_FILL_MAX regions per launch
#define K_FILL_MAX 16
hipError_t mcpet_launch_fill_regions(void* const* p, const size_t* n_words, int n, hipStream_t stream);
// Activity and attenuation line integrals on a rebinned sinogram grid (scatter surrogate
// inputs): `map` is an nvox (activity, mu) scratch map; lors/off the crystal pairs of each
// (angle, radial) group in CSR form; D <= 256.
hipError_t mcpet_launch_project(const KParams* d_params, const int* active_vox, const float* active_act, long long n_active,
                                float2* map, size_t nvox, const int2* lors, const int* off, int D, int H, int W, int gz,
                                float step, float* act_out, float* att_out, hipStream_t stream);
// a[0..n) times float32(1 / max(a)) (a >= 0), in place; mx: one device word of scratch.
hipError_t mcpet_launch_scale_by_max(float* a, int n, unsigned* mx, hipStream_t stream);
hipError_t mcpet_launch_selftest_ranecu(int seed, const long long* d_ids, int nstreams, int n, float* d_out, double* d_out_d,
                                        int as_double, const uint32_t* j1, const uint32_t* j2, hipStream_t stream);
hipError_t mcpet_launch_selftest_fmath(const double* d_x, const double* d_y, int n, double* d_od, float* d_of, hipStream_t stream);
hipError_t mcpet_launch_copy(const uint4* src, uint4* dst, size_t n16, hipStream_t stream);
int mcpet_transport_regs_hint();

// Singles pipeline (coinc.hip).
long long mcpet_rs_tiles(long long n);
void mcpet_scan_levels(long long m, std::vector<long long>& levels);
hipError_t mcpet_launch_keys_from_singles(const mcpet_single* singles, long long n, unsigned long long* keys, unsigned* vals,
                                         hipStream_t st);
hipError_t mcpet_launch_iota(unsigned* vals, long long n, hipStream_t st);
hipError_t mcpet_radix_sort(long long n, int passes, unsigned long long* keys[2], unsigned* vals[2], unsigned* hist,
                            unsigned* const* level_bufs, int* which, hipStream_t st);
hipError_t mcpet_launch_pairing(const KParams& P, const unsigned long long* keys, const unsigned* perm, long long n, hipStream_t st);
hipError_t mcpet_launch_gather_singles(const mcpet_single* s, const unsigned* perm, long long n, mcpet_single* out, hipStream_t st);

// Scatter surrogate (unet.hip): sum sinogram counts into a D x H x W grid (plane groups listed
// in CSR form grp_off[D + 1] / grp_sino, angle groups of gh rows, radial groups of gw bins).
hipError_t mcpet_launch_rebin(const int* sino, const int* grp_off, const int* grp_sino, int nangles, int nrad, int D, int H, int W,
                              int gh, int gw, float* out, hipStream_t st);
// oa[i] = float(a[i]), ob[i] = float(b[i]) for i < n (rebinned tallies to the network's input)
hipError_t mcpet_launch_i2f2(const int* a, const int* b, float* oa, float* ob, size_t n, hipStream_t st);
// the same for two tallies (sa -> oa, sb -> ob) in one launch
hipError_t mcpet_launch_rebin2(const int* sa, const int* sb, const int* grp_off, const int* grp_sino, int nangles, int nrad, int D,
                               int H, int W, int gh, int gw, float* oa, float* ob, hipStream_t st);

#endif
