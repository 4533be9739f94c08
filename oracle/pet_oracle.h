/*
 * pet_oracle.h — CPU restatement of the MCGPU-PET photon-transport path.
 *
 * TEST INFRASTRUCTURE ONLY. This code is the checker for the MI355X library: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and by nothing else. The product (libmcpet.so) never links, loads or calls it.
 *
 * Each function restates one reference function (file:line under the reference tree,
 * K = MCGPU-PET_kernel.cu, C = MCGPU-PET.cu, H = MCGPU-PET.h). One reference GPU
 * thread (voxel, threadIdx) is emulated sequentially with its own RANECU stream, so
 * the sequence of random numbers and arithmetic per stream is the reference's.
 *
 * Two builds: liboracle_libm.so uses the C library's logf/expf/... (the reference used
 * CUDA's fast-math versions, which cannot be reproduced bit-for-bit on a CPU);
 * liboracle_fmath.so uses the library's deterministic elementary functions
 * (fmath.h), so that whole histories can be compared bit-for-bit with the GPU.
 */
#ifndef PET_ORACLE_H
#define PET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MATERIALS 15
#define ORC_MAX_SHELLS 30
#define ORC_NP_RAYLEIGH 128
#define ORC_MAX_ENERGYBINS_RAYLEIGH 25005

/* H:347-360 */
typedef struct orc_rayleigh {
  float xco[ORC_NP_RAYLEIGH * ORC_MAX_MATERIALS], pco[ORC_NP_RAYLEIGH * ORC_MAX_MATERIALS],
      aco[ORC_NP_RAYLEIGH * ORC_MAX_MATERIALS], bco[ORC_NP_RAYLEIGH * ORC_MAX_MATERIALS],
      pmax[ORC_MAX_ENERGYBINS_RAYLEIGH * ORC_MAX_MATERIALS];
  unsigned char itlco[ORC_NP_RAYLEIGH * ORC_MAX_MATERIALS], ituco[ORC_NP_RAYLEIGH * ORC_MAX_MATERIALS];
} orc_rayleigh;

/* H:334-344 */
typedef struct orc_compton {
  float fco[ORC_MAX_MATERIALS * ORC_MAX_SHELLS], uico[ORC_MAX_MATERIALS * ORC_MAX_SHELLS],
      fj0[ORC_MAX_MATERIALS * ORC_MAX_SHELLS];
  int noscco[ORC_MAX_MATERIALS];
} orc_compton;

/* Counters reported by orc_run (same meaning as the library's mcpet_stats). */
enum {
  ORC_C_DECAYS = 0, ORC_C_STREAMS, ORC_C_WOODCOCK, ORC_C_MATCHANGE, ORC_C_COMPTON, ORC_C_SHELLEVAL,
  ORC_C_RAYLEIGH, ORC_C_PHOTO, ORC_C_PAIRS, ORC_C_COINC, ORC_C_RNDRAWS, ORC_C_EMIT_PAIR, ORC_C_EMIT_PROMPT,
  ORC_C_N = 16
};

typedef struct orc_sim {
  /* phantom (load_voxels layout, C:1522-1524) */
  int nx, ny, nz;
  float inv_voxel_size[3], size_bbox[3];
  const float* voxel_mat_dens;           /* 3*Nvox */
  /* tables (load_material layouts) */
  int num_values;
  float e0, ide;
  const float* woodcock;                 /* 2*N */
  const float* mfp_a;                    /* 3*N*15 */
  const float* mfp_b;
  const orc_rayleigh* rayleigh;
  const orc_compton* compton;
  /* source */
  uint64_t acquisition_time_ps;
  float mean_life;
  int threads_per_block;
  int seed;
  int bg_mode;                           /* PETA_DEV[0]==0 */
  /* detector (after the centring rule) */
  float det_center[3], det_radius, det_height;
  int psf_size, tally_psf_sinogram, tally_type;
  /* sinogram parameters */
  float e_resol, e_low, e_high, fovz;
  int nrows, ncrystals, nangles, nrad, nzs, nbins, ne, mrd, span, nsinos;
  /* dose */
  int material_dose;                     /* materials_dose != NULL */
  short roi[6];                          /* x_min,x_max,y_min,y_max,z_min,z_max (0-based) */
  /* outputs (caller-allocated, zeroed) */
  int* trues;
  int* scatter;
  int* image_trues;
  int* image_scatter;
  int* spectrum;
  unsigned char* psf;                    /* psf_size * 44 bytes, may be NULL */
  unsigned long long* voxel_edep;        /* 2*N_ROI, may be NULL */
  unsigned long long* material_edep;     /* 2*15, may be NULL */
  /* scalar outputs */
  int psf_index;
  unsigned long long total_histories;
  int final_seed;
  int final_seed_written;
  unsigned long long counters[ORC_C_N];
  /* execution */
  int nthreads;                          /* CPU worker threads (>=1) */
  int voxel_stride;                      /* simulate every k-th voxel (1 = all) */
  int voxel_offset;
  double wall_seconds;                   /* transport wall time */
  /* extended decay model and singles list (no reference counterpart; the library's
   * mcpet_config extensions): effective positron branching (1 = every decay), prompt-gamma
   * branching and energy, and the singles list (32-byte records, see mcpet_single) */
  float positron_fraction;
  float prompt_fraction;
  float prompt_energy;
  unsigned char* singles;
  long long singles_cap;
  long long singles_n;
  /* balanced decay-stream layout (the library's stream_decays extension): 0 = the reference's
   * threads_per_block streams per emitting voxel; > 0: max(1, ceil(E_v / stream_decays))
   * streams for voxel v, E_v = activity * tau (1 - exp(-T / tau)), stream ids numbered in voxel
   * order. stream_first is filled by orc_run (NULL otherwise). */
  int stream_decays;
  long long* stream_first;               /* nvox + 1 entries while orc_run runs */
} orc_sim;

/* RANECU and skip-ahead (K:898-1083, C:2906-2935) */
int orc_abMODm(int m, int a, int s);
void orc_init_PRNG(long long history_batch, int histories_per_thread, int seed_input, int* sx, int* sy);
float orc_ranecu(int* sx, int* sy);
double orc_ranecu_double(int* sx, int* sy);
void orc_update_seed_PRNG(int batch_number, unsigned long long total_histories, int* seed);
void orc_ranecu_block(int seed, const long long* ids, int nstreams, int n, float* out, double* out_d, int as_double);

/* Input-side restatements (C:1176-1189, C:1565-1831) */
void orc_sinogram_sizes(int mrd, int span, int nzs, int nrad, int nangles, int* nseg, int* nsinos, int* nbins);
int orc_load_material(const char* names /* 15*250 */, const float* density_max /* 15, modified */,
                      float* density_nominal, int* num_values, float* e0, float* ide, float** woodcock,
                      float** mfp_a, float** mfp_b, orc_rayleigh* ray, orc_compton* cmp);
/* The same with the energy grid continued past the files' last point up to e_max (eV). */
int orc_load_material_ext(const char* names, const float* density_max, double e_max, float* density_nominal,
                          int* num_values, float* e0, float* ide, float** woodcock, float** mfp_a, float** mfp_b,
                          orc_rayleigh* ray, orc_compton* cmp);
void orc_free(void* p);

/* Whole transport (K:153-463) */
int orc_run(orc_sim* s);

/* Unit probes used by the physics tests (each runs the restated function n times). */
int orc_locate_voxel(const orc_sim* s, float x, float y, float z);
void orc_probe_rotate(int n, const float* dir_in, const double* costh, const double* phi, float* dir_out);
void orc_probe_gcoa(const orc_sim* s, int seed, long long stream, int n, float energy, int mat, float* e_out,
                    double* costh_out);
void orc_probe_graa(const orc_sim* s, int seed, long long stream, int n, float energy, int mat, double* costh_out);
void orc_probe_source(const orc_sim* s, int seed, long long stream, int n, int ix, int iy, int iz, float* pos1,
                      float* dir1, float* dir2);
/* Transport single photons from pos/dir (energy E) until the first real interaction or exit:
 * writes the first real-interaction distance (or -1 if escaped) and its type
 * (1 Compton, 2 Rayleigh, 3 photoelectric, 0 escaped). */
void orc_probe_first_interaction(const orc_sim* s, int seed, long long stream, int n, const float* pos,
                                 const float* dir, float energy, float* dist, int* type);
/* Sinogram bin of a crystal pair exactly as tally_Sinogram (K:1749-1788); -1 if rejected. */
int orc_sinogram_bin(const orc_sim* s, int ix1, int iz1, int ix2, int iz2);
/* Full detector geometry of tally_Sinogram part 1 from photon end states (no RN use). */
int orc_detector_bin(const orc_sim* s, const float* pos, const float* dir, const float* posf, const float* dirf);

#ifdef __cplusplus
}
#endif
#endif
