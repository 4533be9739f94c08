/*
 * mcpet.h — C ABI of the MI355X Monte Carlo PET photon-transport library (libmcpet.so).
 *
 * The reference (MCGPU-PET v0.1) has no library or FFI: its photon-transport path
 * sits behind (1) the launch of track_particles (MCGPU-PET.cu:452-487) with the
 * device buffers prepared by init_CUDA_device (MCGPU-PET.cu:1842-2146) and
 * (2) the CLI/file contract (MCGPU-PET.in, .vox, .mcgpu.gz in; sinograms, images,
 * spectrum, PSF, dose out). Every entry point below replaces one piece of that,
 * cited next to it. Layouts of the inputs are the reference's, so its files drop in.
 *
 * Conventions
 *   - Plain C types only; no exceptions cross the boundary; every call returns a
 *     status code (the reference calls exit(-1|-2|-3) instead).
 *   - The library owns device memory; the caller owns host buffers it passes in.
 *   - One context per host thread; a context owns one HIP stream and is not
 *     thread-safe. Distinct contexts may be used from distinct threads.
 */
#ifndef MCPET_H
#define MCPET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compile-time limits of the reference (MCGPU-PET.h:96-99). */
#define MCPET_MAX_MATERIALS 15
#define MCPET_MAX_SHELLS 30
#define MCPET_NP_RAYLEIGH 128
#define MCPET_MAX_ENERGYBINS_RAYLEIGH 25005
#define MCPET_NAME_LEN 250
#define MCPET_NPHASES 10               /* phase slots of the persistent kernel's statistics */

typedef enum mcpet_status {
  MCPET_OK = 0,
  MCPET_E_ARG = 1,      /* invalid argument / inconsistent sizes          */
  MCPET_E_IO = 2,       /* file missing or unreadable   (ref exit(-1|-2)) */
  MCPET_E_FORMAT = 3,   /* malformed input file          (ref exit(-2))   */
  MCPET_E_RANGE = 4,    /* value out of the supported range              */
  MCPET_E_NOMEM = 5,    /* host or device allocation failed              */
  MCPET_E_HIP = 6,      /* HIP runtime error (no device, launch failure) */
  MCPET_E_STATE = 7,    /* call out of order (e.g. fetch before run)     */
  MCPET_E_CAPACITY = 8  /* a fixed-capacity buffer was exceeded          */
} mcpet_status;

/* ------------------------------------------------------------------ data layouts */

/* Cylindrical detector: reference detector_struct (MCGPU-PET.h:220-232). */
typedef struct mcpet_detector {
  float center[3];            /* PSF_center (cm)                                 */
  float radius;               /* PSF_radius (cm); <0 in the .in means "centre on bbox" */
  float height;               /* PSF_height (cm)                                 */
  int32_t psf_size;           /* PSF capacity in records                          */
  int32_t tally_psf_sinogram; /* 0 both, 1 PSF only, 2 sinogram only              */
  int32_t tally_type;         /* 0 both, 1 trues only, 2 scatter only             */
} mcpet_detector;

/* Phase-space record: reference PSF_element_struct, packed, 44 bytes (MCGPU-PET.h:235-247). */
#pragma pack(push, 4)
typedef struct mcpet_psf_record {
  uint64_t emission_time_ps;
  float travel_time_ps;
  int32_t emission_absvox;
  float energy, z, phi, vx, vy, vz;
  int16_t index1;             /* scatter state 0 none,1 Compton,2 Rayleigh,3 multiple */
  int16_t index2;             /* 0 for a 511 keV annihilation photon                  */
} mcpet_psf_record;
#pragma pack(pop)

/* Rayleigh sampling data: reference rayleigh_struct (MCGPU-PET.h:347-360). */
typedef struct mcpet_rayleigh_table {
  float xco[MCPET_NP_RAYLEIGH * MCPET_MAX_MATERIALS];
  float pco[MCPET_NP_RAYLEIGH * MCPET_MAX_MATERIALS];
  float aco[MCPET_NP_RAYLEIGH * MCPET_MAX_MATERIALS];
  float bco[MCPET_NP_RAYLEIGH * MCPET_MAX_MATERIALS];
  float pmax[MCPET_MAX_ENERGYBINS_RAYLEIGH * MCPET_MAX_MATERIALS];
  uint8_t itlco[MCPET_NP_RAYLEIGH * MCPET_MAX_MATERIALS];
  uint8_t ituco[MCPET_NP_RAYLEIGH * MCPET_MAX_MATERIALS];
} mcpet_rayleigh_table;

/* Compton shell data: reference compton_struct (MCGPU-PET.h:334-344). */
typedef struct mcpet_compton_table {
  float fco[MCPET_MAX_MATERIALS * MCPET_MAX_SHELLS];
  float uico[MCPET_MAX_MATERIALS * MCPET_MAX_SHELLS];
  float fj0[MCPET_MAX_MATERIALS * MCPET_MAX_SHELLS];
  int32_t noscco[MCPET_MAX_MATERIALS];
} mcpet_compton_table;

/* Everything read_input returns (MCGPU-PET.cu:866-1280). */
typedef struct mcpet_config {
  int32_t seed;                  /* RANECU seed; 0 in the file means time-based (C:920-928) */
  int32_t gpu_id;
  int32_t threads_per_block;     /* "GPU THREADS PER CUDA BLOCK": decay streams per voxel */
  float fact_object;             /* density scale factor (C:938,1494)                      */
  uint64_t acquisition_time_ps;  /* (uint64)(T*1e12+0.5) (C:977)                            */
  float mean_life;               /* isotope mean life (s)                                   */
  float activity[MCPET_MAX_MATERIALS]; /* parsed but unused by the transport (C:987-995)    */
  char psf_file[MCPET_NAME_LEN];
  mcpet_detector detector;
  int32_t material_dose;         /* 1 tally dose per material                               */
  int32_t voxel_dose;            /* 1 tally dose per voxel in the ROI                       */
  char dose_file[MCPET_NAME_LEN];
  int16_t roi_x_min, roi_x_max, roi_y_min, roi_y_max, roi_z_min, roi_z_max; /* 0-based */
  float e_resol, e_low, e_high;  /* energy resolution (FWHM fraction), window in eV         */
  float fovz;
  int32_t nrows, ncrystals, nangles, nrad, nzs, res, ne, mrd, span;
  int32_t nseg, nsinos, nbins, nvoxs; /* derived (C:1176-1189)                               */
  char voxel_file[MCPET_NAME_LEN];
  char material_files[MCPET_MAX_MATERIALS][MCPET_NAME_LEN];
  int32_t bg_mode;               /* 1: second photon isotropic ("BG" run, C:81-84)          */
  int32_t voxel_format;          /* device voxel layout: 0 auto (palette indices when the phantom
                                    has <= 65536 distinct (material, density) pairs), 1 always the
                                    8-byte (density, material) word. Results are identical.    */
  /* Extensions beyond the reference; all-zero values give the reference's behaviour.
   * tally_psf_sinogram = MCPET_TALLY_SINGLES selects the singles pipeline: the transport
   * writes every detected photon to a singles list, which is sorted by detection time and
   * paired with a coincidence timing window into trues, scatter, randoms and spurious
   * (prompt-gamma) coincidences. The reference's files cannot request it. */
  float positron_fraction;       /* fraction of decays emitting an annihilation pair (<= 0: 1) */
  float prompt_gamma_fraction;   /* fraction of decays emitting a prompt gamma (0: none)       */
  float prompt_gamma_energy;     /* eV, inside the interaction tables                          */
  int32_t stream_decays;         /* 0: the reference's threads_per_block decay streams per emitting
                                    voxel. > 0: voxel v gets max(1, ceil(E_v / stream_decays))
                                    streams, E_v = activity * tau (1 - exp(-T / tau)) its expected
                                    decays over acquisition_time_ps, each at activity / streams;
                                    stream ids count the streams in voxel order. Balances the work
                                    of hot voxels (statistically equal, other random numbers).  */
  int64_t singles_capacity;      /* records of the device singles list                         */
  uint64_t coincidence_window_ps;/* coincidence timing window (ps)                             */
} mcpet_config;

/* tally_psf_sinogram value of the singles pipeline (the reference uses 0, 1, 2). */
#define MCPET_TALLY_SINGLES 3

/* One detected photon of the singles pipeline (32 bytes). */
typedef struct mcpet_single {
  uint64_t time_ps;              /* emission time + flight time to the detector (ps)           */
  uint64_t event;                /* decay identifier: (reference thread id << 32) | decay number */
  float energy;                  /* energy after the detector's energy blur (eV)               */
  float z, phi;                  /* hit on the detector cylinder (cm, rad), relative to its centre */
  int16_t scatter;               /* 0 none, 1 Compton, 2 Rayleigh, 3 multiple                  */
  int16_t kind;                  /* 0, 1: annihilation photon 1, 2; 2: prompt gamma             */
} mcpet_single;

/* Voxelized phantom (load_voxels, MCGPU-PET.cu:1381-1533). */
typedef struct mcpet_phantom {
  int32_t nx, ny, nz;
  float voxel_size[3];
  float inv_voxel_size[3];       /* 1/voxel_size as float (C:1460-1462)      */
  float size_bbox[3];            /* n*voxel_size as float (C:1433-1435)      */
  float density_max[MCPET_MAX_MATERIALS]; /* -999 where a material is absent */
  float* voxel_mat_dens;         /* 3*nx*ny*nz floats, x fastest: (material+0.0001, density, activity Bq) */
  int32_t owned;                 /* 1 if voxel_mat_dens was allocated by the library */
} mcpet_phantom;

/* Interaction tables (load_material, MCGPU-PET.cu:1565-1831). */
typedef struct mcpet_tables {
  int32_t num_values;            /* energy grid points (10200 in the sample files) */
  float e0, ide;                 /* first energy (eV), 1/bin width (1/eV)          */
  float density_nominal[MCPET_MAX_MATERIALS]; /* -1 for unused slots             */
  float* woodcock;               /* 2*num_values: (a, b), mfp_W(E) = a + E b         */
  float* mfp_a;                  /* 3*num_values*15, [bin][mat]: (total, Compton, Rayleigh) */
  float* mfp_b;                  /* same layout, slopes                             */
  mcpet_rayleigh_table* rayleigh;
  mcpet_compton_table* compton;
  int32_t owned;
  int32_t num_file_values;       /* grid points read from the files; points beyond are extrapolated
                                    (mcpet_load_materials_extended), 0 when none are   */
} mcpet_tables;

/* Results copied back by mcpet_fetch (reference D2H copies, MCGPU-PET.cu:499-535,642-652). */
typedef struct mcpet_tallies {
  int32_t* trues;                /* nbins     (may be NULL to skip)    */
  int32_t* scatter;              /* nbins                              */
  int32_t* image_trues;          /* nx*ny*nz                           */
  int32_t* image_scatter;        /* nx*ny*nz                           */
  int32_t* spectrum;             /* ne                                 */
  mcpet_psf_record* psf;         /* psf_capacity records (may be NULL) */
  int64_t psf_capacity;
  uint64_t* voxel_edep;          /* 2*N_ROI (Edep, Edep^2) pairs       */
  uint64_t* material_edep;       /* 2*15                                */
  /* scalars, always filled */
  uint64_t total_histories;      /* photons: 2 per decay (K:282)       */
  int64_t psf_count;             /* records written (index_PSF, clamped to capacity) */
  int64_t psf_requested;         /* records that reached the PSF stage (index_PSF before clamp) */
  int64_t sum_trues, sum_scatter;
  int32_t final_seed;            /* seed.x of the last stream (K:454-455); 0 if not written */
  /* singles pipeline (tally_psf_sinogram = MCPET_TALLY_SINGLES); trues/scatter above hold
     its true and scattered coincidences */
  int32_t* randoms;              /* nbins: coincidences of singles from different decays    */
  int32_t* spurious;             /* nbins: a prompt gamma paired with a photon of its decay  */
  mcpet_single* singles;         /* the singles list sorted by detection time (may be NULL) */
  int64_t singles_buffer;        /* capacity of `singles` in records                          */
  int64_t singles_count;         /* singles stored (clamped to the device capacity)          */
  int64_t singles_requested;     /* singles detected (before the clamp)                      */
  int64_t coincidences[5];       /* binned trues, scatter, randoms, spurious; clusters of more
                                    than two singles (multiples, discarded)                  */
} mcpet_tallies;

/* Transport launch variants. */
typedef enum mcpet_kernel_kind {
  MCPET_KERNEL_AUTO = 0,
  MCPET_KERNEL_PLAIN = 1,        /* one workgroup per emitting voxel, lanes = decay streams */
  MCPET_KERNEL_PERSISTENT = 2    /* persistent waves pulling streams from a device queue    */
} mcpet_kernel_kind;

typedef struct mcpet_run_params {
  int32_t kernel;                /* mcpet_kernel_kind                                  */
  int32_t block_size;            /* workgroup size for the persistent kernel (0: auto) */
  int32_t blocks_per_cu;         /* persistent grid = CUs * blocks_per_cu (0: auto)    */
  int32_t zero_tallies;          /* 1: clear tallies before the run (default)          */
  int32_t seed;                  /* 0: use the config seed                             */
  uint64_t acquisition_time_ps;  /* 0: use the config value                            */
  int32_t phase_policy;          /* persistent kernel: 0 auto, 1 most lanes, 2 least waste */
  int32_t phase_cost[MCPET_NPHASES]; /* relative phase costs for policy 2 (0: built-in) */
  int32_t count_events;          /* 1: instrumented persistent kernel that counts every event
                                    in mcpet_stats (same tallies; slower). 0: only decays,
                                    streams, pairs_tallied and coincidences are counted. The
                                    plain kernel always counts every event.              */
  int32_t use_graph;             /* batches: replay the clearing + launch as a HIP graph  */
  int32_t lds_layout;            /* persistent kernel, single context: 0 = the autotuned or
                                    default layout; MCPET_LDS_SET | MCPET_LDS_* bits = that
                                    layout (tallies are identical in every layout)        */
} mcpet_run_params;

/* Per-run statistics: HIP-event kernel time and algorithmic event counts (see
 * mcpet_run_params.count_events for which counts a run fills). */
typedef struct mcpet_stats {
  double kernel_ms;              /* transport kernel(s), HIP events on the context stream */
  double zero_ms;                /* tally clearing                                       */
  uint64_t decays;
  uint64_t streams;
  uint64_t woodcock_steps;
  uint64_t material_changes;     /* (a,b) table reloads                                  */
  uint64_t compton;
  uint64_t compton_shell_evals;  /* sum over Compton events of n_shell * n_trials        */
  uint64_t rayleigh;
  uint64_t photoelectric;
  uint64_t pairs_tallied;        /* coincidences reaching tally_Sinogram/PSF             */
  uint64_t coincidences;         /* trues + scatter binned                               */
  uint64_t rn_draws;
  /* persistent kernel, diagnostic builds only: executions, active lanes and shader clock
     cycles (summed over waves) per phase: DECAY, SOURCE2, STEP, INTERACT, TALLY, CTAU (one
     trial of GCOa's tau loop), CPZ (one trial of its p_z loop), DETECT (singles list),
     SOURCE3 (prompt gamma); zero otherwise */
  uint64_t phase_exec[MCPET_NPHASES];
  uint64_t phase_lanes[MCPET_NPHASES];
  uint64_t phase_cycles[MCPET_NPHASES];
  /* singles pipeline: time of the sort and pairing (HIP events), singles detected; in this
     mode `coincidences` counts the binned true and scattered coincidences of the pairing */
  double coinc_ms;
  uint64_t singles;
  uint64_t pairs_emitted;        /* decays that emitted an annihilation pair              */
  uint64_t prompts_emitted;      /* decays that emitted a prompt gamma                    */
} mcpet_stats;

/* ------------------------------------------------------------------ host-side input (boundary) */

/* Parse a MCGPU-PET.in file. Replaces read_input (MCGPU-PET.cu:866-1280), fgets_trimmed
 * (C:1320-1350), trim_name (C:1292-1311) and the sinogram-size formula (C:1176-1189).
 * The seed 0 -> time-based rule is applied. err (may be NULL) receives a message. */
mcpet_status mcpet_read_input(const char* path, mcpet_config* cfg, char* err, size_t errlen);

/* Load a .vox[.gz] phantom. Replaces load_voxels (MCGPU-PET.cu:1381-1533): applies
 * fact_object, clamps the ROI maxima of cfg (may be NULL) to the phantom size. */
mcpet_status mcpet_load_voxels(const char* path, mcpet_config* cfg, mcpet_phantom* out,
                               char* err, size_t errlen);

/* Build a phantom from caller arrays (material id 1..15, density g/cm3, activity Bq), x fastest.
 * Equivalent to writing and re-reading a .vox file. */
mcpet_status mcpet_phantom_from_arrays(int32_t nx, int32_t ny, int32_t nz, const float voxel_size[3],
                                       const int32_t* material, const float* density,
                                       const float* activity, float fact_object, mcpet_phantom* out,
                                       char* err, size_t errlen);
void mcpet_free_phantom(mcpet_phantom* ph);

/* Read material files and build the interaction tables. Replaces load_material
 * (MCGPU-PET.cu:1565-1831). Slots with an empty name are skipped; slots absent from
 * the phantom (density_max <= 0) are skipped except slot 0 (C:1608-1621). */
mcpet_status mcpet_load_materials(const char names[MCPET_MAX_MATERIALS][MCPET_NAME_LEN],
                                  const float density_max[MCPET_MAX_MATERIALS], mcpet_tables* out,
                                  char* err, size_t errlen);
/* The same tables with the energy grid continued past the files' last energy up to e_max_ev
 * (no reference counterpart: the reference reads past its tables there, K:302,359, and only
 * checks 511 keV, C:344-351). The grid keeps the files' first energy and step; every added
 * point is built from the files' last point (E_L) as raw mean free paths, then through the same
 * float arithmetic as the file points:
 *   Compton        mfp(E) = mfp(E_L) sigma_KN(E_L) / sigma_KN(E)   (Klein-Nishina total cross
 *                  section per electron; the files' Compton data follow its scaling within 2e-4
 *                  between 450 and 515 keV)
 *   Rayleigh, photoelectric: mfp(E) = mfp(E_L) (E / E_L)^s, s the log-log slope between E_L
 *                  and the file point nearest E_L / 2 (about 1.99 and 2.9)
 *   total          1 / (1/Rayleigh + 1/Compton + 1/photoelectric)  (as in the files below 1.022 MeV)
 *   Rayleigh pmax  the last file value (the form-factor cumulative has saturated, 1 - 3e-8)
 * e_max_ev <= the last file energy: the files' grid unchanged. At most
 * MCPET_MAX_ENERGYBINS_RAYLEIGH points in all. */
mcpet_status mcpet_load_materials_extended(const char names[MCPET_MAX_MATERIALS][MCPET_NAME_LEN],
                                           const float density_max[MCPET_MAX_MATERIALS], double e_max_ev,
                                           mcpet_tables* out, char* err, size_t errlen);
void mcpet_free_tables(mcpet_tables* t);

/* Sinogram sizes from (MRD, SPAN, NZS, NRAD, NANGLES): NSEG, NSINOS, NBINS (C:1176-1189). */
void mcpet_sinogram_sizes(mcpet_config* cfg);

/* ------------------------------------------------------------------ device path */

typedef struct mcpet_ctx mcpet_ctx;

/* Select the device, upload phantom and tables, allocate tallies. Replaces
 * init_CUDA_device (MCGPU-PET.cu:1842-2146) and the zeroing at C:408-413.
 * Applies the "negative radius -> centre the cylinder on the bbox" rule (C:316-322)
 * and the 511 keV table-range check (C:344-351). */
mcpet_status mcpet_create(const mcpet_config* cfg, const mcpet_phantom* ph, const mcpet_tables* tab,
                          mcpet_ctx** out);

/* Launch the transport (asynchronous on the context stream). Replaces
 * track_particles<<<blocks,threads>>> (MCGPU-PET.cu:452-487; kernel K:153-463).
 * Waits for the context's previous run to finish first (its tallies are overwritten);
 * in the singles pipeline it also waits for the transport, to size the sort. */
mcpet_status mcpet_run(mcpet_ctx* ctx, const mcpet_run_params* params);

/* Wait for the context stream; reports kernel faults (C:486-487). */
mcpet_status mcpet_sync(mcpet_ctx* ctx);

/* Copy tallies to caller buffers (NULL pointers are skipped). Replaces the D2H copies
 * at MCGPU-PET.cu:503-535,646,652. Implies mcpet_sync. */
mcpet_status mcpet_fetch(mcpet_ctx* ctx, mcpet_tallies* out);

/* Statistics of the last run (implies mcpet_sync). */
mcpet_status mcpet_get_stats(mcpet_ctx* ctx, mcpet_stats* out);

/* Free device memory and the stream (MCGPU-PET.cu:654-669). */
void mcpet_destroy(mcpet_ctx* ctx);

/* Message of the last failing call on ctx (or of the last context-free call if ctx is NULL). */
const char* mcpet_last_error(const mcpet_ctx* ctx);

/* Effective detector after mcpet_create (centre/radius rule applied). */
mcpet_status mcpet_get_detector(const mcpet_ctx* ctx, mcpet_detector* out);

/* LDS layout of the persistent transport kernel (mcpet_tune_result.lds_layout bits): what each
 * workgroup keeps in LDS rather than reading from global memory. No placement changes a tally. */
#define MCPET_LDS_PALETTE 1      /* the u8 voxel palette (phantoms with <= 256 distinct voxel words) */
#define MCPET_LDS_SPECTRUM 2     /* the energy spectrum, privatised (else global atomics)            */
#define MCPET_LDS_XS511 4        /* the 511 keV total-cross-section row of every material            */
#define MCPET_LDS_SET 8          /* mcpet_run_params.lds_layout holds an explicit layout             */

typedef struct mcpet_tune_result {
  int32_t block_size, blocks_per_cu;
  int32_t phase_policy;          /* 1 most lanes, 2 least waste (mcpet_run_params)        */
  double ms;                     /* kernel time of the sample (HIP events)                */
  double decays_per_ms;
  int32_t lds_layout;            /* MCPET_LDS_* bits                                      */
  int64_t stream_decays;         /* decays per stream of the layout (0: threads_per_block
                                    streams per emitting voxel, the reference's layout)   */
} mcpet_tune_result;

/* Autotuner of one context. No reference counterpart: the reference fixes its one launch from
 * threads_per_block (MCGPU-PET.in line 14, MCGPU-PET.cu:452-487). Each candidate runs a sample
 * of the context's decay streams (every k-th stream over the whole acquisition, k =
 * round(1 / stream_fraction)) and is timed with HIP events; the fastest is kept for later
 * mcpet_run calls whose block_size, blocks_per_cu and phase_policy are 0. Searched in turn:
 *   1. launch shape (workgroup size x workgroups per CU) x phase-selection policy;
 *   2. tune_layout: the LDS layout (MCPET_LDS_* placements moved to global memory one at a
 *      time, kept when faster); tallies are identical in every layout;
 *   3. tune_streams: the decays per stream (the reference's histories per thread), among
 *      stream_decays[0..n_stream_decays) (0 = threads_per_block streams per emitting voxel).
 *      Another layout draws other random-number streams (statistically equal tallies, not the
 *      same bits), so it is opt-in; the chosen value is applied to the context and reported,
 *      and a context created with mcpet_config.stream_decays = that value reproduces its runs.
 * results receives every timed candidate, then the choice (up to max_results entries);
 * n_results the number written. The context's tallies afterwards are not a run's. */
typedef struct mcpet_tune_options {
  double stream_fraction;        /* (0, 1]: share of the streams each candidate simulates */
  int32_t tune_layout;           /* 1: search the LDS layout                              */
  int32_t tune_streams;          /* 1: search the decays per stream (opt-in, see above)   */
  int32_t n_stream_decays;       /* 0..8 candidates in stream_decays                     */
  int64_t stream_decays[8];
} mcpet_tune_options;
mcpet_status mcpet_autotune(mcpet_ctx* ctx, const mcpet_tune_options* opt, mcpet_tune_result* results, int32_t max_results,
                            int32_t* n_results);

/* ------------------------------------------------------------------ batches of phantoms
 * No reference counterpart (the reference runs one phantom per process). A batch simulates
 * n phantoms that share the scanner, tally settings and interaction tables (cfg, tab), each
 * with its own voxels, activity and tallies, in ONE persistent launch with a stream queue
 * per phantom. Phantom k's tallies equal those of mcpet_create(cfg, &phantoms[k], tab) +
 * mcpet_run with the same run parameters, bit for bit. The pair tallies (modes 0-2) only. */
typedef struct mcpet_batch mcpet_batch;


mcpet_status mcpet_batch_create(const mcpet_config* cfg, const mcpet_phantom* phantoms, int32_t n, const mcpet_tables* tab,
                                mcpet_batch** out);
/* kernel is ignored (always the persistent batch kernel); block_size/blocks_per_cu 0 take the
 * autotuned values if mcpet_batch_autotune ran, else the defaults; use_graph replays the
 * clearing and the launch as a captured HIP graph (rebuilt when the launch shape changes). */
mcpet_status mcpet_batch_run(mcpet_batch* b, const mcpet_run_params* params);
mcpet_status mcpet_batch_sync(mcpet_batch* b);
mcpet_status mcpet_batch_fetch(mcpet_batch* b, int32_t k, mcpet_tallies* out);
/* Counts of phantom k; kernel_ms is the whole batch launch. */
mcpet_status mcpet_batch_get_stats(mcpet_batch* b, int32_t k, mcpet_stats* out);
/* Time the batch on the first `fraction` of every phantom's decay streams (full acquisition
 * per stream) for each launch shape (block size x workgroups per CU), keep the fastest for
 * later runs; writes up to max_results timings. */
mcpet_status mcpet_batch_autotune(mcpet_batch* b, double acquisition_fraction, mcpet_tune_result* results,
                                  int32_t max_results, int32_t* n_results);
void mcpet_batch_destroy(mcpet_batch* b);
const char* mcpet_batch_last_error(const mcpet_batch* b);

/* ------------------------------------------------------------------ serving loop
 * No reference counterpart. A server owns the interaction tables and runs simulation
 * requests (a config + phantom each, pair tallies) with continuous batching: every
 * mcpet_server_step admits queued requests in submission order while the device memory of
 * the admitted ones stays within the HBM budget, then advances every admitted request by
 * up to `streams_per_step` of its decay streams in ONE batched launch; requests whose
 * streams are all done are copied to host memory and release their device memory before the
 * next step. Streams are independent with fixed ids, so a request's tallies equal its
 * standalone mcpet_run bit for bit, whatever else it was batched with. */
typedef struct mcpet_server mcpet_server;
typedef struct mcpet_surrogate mcpet_surrogate;   /* the scatter network (below) */

typedef enum mcpet_request_state {
  MCPET_REQ_QUEUED = 0, MCPET_REQ_RUNNING = 1, MCPET_REQ_DONE = 2, MCPET_REQ_REJECTED = 3, MCPET_REQ_UNKNOWN = 4,
  MCPET_REQ_FAILED = 5     /* admitted, but its context could not be created or its results not computed */
} mcpet_request_state;

typedef struct mcpet_server_stats {
  int64_t submitted, completed, rejected, steps;
  int32_t queued, running;
  size_t budget_bytes;
  size_t peak_admitted_bytes;    /* largest device footprint of the admitted requests        */
  int32_t peak_running;
  double latency_p50_ms, latency_p99_ms, latency_max_ms;   /* submit -> done, completed requests */
  double busy_ms;                /* time spent in steps                                       */
  int64_t failed;                /* requests that failed after admission                      */
} mcpet_server_stats;

mcpet_status mcpet_server_create(const mcpet_tables* tab, size_t hbm_budget_bytes, int32_t max_running,
                                 int64_t streams_per_step, mcpet_server** out);
/* Queue a request (cfg and phantom are copied; seed 0: the config's). A request that
 * mcpet_create would refuse (mcpet_create's argument and range checks, every voxel's material
 * slot) or whose seed cannot be used is rejected with that status, and one whose device
 * footprint alone exceeds the budget with MCPET_E_CAPACITY; a rejected request never enters the
 * queue. A request that fails after admission (its context cannot be created, its results
 * cannot be copied or computed) fails alone: its state becomes MCPET_REQ_FAILED, its device
 * memory is released and the requests behind it proceed. Fetching a rejected or failed request
 * returns its error status and message and forgets it. */
mcpet_status mcpet_server_submit(mcpet_server* s, const mcpet_config* cfg, const mcpet_phantom* ph, int32_t seed,
                                 int64_t* request_id);
/* One scheduling step (admission, one batched launch, completions); blocks until done. */
mcpet_status mcpet_server_step(mcpet_server* s, int32_t* running, int32_t* completed);
/* Step until no request is queued or running. */
mcpet_status mcpet_server_drain(mcpet_server* s);
mcpet_status mcpet_server_poll(mcpet_server* s, int64_t request_id, int32_t* state);
/* Results of a finished request (the sizes of its config / phantom); the request is then forgotten. */
mcpet_status mcpet_server_fetch(mcpet_server* s, int64_t request_id, mcpet_tallies* out);
/* A request whose result is only its trues and scatter rebinned on the device onto a D x H x W
 * grid (mcpet_rebin_sinogram) and its number of decays: no full tallies are copied to the host
 * (the scatter surrogate's inputs and its Monte Carlo fallbacks). with_maps: also the phantom's
 * line integrals on that grid (mcpet_project_maps; D <= 256). */
mcpet_status mcpet_server_submit_rebinned(mcpet_server* s, const mcpet_config* cfg, const mcpet_phantom* ph, int32_t seed,
                                          int32_t D, int32_t H, int32_t W, int32_t with_maps, int64_t* request_id);
/* trues, scatter, act, att: D*H*W floats each (may be NULL; act and att need with_maps);
 * decays (may be NULL). The request is then forgotten. */
mcpet_status mcpet_server_fetch_rebinned(mcpet_server* s, int64_t request_id, float* trues, float* scatter, float* act,
                                         float* att, int64_t* decays);
/* A request answered on the device by a scatter network (mcpet_surrogate_estimate on the
 * request's context when its run completes: rebinning, maps, normalisation, network, scaling).
 * The network must outlive the request; the server's requests use it one at a time. */
mcpet_status mcpet_server_submit_estimate(mcpet_server* s, const mcpet_config* cfg, const mcpet_phantom* ph, int32_t seed,
                                          mcpet_surrogate* net, int64_t* request_id);
/* estimate, scatter_counts: the network's grid (may be NULL); decays (may be NULL). The request
 * is then forgotten. */
mcpet_status mcpet_server_fetch_estimate(mcpet_server* s, int64_t request_id, float* estimate, float* scatter_counts,
                                         int64_t* decays);
mcpet_status mcpet_server_get_stats(mcpet_server* s, mcpet_server_stats* out);
void mcpet_server_destroy(mcpet_server* s);
const char* mcpet_server_last_error(const mcpet_server* s);

/* ------------------------------------------------------------------ scatter surrogate (no reference counterpart)
 *
 * The rebinned scatter grid: sinogram izm of segment iseg holds the ring pairs with ring sum
 * s = iz1 + iz2 = izm - ofseg(iseg) + ofz(incl) (the michelogram offsets of K:1749-1788), and
 * single-slice rebinning puts it at plane s (0 .. 2 NROWS - 2). A D x H x W grid sums planes in
 * groups of ceil(NZS / D), angles in groups of ceil(NANGLES / H) and radial bins in groups of
 * ceil(NRAD / W). */

/* plane_of_sinogram[izm] = s for every sinogram of cfg's michelogram (cfg->nsinos entries). */
mcpet_status mcpet_sinogram_planes(const mcpet_config* cfg, int32_t* plane_of_sinogram);

/* Rebin tally `which` (0 trues, 1 scatter, 2 randoms, 3 spurious) of the last run on the device
 * into the D x H x W grid (out: host, D*H*W floats of counts). Waits for the run. */
mcpet_status mcpet_rebin_sinogram(mcpet_ctx* c, int32_t which, int32_t D, int32_t H, int32_t W, float* out);
/* From the next run on, every detected trues or scatter coincidence is also added to its bin
 * of the D x H x W rebinning grid (the grid of mcpet_rebin_sinogram), so that the fused
 * estimate (mcpet_surrogate_estimate) on that grid reads these counts instead of rebinning the
 * sinograms after the run: the same integer sums. The sinograms are tallied as before. The
 * estimate reads the counts only while every run since the last clear of the tallies added to
 * them (a run with zero_tallies = 0 after one with the tally off: it rebins the sinograms).
 * D = H = W = 0 turns it off. Not in the singles mode (MCPET_E_STATE). No reference
 * counterpart: an extension for the scatter surrogate's input runs. */
mcpet_status mcpet_tally_rebinned(mcpet_ctx* c, int32_t D, int32_t H, int32_t W);
mcpet_status mcpet_batch_rebin_sinogram(mcpet_batch* b, int32_t k, int32_t which, int32_t D, int32_t H, int32_t W, float* out);

/* The phantom seen along the lines of response of the D x H x W grid (1 <= D <= 256): for every
 * grid bin, the mean over the crystal pairs of its angle and radial groups (the pairs of the
 * coincidence binning, K:1776-1786) of the line between the two crystal centres (angle
 * 2 pi (k + 1/2) / NCRYSTALS - pi on the detector radius) in the axial plane of the bin's
 * central ring sum, sampled every half transaxial voxel inside the phantom:
 *   act: integral of the activity (Bq cm); att: exp(-integral of mu at 511 keV) (1 outside).
 * Computed on the device at the first call for a grid and kept by the context (act and att may
 * be NULL: compute and keep only); no run needed. */
mcpet_status mcpet_project_maps(mcpet_ctx* c, int32_t D, int32_t H, int32_t W, float* act, float* att);
mcpet_status mcpet_batch_project_maps(mcpet_batch* b, int32_t k, int32_t D, int32_t H, int32_t W, float* act, float* att);

/* A 3D U-Net on the rebinned grid (D, H, W multiples of 4): in_channels -> width (two 3^3
 * convolutions) -> stride-2 convolution to 2 width -> convolution -> stride-2 to 4 width ->
 * convolution -> nearest upsampling, concatenation with the 2-width skip -> convolution to
 * 2 width -> upsampling, concatenation with the width skip -> convolution to width -> 1^3
 * convolution; ReLU after every 3^3 convolution; the output is input channel 0 plus that
 * residual. Every convolution runs as an implicit GEMM on BF16 MFMA tiles with FP32
 * accumulation; parameters, activations and optimiser state are FP32. Parameter layout: per
 * layer W[Cout][Cin][k^3] then b[Cout], layers in the order above. Created with He-normal
 * weights from `seed`, zero biases and a zero output layer (the identity on channel 0). */
mcpet_status mcpet_surrogate_create(int32_t D, int32_t H, int32_t W, int32_t in_channels, int32_t width, uint64_t seed,
                                    mcpet_surrogate** out);
int64_t mcpet_surrogate_param_count(const mcpet_surrogate* s);
mcpet_status mcpet_surrogate_get_params(mcpet_surrogate* s, float* out);
mcpet_status mcpet_surrogate_set_params(mcpet_surrogate* s, const float* in);
/* x: batch x in_channels x D x H x W (host); y: batch x D x H x W (host). gflop: forward GFLOP per sample. */
mcpet_status mcpet_surrogate_forward(mcpet_surrogate* s, const float* x, int32_t batch, float* y, double* gflop);
/* Device time (ms, HIP events on the network's stream) of the network proper in the last
 * mcpet_surrogate_forward call (its last sample; no host copies). */
mcpet_status mcpet_surrogate_forward_ms(mcpet_surrogate* s, double* ms);
/* One estimate from the last run of ctx (a 2- or 4-input-channel network; same device), all on
 * the device: its scatter and trues rebinned onto the network's grid (mcpet_rebin_sinogram),
 * normalised (counts times float(voxels / total trues), the total summed exactly in double),
 * with 4 channels also the activity integrals times float(1 / their maximum) and the
 * attenuation factors (mcpet_project_maps), the network, and estimate = max(output, 0) times
 * float(total trues / (voxels * decays)): scatter per decay (D*H*W floats, host).
 * scatter_counts (may be NULL): the rebinned scatter; decays (may be NULL): the run's decays.
 * Call it right after mcpet_run: its device work is ordered behind the run on the device, and
 * the host waits once, for the outputs. */
mcpet_status mcpet_surrogate_estimate(mcpet_surrogate* s, mcpet_ctx* ctx, float* estimate, float* scatter_counts,
                                      int64_t* decays);
/* loss = mean over batch and voxels of (y - target)^2; grad (host, param_count floats; may be NULL). */
mcpet_status mcpet_surrogate_gradients(mcpet_surrogate* s, const float* x, const float* target, int32_t batch, float* loss,
                                       float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss; loss is the value before the step. */
mcpet_status mcpet_surrogate_train_step(mcpet_surrogate* s, const float* x, const float* target, int32_t batch, float lr,
                                        float* loss);
void mcpet_surrogate_destroy(mcpet_surrogate* s);
const char* mcpet_surrogate_last_error(const mcpet_surrogate* s);

/* ------------------------------------------------------------------ host-side output (boundary) */

/* Write the reference's output files into directory dir (NULL = cwd):
 * sinogram_Trues/Scatter.raw.gz and image_Trues/Scatter.raw.gz (C:544-562; BG_ names in
 * BG mode, C:580-594), Energy_Sinogram_Spectrum.dat (C:612-616), the PSF text and .raw
 * (report_PSF, C:3193-3322) and the voxel dose files (report_voxels_dose, C:2389-2650).
 * Which files are written follows tally_psf_sinogram and the dose flags, as in main. */
mcpet_status mcpet_write_outputs(const char* dir, const mcpet_config* cfg, const mcpet_phantom* ph,
                                 const mcpet_tables* tab, const mcpet_tallies* t,
                                 double time_elapsed_s, char* err, size_t errlen);

/* The dose summary of report_voxels_dose (MCGPU-PET.cu:2476-2641) from the voxel dose tally of
 * the ROI (t->voxel_edep: ROI voxels x fastest, (sum E, sum E^2) per voxel in eV): the energy
 * deposited inside the ROI with voxels below 0.1 g/cm3 left out (C:2523-2527), the peak voxel
 * dose over all materials, and per material slot the ROI mass, voxel count, energy (all ROI
 * voxels of the slot, as C:2512-2515 sums them before the low-density exclusion), dose and
 * peak voxel. Doses in eV/g per history, standard deviations as the reference computes them. */
typedef struct mcpet_dose_report {
  double total_energy_kev;                  /* inside the ROI, keV per history            */
  double max_dose, max_dose_std_dev;        /* peak voxel dose of all materials           */
  double mass[MCPET_MAX_MATERIALS];         /* g, ROI voxels of the slot                  */
  double edep[MCPET_MAX_MATERIALS];         /* eV per history                             */
  double dose[MCPET_MAX_MATERIALS], dose_std_dev[MCPET_MAX_MATERIALS];
  double peak_dose[MCPET_MAX_MATERIALS], peak_std_dev[MCPET_MAX_MATERIALS];
  int32_t peak_voxel[MCPET_MAX_MATERIALS][3];   /* 0-based voxel coordinates in the phantom  */
  uint32_t voxels[MCPET_MAX_MATERIALS];     /* ROI voxels of the slot (0: not reported)    */
} mcpet_dose_report;

mcpet_status mcpet_voxel_dose_report(const mcpet_config* cfg, const mcpet_phantom* ph, const uint64_t* voxel_edep,
                                     uint64_t total_histories, mcpet_dose_report* out);

/* RANECU seeds (x, y) at the start of stream ids[i] for input seed `seed`, computed on the
 * host with the same jump-table skip-ahead as the kernels; equal to init_PRNG(ids[i], 555555,
 * seed) (K:924-978). Lets a caller resume or reproduce any decay stream. */
mcpet_status mcpet_stream_seeds(int32_t seed, const int64_t* ids, int32_t n, int32_t* sx, int32_t* sy);

/* ------------------------------------------------------------------ self-test entry points */

/* Draw n RANECU floats for each of nstreams streams on the device, stream i initialised as
 * init_PRNG(stream_ids[i], 555555, seed) (K:924-978, K:1044-1061); out is [nstreams][n].
 * If as_double, draws ranecu_double (K:1067-1083) into out_d instead. */
mcpet_status mcpet_selftest_ranecu(int32_t seed, const int64_t* stream_ids, int32_t nstreams,
                                   int32_t n, float* out, double* out_d, int32_t as_double);

/* Evaluate the device elementary functions (log, exp, sincos, atan2, acos, logf, expf,
 * powf, sincosf, atan2f, acosf) on n inputs; used to check host/device bit equality. */
mcpet_status mcpet_selftest_fmath(const double* x, const double* y, int32_t n, double* out_d /* 7*n */,
                                  float* out_f /* 8*n */);

/* Stable device radix sort of n 64-bit keys by their low 8*passes bits (passes 1..8), the
 * sort of the singles pipeline; perm receives the sorted order of the input indices. */
mcpet_status mcpet_selftest_sort(const uint64_t* keys, int64_t n, int32_t passes, uint32_t* perm);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same sort as
 * mcpet_selftest_sort, same arguments and results, calling no hipMalloc/hipFree. Its device
 * buffers come from one static device array of about 24.3 MiB that each device the library runs
 * on holds until the process ends; n is limited to 2^20 keys (MCPET_E_CAPACITY beyond). Calls
 * that overlap on one device share that array and race. */
mcpet_status mcpet_selftest_sort_noalloc(const uint64_t* keys, int64_t n, int32_t passes, uint32_t* perm);

/* C = A (MxK) x B (KxN), row-major FP32 in and out, computed by the surrogate's BF16 MFMA GEMM
 * kernel (operands rounded to BF16, FP32 accumulation; split > 1: K split over workgroups). */
mcpet_status mcpet_selftest_gemm_bf16(const float* a, const float* b, int32_t M, int32_t N, int32_t K, int32_t split, float* c);

/* Device copy of nbytes (for HBM-counter calibration: FETCH_SIZE/WRITE_SIZE of a known copy). */
mcpet_status mcpet_selftest_copy(size_t nbytes, int32_t iters, double* ms_out);

/* Library version string. */
const char* mcpet_version(void);

/* sizeof of the ABI structs, in this order: mcpet_config, mcpet_phantom, mcpet_tables,
 * mcpet_tallies, mcpet_run_params, mcpet_stats, mcpet_psf_record, mcpet_single (lets a
 * binding check its struct layouts). Returns the number of sizes written (<= n). */
int32_t mcpet_struct_sizes(size_t* out, int32_t n);

#ifdef __cplusplus
}
#endif

#endif /* MCPET_H */
