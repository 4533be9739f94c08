// runtime.cpp — C ABI of the device path (mcpet_create / run / sync / fetch / destroy).
//
// Replaces the reference's init_CUDA_device (MCGPU-PET.cu:1842-2146), the single launch
// of track_particles (C:452-487) and the result copies (C:499-535, 642-652). The
// library owns all device memory; the caller's host buffers are only read or written
// during the calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_internal.h"
#include "runtime_internal.h"
#include "kparams.h"
#include "mcpet.h"

struct mcpet_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr, ev_z0 = nullptr;
  hipEvent_t ev_done = nullptr;   // recorded behind every launch that runs this context (any stream)
  mcpet_config cfg{};
  mcpet_detector det{};
  int nx = 0, ny = 0, nz = 0;
  long long nvox = 0;
  long long n_active = 0;
  long long n_roi = 0;
  KParams P{};
  // device allocations
  float2* d_vox = nullptr;
  uint8_t* d_vox8 = nullptr;
  uint16_t* d_vox16 = nullptr;
  float2* d_palette = nullptr;
  int* d_active_vox = nullptr;
  float* d_active_act = nullptr;
  long long* d_stream_start = nullptr;   // balanced stream layout (stream_decays > 0)
  long long* d_stream_id = nullptr;
  int* d_stream_hint = nullptr;
  int* d_rbt = nullptr;                  // rebinned trues and scatter (mcpet_tally_rebinned), 2 x D*H*W
  int* d_rbt_plane = nullptr;            // their plane group of every sinogram
  int rbt_dims[3] = {0, 0, 0};
  bool rbt_ran = false;                  // the last run added to them
  size_t hint_cap = 0;                   // entries allocated for d_stream_hint
  float2* d_woodcock = nullptr;
  float2* d_tot = nullptr;
  float4* d_cr = nullptr;
  mcpet_rayleigh_table* d_ray = nullptr;
  mcpet_compton_table* d_cmp = nullptr;
  int* d_trues = nullptr;
  int* d_scatter = nullptr;
  int* d_img_t = nullptr;
  int* d_img_sc = nullptr;
  int* d_spectrum = nullptr;
  mcpet_psf_record* d_psf = nullptr;
  int* d_psf_index = nullptr;
  unsigned long long* d_voxel_edep = nullptr;
  unsigned long long* d_mat_edep = nullptr;
  unsigned long long* d_counters = nullptr;  // KC_N counters + stream queue
  int* d_final_seed = nullptr;
  size_t dev_bytes = 0;          // device memory held by the context
  KParams* d_params = nullptr;   // the run's parameter block (device copy)
  KParams* h_params = nullptr;   // pinned staging copy
  // singles pipeline
  bool singles_mode = false;
  mcpet_single* d_singles = nullptr;
  unsigned long long* d_singles_n = nullptr;
  unsigned long long* d_keys[2] = {nullptr, nullptr};
  unsigned* d_vals[2] = {nullptr, nullptr};
  unsigned* d_hist = nullptr;
  std::vector<unsigned*> d_scan;
  int* d_randoms = nullptr;
  int* d_spurious = nullptr;
  unsigned long long* d_coinc = nullptr;
  hipEvent_t ev_c1 = nullptr;
  unsigned long long singles_req = 0;
  long long n_singles = 0;
  int sorted = 0;
  bool ran = false;
  bool timed = false;
  double zero_ms = 0.0;
  // rebinning onto the surrogate's grid (allocated at the first rebin)
  int* d_rb_off = nullptr;
  int* d_rb_sino = nullptr;
  float* d_rb_out = nullptr;
  int rb_dims[3] = {0, 0, 0};
  // line integrals on the surrogate's grid (computed at the first request, kept)
  KParams* d_pj_params = nullptr;
  float* d_pj_out = nullptr;      // [3][D][H][W]: activity integral, the same / its maximum, attenuation factor; + 1 word
  int pj_dims[3] = {0, 0, 0};
  size_t pj_alloc = 0;            // grid size of d_pj_out's allocation
  // mcpet_autotune's choices for runs that leave them 0 (0 / -1: none)
  int tuned_bs = 0, tuned_per_cu = 0, tuned_policy = 0, tuned_lds = -1;
  std::string err;
};

namespace {

thread_local std::string g_err;

mcpet_status set_err(mcpet_ctx* c, mcpet_status st, const std::string& msg)
{
  if (c) c->err = msg;
  g_err = msg;
  mcpet_host_set_error(msg.c_str());
  return st;
}

// Seeds for which the reference's abMODm (K:1000-1032) divides by zero: 0 (replaced by a
// time-based seed when read from a file) and values below -m2.
static bool seed_usable(int s) { return s != 0 && s >= -MCR_M2; }
static std::string seed_msg(int s)
{
  return "seed " + std::to_string(s) + " is not a usable RANECU seed (must be non-zero and >= -" +
         std::to_string(MCR_M2) + ")";
}

#define HIPCHK(ctx, call)                                                                              \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      return set_err(ctx, MCPET_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

template <class T>
mcpet_status dalloc(mcpet_ctx* c, T** p, size_t n)
{
  *p = nullptr;
  if (n == 0) return MCPET_OK;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(c, MCPET_E_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
  }
  c->dev_bytes += n * sizeof(T);
  return MCPET_OK;
}

void free_all(mcpet_ctx* c)
{
  void* ptrs[] = {c->d_vox, c->d_vox8, c->d_vox16, c->d_palette, c->d_active_vox, c->d_active_act, c->d_stream_start, c->d_stream_id, c->d_stream_hint, c->d_rbt, c->d_rbt_plane, c->d_woodcock, c->d_tot, c->d_cr, c->d_ray, c->d_cmp,
                  c->d_trues, c->d_scatter, c->d_img_t, c->d_img_sc, c->d_spectrum, c->d_psf, c->d_psf_index,
                  c->d_voxel_edep, c->d_mat_edep, c->d_counters, c->d_final_seed, c->d_singles, c->d_singles_n,
                  c->d_keys[0], c->d_keys[1], c->d_vals[0], c->d_vals[1], c->d_hist, c->d_randoms, c->d_spurious,
                  c->d_coinc, c->d_rb_off, c->d_rb_sino, c->d_rb_out, c->d_pj_params, c->d_pj_out};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (unsigned* p : c->d_scan)
    if (p) (void)hipFree(p);
  if (c->ev_c1) (void)hipEventDestroy(c->ev_c1);
  if (c->d_params) (void)hipFree(c->d_params);
  if (c->h_params) (void)hipHostFree(c->h_params);
  if (c->ev_k0) (void)hipEventDestroy(c->ev_k0);
  if (c->ev_k1) (void)hipEventDestroy(c->ev_k1);
  if (c->ev_done) (void)hipEventDestroy(c->ev_done);
  if (c->ev_z0) (void)hipEventDestroy(c->ev_z0);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

}  // namespace

bool mcpet_seed_usable(int seed) { return seed_usable(seed); }

extern "C" const char* mcpet_version(void) { return "mcpet-mi355x 0.2 (gfx950)"; }

extern "C" int32_t mcpet_struct_sizes(size_t* out, int32_t n)
{
  const size_t sz[] = {sizeof(mcpet_config),      sizeof(mcpet_phantom), sizeof(mcpet_tables),
                       sizeof(mcpet_tallies),     sizeof(mcpet_run_params), sizeof(mcpet_stats),
                       sizeof(mcpet_psf_record),  sizeof(mcpet_single), sizeof(mcpet_server_stats)};
  int32_t k = 0;
  for (; out && k < n && k < (int32_t)(sizeof(sz) / sizeof(sz[0])); k++) out[k] = sz[k];
  return k;
}

extern "C" const char* mcpet_last_error(const mcpet_ctx* ctx)
{
  if (ctx) return ctx->err.c_str();
  const char* h = mcpet_host_last_error();
  if (h && h[0]) return h;
  return g_err.c_str();
}

// The argument and range checks of mcpet_create that need no device and no allocation (the
// serving loop runs them when a request is submitted, so that a request that cannot be created
// is rejected then and never blocks the queue).
mcpet_status mcpet_check_create_args(const mcpet_config* cfg, const mcpet_phantom* ph, const mcpet_tables* tab,
                                     bool scan_voxels)
{
  if (!cfg || !ph || !tab || !ph->voxel_mat_dens || !tab->woodcock || !tab->mfp_a || !tab->mfp_b || !tab->rayleigh || !tab->compton)
    return set_err(nullptr, MCPET_E_ARG, "mcpet_create: null argument");
  if (cfg->threads_per_block < 1) return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: threads_per_block must be >= 1");
  if (cfg->nbins < 1 || cfg->ne < 0 || cfg->nrows < 1 || cfg->ncrystals < 1 || cfg->nangles < 1 || cfg->nrad < 1 || cfg->span < 1)
    return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: invalid sinogram parameters");
  if (tab->num_values < 2 || !(tab->ide > 0.0f))
    return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: invalid interaction tables");
  // C:344-351: the tables must reach the annihilation energy.
  if (510998.95f > (tab->e0 + (tab->num_values - 1) / tab->ide) || !(tab->e0 < 510998.95f))
    return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: the material files do not have data at 510.999 keV");

  {
    // The primary's energy bin must lie inside the tables (index+1 is read for Rayleigh).
    float fi = floorf((510998.95f - tab->e0) * tab->ide);
    if (!(fi >= 0.0f && fi <= (float)(tab->num_values - 2)))
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: 511 keV outside the interaction tables");
  }
  const int tmode = cfg->detector.tally_psf_sinogram;
  if (tmode < 0 || tmode > MCPET_TALLY_SINGLES)
    return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: tally_psf_sinogram must be 0, 1, 2 or 3 (singles)");
  if (!(cfg->prompt_gamma_fraction >= 0.0f && cfg->prompt_gamma_fraction <= 1.0f))
    return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: prompt_gamma_fraction must lie in [0, 1]");
  if (cfg->prompt_gamma_fraction > 0.0f) {
    // the prompt gamma's energy bin (and the next one, read by Rayleigh sampling) must exist
    const float fi = floorf((cfg->prompt_gamma_energy - tab->e0) * tab->ide);
    if (!(cfg->prompt_gamma_energy > 0.0f && fi >= 0.0f && fi <= (float)(tab->num_values - 2)))
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: the prompt gamma energy lies outside the interaction tables");
  }
  if (tmode == MCPET_TALLY_SINGLES) {
    if (cfg->singles_capacity < 1 || cfg->singles_capacity > 0x7fffffffLL)
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: singles_capacity must lie in [1, 2^31-1]");
    if (cfg->coincidence_window_ps == 0)
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: coincidence_window_ps must be > 0");
    if ((long long)ph->nx * ph->ny * ph->nz * (long long)cfg->threads_per_block >= (1LL << 32))
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: voxels x threads_per_block must be < 2^32 for singles event ids");
  }
  if (cfg->voxel_dose && cfg->roi_x_max > -1 &&
      (cfg->roi_x_min > cfg->roi_x_max || cfg->roi_y_min > cfg->roi_y_max || cfg->roi_z_min > cfg->roi_z_max ||
       cfg->roi_x_min < 0 || cfg->roi_y_min < 0 || cfg->roi_z_min < 0 || cfg->roi_x_max >= ph->nx ||
       cfg->roi_y_max >= ph->ny || cfg->roi_z_max >= ph->nz))
    return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: dose region of interest outside the phantom");
  if (scan_voxels) {
    const long long nvox = (long long)ph->nx * ph->ny * ph->nz;
    for (long long v = 0; v < nvox; v++) {
      const int m0 = (int)(ph->voxel_mat_dens[3 * v] - 1);
      if (m0 < 0 || m0 >= MCPET_MAX_MATERIALS) return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: voxel material out of range");
    }
  }
  return MCPET_OK;
}

// Balanced stream layout: voxel v gets max(1, ceil(E_v / stream_decays)) streams (E_v: its
// expected decays over the acquisition, teff = tau (1 - exp(-T / tau)) times its activity);
// random-number stream ids count the streams in voxel order, the device processes the voxels
// in its own (hottest first) order. act_dev: activities in device order; src_of[k]: the
// position of device entry k in voxel order. Fills stream_start (n + 1) and stream_id (n) in
// device order and returns the number of streams.
static long long balanced_layout(const std::vector<float>& act_dev, const std::vector<size_t>& src_of, double teff,
                                 long long stream_decays, std::vector<long long>& stream_start, std::vector<long long>& stream_id)
{
  const size_t n = act_dev.size();
  auto streams_of = [&](float act) {
    const double m = std::ceil((double)act * teff / (double)stream_decays);
    return m < 1.0 ? 1LL : (m > 1.0e9 ? 1000000000LL : (long long)m);
  };
  std::vector<long long> first_id(n);   // in voxel order
  for (size_t k = 0; k < n; k++) first_id[src_of[k]] = streams_of(act_dev[k]);
  long long id = 0;
  for (size_t j = 0; j < n; j++) {
    const long long m = first_id[j];
    first_id[j] = id;
    id += m;
  }
  stream_start.resize(n + 1);
  stream_id.resize(n);
  long long s0 = 0;
  for (size_t k = 0; k < n; k++) {
    stream_start[k] = s0;
    stream_id[k] = first_id[src_of[k]];
    s0 += streams_of(act_dev[k]);
  }
  stream_start[n] = s0;
  return id;
}

// Owner hints of a balanced layout (start: n + 1 stream offsets, strictly increasing): with
// g the smallest shift leaving at most n blocks of 2^g streams, hint[b] = the voxel owning
// stream b 2^g and hint[last + 1] = n - 1. The device's owner search then starts from the
// voxels between two consecutive hints instead of all n.
static int stream_hints(const std::vector<long long>& start, std::vector<int>& hint)
{
  const long long n = (long long)start.size() - 1, total = start.back();
  int g = 0;
  while ((total >> g) > n) g++;
  const long long nb = (total - 1) >> g;
  hint.resize((size_t)nb + 2);
  long long k = 0;
  for (long long b = 0; b <= nb; b++) {
    const long long s = b << g;
    while (start[(size_t)k + 1] <= s) k++;
    hint[(size_t)b] = (int)k;
  }
  hint[(size_t)nb + 1] = (int)(n - 1);
  return g;
}

static double effective_time(const mcpet_config* cfg)
{
  const double tau = (double)cfg->mean_life, T = (double)cfg->acquisition_time_ps * 1.0e-12;
  return tau * (1.0 - std::exp(-T / tau));
}

extern "C" mcpet_status mcpet_create(const mcpet_config* cfg, const mcpet_phantom* ph, const mcpet_tables* tab, mcpet_ctx** out)
{
  if (!out) return set_err(nullptr, MCPET_E_ARG, "mcpet_create: null output pointer");
  *out = nullptr;
  {
    const mcpet_status st0 = mcpet_check_create_args(cfg, ph, tab, false);   // voxels are checked while packing
    if (st0 != MCPET_OK) return st0;
  }
  mcpet_ctx* c = new mcpet_ctx();
  c->cfg = *cfg;
  c->nx = ph->nx;
  c->ny = ph->ny;
  c->nz = ph->nz;
  c->nvox = (long long)ph->nx * ph->ny * ph->nz;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    delete c;
    return set_err(nullptr, MCPET_E_HIP, "mcpet_create: no HIP device available");
  }
  c->device = (cfg->gpu_id >= 0 && cfg->gpu_id < ndev) ? cfg->gpu_id : 0;   // C:1940-1952 fallback
  mcpet_status st;
#define CK(x)                  \
  do {                         \
    st = (x);                  \
    if (st != MCPET_OK) {      \
      std::string m = c->err;  \
      free_all(c);             \
      delete c;                \
      set_err(nullptr, st, m); \
      return st;               \
    }                          \
  } while (0)
#define HK(x) CK(([&]() -> mcpet_status { HIPCHK(c, (x)); return MCPET_OK; })())

  HK(hipSetDevice(c->device));
  {
    // compute units per device, queried once per process (the property query is slow)
    static std::mutex cu_mu;
    static std::map<int, int> cus;
    std::lock_guard<std::mutex> lock(cu_mu);
    auto it = cus.find(c->device);
    if (it == cus.end()) {
      hipDeviceProp_t prop;
      HK(hipGetDeviceProperties(&prop, c->device));
      it = cus.emplace(c->device, prop.multiProcessorCount).first;
    }
    c->num_cus = it->second;
  }
  HK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HK(hipEventCreate(&c->ev_k0));
  HK(hipEventCreate(&c->ev_k1));
  HK(hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming));
  HK(hipEventCreate(&c->ev_z0));

  // Detector: negative radius -> centre on the bounding box (C:316-322).
  c->det = cfg->detector;
  if (c->det.radius < 0.0f) {
    c->det.center[0] = ph->size_bbox[0] * 0.5f;
    c->det.center[1] = ph->size_bbox[1] * 0.5f;
    c->det.center[2] = ph->size_bbox[2] * 0.5f;
    c->det.radius = -c->det.radius;
  }

  // Table id per material slot: slots whose total cross-section columns are bitwise
  // identical (the same material file) share the smallest such slot's id.
  int table_id[MCPET_MAX_MATERIALS];
  for (int m = 0; m < MCPET_MAX_MATERIALS; m++) {
    table_id[m] = m;
    for (int m2 = 0; m2 < m; m2++) {
      bool same = true;
      for (int b = 0; b < tab->num_values && same; b++) {
        const size_t i1 = 3 * ((size_t)b * MCPET_MAX_MATERIALS + m), i2 = 3 * ((size_t)b * MCPET_MAX_MATERIALS + m2);
        same = std::memcmp(&tab->mfp_a[i1], &tab->mfp_a[i2], 4) == 0 && std::memcmp(&tab->mfp_b[i1], &tab->mfp_b[i2], 4) == 0;
      }
      if (same) {
        table_id[m] = table_id[m2];
        break;
      }
    }
  }
  // Voxels -> packed (density, material word) pairs; the emitting voxels, in voxel order.
  std::vector<float2> hv((size_t)c->nvox);
  std::vector<int> act_idx;
  for (long long v = 0; v < c->nvox; v++) {
    const float* p = ph->voxel_mat_dens + 3 * v;
    int m0 = (int)(p[0] - 1);
    if (m0 < 0 || m0 >= MCPET_MAX_MATERIALS) {
      free_all(c);
      delete c;
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: voxel material out of range");
    }
    float bits;
    const int word = m0 | (table_id[m0] << 8);
    std::memcpy(&bits, &word, 4);
    hv[(size_t)v] = make_float2(p[1], bits);
    if (p[2] > 1.0e-7f) act_idx.push_back((int)v);   // K:232-237
  }
  // Palette of the distinct (density, material) words: a u8 or u16 index per voxel keeps
  // the volume small enough to stay in the L2 of each XCD (the Woodcock gather is random).
  std::vector<float2> palette;
  std::vector<uint16_t> pidx;
  if (cfg->voxel_format == 0) {
    std::unordered_map<uint64_t, int> seen;
    uint64_t last_key = 0;
    int last_idx = 0;
    pidx.resize((size_t)c->nvox);
    for (long long v = 0; v < c->nvox && palette.size() <= 65536; v++) {
      uint32_t d, m;
      std::memcpy(&d, &hv[(size_t)v].x, 4);
      std::memcpy(&m, &hv[(size_t)v].y, 4);
      const uint64_t key = ((uint64_t)m << 32) | d;
      if (v > 0 && key == last_key) {   // runs of equal voxels (air, tissue interiors) skip the map
        pidx[(size_t)v] = (uint16_t)last_idx;
        continue;
      }
      auto it = seen.find(key);
      if (it == seen.end()) {
        it = seen.emplace(key, (int)palette.size()).first;
        palette.push_back(hv[(size_t)v]);
      }
      last_key = key;
      last_idx = it->second;
      pidx[(size_t)v] = (uint16_t)it->second;
    }
    if (palette.size() > 65536) {
      palette.clear();
      pidx.clear();
    }
  }
  c->P.vox_fmt = palette.empty() ? 0 : (palette.size() <= 256 ? 1 : 2);
  c->P.n_palette = (int)palette.size();
  // Device order of the emitting voxels: hottest first (long streams start early), by a
  // counting sort on the top 12 bits of the activity's float representation (exponent and 3
  // mantissa bits: classes within a factor 1.125), stable within a class. The order only
  // balances the load: every stream draws the same random numbers in any order.
  std::vector<int> act_vox(act_idx);   // voxel order
  std::vector<size_t> src_of(act_idx.size());   // device position -> position in voxel order
  {
    std::vector<uint32_t> cls(act_idx.size());
    std::vector<size_t> count(4097, 0);
    for (size_t k = 0; k < act_idx.size(); k++) {
      uint32_t b;
      std::memcpy(&b, &ph->voxel_mat_dens[3 * (size_t)act_idx[k] + 2], 4);
      cls[k] = 4095u - (b >> 20);   // activity > 0: larger bits = hotter; class 0 = hottest
      count[cls[k] + 1]++;
    }
    for (size_t i = 1; i < count.size(); i++) count[i] += count[i - 1];
    for (size_t k = 0; k < act_vox.size(); k++) {
      const size_t dst = count[cls[k]]++;
      act_idx[dst] = act_vox[k];
      src_of[dst] = k;
    }
  }
  std::vector<float> act_val(act_idx.size());
  for (size_t i = 0; i < act_idx.size(); i++) act_val[i] = ph->voxel_mat_dens[3 * (size_t)act_idx[i] + 2];
  c->n_active = (long long)act_idx.size();
  // Balanced stream layout: voxel v gets max(1, ceil(E_v / stream_decays)) streams (E_v: its
  // expected decays over the acquisition); random-number stream ids count the streams in voxel
  // order, the device processes the voxels hottest first.
  std::vector<long long> stream_start, stream_id;
  long long balanced_streams = 0;
  if (cfg->stream_decays > 0 && !act_idx.empty()) {
    if (c->n_active > 0x7fffffffLL) {
      free_all(c);
      delete c;
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: too many emitting voxels for the balanced stream layout");
    }
    balanced_streams = balanced_layout(act_val, src_of, effective_time(cfg), cfg->stream_decays, stream_start, stream_id);
    if (cfg->detector.tally_psf_sinogram == MCPET_TALLY_SINGLES && balanced_streams >= (1LL << 32)) {
      free_all(c);
      delete c;
      return set_err(nullptr, MCPET_E_RANGE, "mcpet_create: more than 2^32 decay streams (singles event ids)");
    }
  }

  // Tables: split (a,b) of the total cross section (read per Woodcock step) from the
  // Compton/Rayleigh pairs (read per real interaction).
  const int N = tab->num_values;
  std::vector<float2> hw((size_t)N), htot((size_t)N * MCPET_MAX_MATERIALS);
  std::vector<float4> hcr((size_t)N * MCPET_MAX_MATERIALS);
  for (int i = 0; i < N; i++) hw[i] = make_float2(tab->woodcock[2 * i], tab->woodcock[2 * i + 1]);
  for (size_t b = 0; b < (size_t)N * MCPET_MAX_MATERIALS; b++) {
    const float* a = tab->mfp_a + 3 * b;
    const float* s = tab->mfp_b + 3 * b;
    htot[b] = make_float2(a[0], s[0]);
    hcr[b] = make_float4(a[1], s[1], a[2], s[2]);
  }

  // ROI (dose) size.
  const bool voxel_dose = cfg->voxel_dose && cfg->roi_x_max > -1;
  if (voxel_dose) {   // inside the phantom (checked above)
    c->n_roi = (long long)(cfg->roi_x_max - cfg->roi_x_min + 1) * (cfg->roi_y_max - cfg->roi_y_min + 1) *
               (cfg->roi_z_max - cfg->roi_z_min + 1);
  }
  const int mode = cfg->detector.tally_psf_sinogram;
  const bool want_psf = (mode == 0 || mode == 1) && cfg->detector.psf_size > 0;

  if (c->P.vox_fmt == 0) CK(dalloc(c, &c->d_vox, (size_t)c->nvox));
  if (c->P.vox_fmt == 1) CK(dalloc(c, &c->d_vox8, (size_t)c->nvox));
  if (c->P.vox_fmt == 2) CK(dalloc(c, &c->d_vox16, (size_t)c->nvox));
  if (c->P.vox_fmt != 0) CK(dalloc(c, &c->d_palette, palette.size()));
  CK(dalloc(c, &c->d_active_vox, std::max<size_t>(1, act_idx.size())));
  CK(dalloc(c, &c->d_active_act, std::max<size_t>(1, act_idx.size())));
  std::vector<int> stream_hint;
  int hint_shift = 0;
  if (!stream_start.empty()) {
    hint_shift = stream_hints(stream_start, stream_hint);
    CK(dalloc(c, &c->d_stream_start, stream_start.size()));
    CK(dalloc(c, &c->d_stream_id, stream_id.size()));
    CK(dalloc(c, &c->d_stream_hint, stream_hint.size()));
    c->hint_cap = stream_hint.size();
  }
  CK(dalloc(c, &c->d_woodcock, (size_t)N));
  CK(dalloc(c, &c->d_tot, htot.size()));
  CK(dalloc(c, &c->d_cr, hcr.size()));
  CK(dalloc(c, &c->d_ray, 1));
  CK(dalloc(c, &c->d_cmp, 1));
  CK(dalloc(c, &c->d_trues, (size_t)cfg->nbins));
  CK(dalloc(c, &c->d_scatter, (size_t)cfg->nbins));
  CK(dalloc(c, &c->d_img_t, (size_t)c->nvox));
  CK(dalloc(c, &c->d_img_sc, (size_t)c->nvox));
  CK(dalloc(c, &c->d_spectrum, (size_t)std::max(1, cfg->ne)));
  if (want_psf) CK(dalloc(c, &c->d_psf, (size_t)cfg->detector.psf_size));
  CK(dalloc(c, &c->d_psf_index, 1));
  if (voxel_dose) CK(dalloc(c, &c->d_voxel_edep, 2 * (size_t)c->n_roi));
  CK(dalloc(c, &c->d_mat_edep, 2 * MCPET_MAX_MATERIALS));
  CK(dalloc(c, &c->d_counters, KC_N + 1));
  CK(dalloc(c, &c->d_params, 1));
  HK(hipHostMalloc((void**)&c->h_params, sizeof(KParams), hipHostMallocDefault));
  c->singles_mode = (mode == MCPET_TALLY_SINGLES);
  if (c->singles_mode) {
    const long long cap = cfg->singles_capacity;
    CK(dalloc(c, &c->d_singles, (size_t)cap));
    CK(dalloc(c, &c->d_singles_n, 1));
    for (int k = 0; k < 2; k++) {
      CK(dalloc(c, &c->d_keys[k], (size_t)cap));
      CK(dalloc(c, &c->d_vals[k], (size_t)cap));
    }
    const long long m = 256LL * mcpet_rs_tiles(cap);
    CK(dalloc(c, &c->d_hist, (size_t)m));
    std::vector<long long> levels;
    mcpet_scan_levels(m, levels);
    c->d_scan.assign(levels.size(), nullptr);
    for (size_t k = 0; k < levels.size(); k++) CK(dalloc(c, &c->d_scan[k], (size_t)levels[k]));
    CK(dalloc(c, &c->d_randoms, (size_t)cfg->nbins));
    CK(dalloc(c, &c->d_spurious, (size_t)cfg->nbins));
    CK(dalloc(c, &c->d_coinc, 5));
    HK(hipEventCreate(&c->ev_c1));
  }
  CK(dalloc(c, &c->d_final_seed, 1));

  if (c->P.vox_fmt == 0) HK(hipMemcpy(c->d_vox, hv.data(), hv.size() * sizeof(float2), hipMemcpyHostToDevice));
  if (c->P.vox_fmt == 1) {
    std::vector<uint8_t> p8(pidx.begin(), pidx.end());
    HK(hipMemcpy(c->d_vox8, p8.data(), p8.size(), hipMemcpyHostToDevice));
  }
  if (c->P.vox_fmt == 2) HK(hipMemcpy(c->d_vox16, pidx.data(), pidx.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  if (c->P.vox_fmt != 0)
    HK(hipMemcpy(c->d_palette, palette.data(), palette.size() * sizeof(float2), hipMemcpyHostToDevice));
  if (!act_idx.empty()) {
    HK(hipMemcpy(c->d_active_vox, act_idx.data(), act_idx.size() * sizeof(int), hipMemcpyHostToDevice));
    HK(hipMemcpy(c->d_active_act, act_val.data(), act_val.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  if (!stream_start.empty()) {
    HK(hipMemcpy(c->d_stream_start, stream_start.data(), stream_start.size() * sizeof(long long), hipMemcpyHostToDevice));
    HK(hipMemcpy(c->d_stream_id, stream_id.data(), stream_id.size() * sizeof(long long), hipMemcpyHostToDevice));
    HK(hipMemcpy(c->d_stream_hint, stream_hint.data(), stream_hint.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  HK(hipMemcpy(c->d_woodcock, hw.data(), hw.size() * sizeof(float2), hipMemcpyHostToDevice));
  HK(hipMemcpy(c->d_tot, htot.data(), htot.size() * sizeof(float2), hipMemcpyHostToDevice));
  HK(hipMemcpy(c->d_cr, hcr.data(), hcr.size() * sizeof(float4), hipMemcpyHostToDevice));
  HK(hipMemcpy(c->d_ray, tab->rayleigh, sizeof(mcpet_rayleigh_table), hipMemcpyHostToDevice));
  HK(hipMemcpy(c->d_cmp, tab->compton, sizeof(mcpet_compton_table), hipMemcpyHostToDevice));

  KParams& P = c->P;
  P.positron_frac = (cfg->positron_fraction > 0.0f && cfg->positron_fraction < 1.0f) ? cfg->positron_fraction : 1.0f;
  P.prompt_frac = cfg->prompt_gamma_fraction;
  P.prompt_energy = cfg->prompt_gamma_energy;
  P.singles = c->d_singles;
  P.singles_n = c->d_singles_n;
  P.singles_cap = c->singles_mode ? cfg->singles_capacity : 0;
  P.window_ps = cfg->coincidence_window_ps;
  P.randoms = c->d_randoms;
  P.spurious = c->d_spurious;
  P.coinc_counts = c->d_coinc;
  P.vox = c->d_vox;
  P.vox8 = c->d_vox8;
  P.vox16 = c->d_vox16;
  P.palette = c->d_palette;
  P.nx = ph->nx;
  P.ny = ph->ny;
  P.nz = ph->nz;
  P.ivx = ph->inv_voxel_size[0];
  P.ivy = ph->inv_voxel_size[1];
  P.ivz = ph->inv_voxel_size[2];
  P.bx = ph->size_bbox[0];
  P.by = ph->size_bbox[1];
  P.bz = ph->size_bbox[2];
  P.active_vox = c->d_active_vox;
  P.active_act = c->d_active_act;
  P.tpb = cfg->threads_per_block;
  P.stream_start = c->d_stream_start;
  P.stream_id = c->d_stream_id;
  P.stream_hint = c->d_stream_hint;
  P.hint_shift = hint_shift;
  P.n_active = (int)std::min<long long>(c->n_active, 0x7fffffffLL);
  P.n_streams = c->d_stream_start ? balanced_streams : c->n_active * (long long)cfg->threads_per_block;
  P.queue = c->d_counters + KC_N;
  P.seed = cfg->seed;
  P.bg_mode = cfg->bg_mode;
  P.acq_ps = cfg->acquisition_time_ps;
  P.inv_mean_life = 1.0 / (double)cfg->mean_life;
  P.inv_nx = 1.0 / (double)ph->nx;
  P.inv_ny = 1.0 / (double)ph->ny;
  mcr_build_jump_table(MCR_A1, MCR_M1, P.j1);
  mcr_build_jump_table(MCR_A2, MCR_M2, P.j2);
  P.woodcock = c->d_woodcock;
  P.tot = c->d_tot;
  P.cr = c->d_cr;
  P.ray = c->d_ray;
  P.cmp = c->d_cmp;
  P.num_values = N;
  P.e0 = tab->e0;
  P.ide = tab->ide;
  P.cx = c->det.center[0];
  P.cy = c->det.center[1];
  P.cz = c->det.center[2];
  P.radius = c->det.radius;
  P.height = c->det.height;
  P.psf_size = want_psf ? cfg->detector.psf_size : 0;
  P.mode = mode;
  P.type = cfg->detector.tally_type;
  P.e_resol = cfg->e_resol;
  P.e_low = cfg->e_low;
  P.e_high = cfg->e_high;
  P.fovz = cfg->fovz;
  P.nrows = cfg->nrows;
  P.ncrystals = cfg->ncrystals;
  P.nangles = cfg->nangles;
  P.nrad = cfg->nrad;
  P.nsinos = cfg->nsinos;
  P.nzs = cfg->nzs;
  P.ne = cfg->ne;
  P.mrd = cfg->mrd;
  P.span = cfg->span;
  P.material_dose = cfg->material_dose ? 1 : 0;
  P.voxel_dose = voxel_dose ? 1 : 0;
  P.roi[0] = cfg->roi_x_min;
  P.roi[1] = cfg->roi_x_max;
  P.roi[2] = cfg->roi_y_min;
  P.roi[3] = cfg->roi_y_max;
  P.roi[4] = cfg->roi_z_min;
  P.roi[5] = cfg->roi_z_max;
  P.trues = c->d_trues;
  P.scatter = c->d_scatter;
  P.img_t = c->d_img_t;
  P.img_sc = c->d_img_sc;
  P.spectrum = c->d_spectrum;
  P.psf = c->d_psf;
  P.psf_index = c->d_psf_index;
  P.voxel_edep = c->d_voxel_edep;
  P.mat_edep = c->d_mat_edep;
  P.counters = c->d_counters;
  P.final_seed = c->d_final_seed;
  P.last_thread_id = c->d_stream_start ? balanced_streams - 1
                                        : (c->nvox - 1) * (long long)cfg->threads_per_block + (cfg->threads_per_block - 1);
  P.lds_spectrum = (cfg->ne > 0 && cfg->ne <= 8192) ? 1 : 0;
  P.lds_palette = 1;
  P.lds_xs511 = 1;
  P.stream_stride = 1;
  *out = c;
#undef CK
#undef HK
  return MCPET_OK;
}

// Clear the tallies (and with counters, the event counters and stream queue) on `stream` with
// one fill kernel for all of them (one launch instead of a memset per buffer, and no memset
// nodes when a graph captures it).
static mcpet_status clear_buffers(mcpet_ctx* c, hipStream_t stream, bool tallies, bool counters)
{
  struct Region {
    void* p;
    size_t bytes;
  };
  std::vector<Region> r;
  if (tallies) {
    r.push_back({c->d_trues, sizeof(int) * (size_t)c->cfg.nbins});
    r.push_back({c->d_scatter, sizeof(int) * (size_t)c->cfg.nbins});
    r.push_back({c->d_img_t, sizeof(int) * (size_t)c->nvox});
    r.push_back({c->d_img_sc, sizeof(int) * (size_t)c->nvox});
    r.push_back({c->d_spectrum, sizeof(int) * (size_t)std::max(1, c->cfg.ne)});
    r.push_back({c->d_psf_index, sizeof(int)});
    if (c->d_voxel_edep) r.push_back({c->d_voxel_edep, 16 * (size_t)c->n_roi});
    if (c->d_rbt) r.push_back({c->d_rbt, 2 * sizeof(int) * (size_t)c->rbt_dims[0] * c->rbt_dims[1] * c->rbt_dims[2]});
    r.push_back({c->d_mat_edep, 16 * MCPET_MAX_MATERIALS});
    r.push_back({c->d_final_seed, sizeof(int)});
    if (c->singles_mode) {
      r.push_back({c->d_randoms, sizeof(int) * (size_t)c->cfg.nbins});
      r.push_back({c->d_spurious, sizeof(int) * (size_t)c->cfg.nbins});
    }
  }
  if (counters) r.push_back({c->d_counters, sizeof(unsigned long long) * (KC_N + 1)});
  std::vector<void*> ptrs;
  std::vector<size_t> words;
  for (const Region& x : r)
    if (x.p && x.bytes) {
      ptrs.push_back(x.p);
      words.push_back(x.bytes / 4);
    }
  const hipError_t e = mcpet_launch_fill_regions(ptrs.data(), words.data(), (int)ptrs.size(), stream);
  if (e != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("clearing tallies: ") + hipGetErrorString(e));
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_run(mcpet_ctx* c, const mcpet_run_params* rp)
{
  if (!c) return set_err(nullptr, MCPET_E_ARG, "mcpet_run: null context");
  mcpet_run_params p{};
  p.zero_tallies = 1;
  if (rp) p = *rp;
  HIPCHK(c, hipSetDevice(c->device));
  // the parameter block's staging copy may still be read by the previous run's copy
  HIPCHK(c, hipStreamSynchronize(c->stream));
  KParams P = c->P;
  if (p.seed != 0) P.seed = p.seed;
  if (!seed_usable(P.seed)) return set_err(c, MCPET_E_ARG, seed_msg(P.seed));
  if (p.acquisition_time_ps != 0) P.acq_ps = p.acquisition_time_ps;
  // Phase selection (persistent kernel): most lanes by default (measured best on the head
  // workload); policy 2 weighs the phases by relative cost (DECAY, SOURCE2, STEP, INTERACT, TALLY).
  // Launch shape, policy and LDS layout left 0 take mcpet_autotune's choices when it ran.
  if (p.block_size <= 0 && c->tuned_bs > 0) p.block_size = c->tuned_bs;
  if (p.blocks_per_cu <= 0 && c->tuned_per_cu > 0) p.blocks_per_cu = c->tuned_per_cu;
  if (p.phase_policy <= 0 && c->tuned_policy > 0) p.phase_policy = c->tuned_policy;
  const int lds = (p.lds_layout & MCPET_LDS_SET) ? p.lds_layout : c->tuned_lds;
  if (lds >= 0) {
    P.lds_palette = (lds & MCPET_LDS_PALETTE) ? 1 : 0;
    P.lds_spectrum = (lds & MCPET_LDS_SPECTRUM) && c->P.lds_spectrum ? 1 : 0;   // only where the spectrum fits
    P.lds_xs511 = (lds & MCPET_LDS_XS511) ? 1 : 0;
  }
  P.sm_policy = (p.phase_policy == 2) ? 1 : 0;
  static const int default_cost[MCPET_NPHASES] = {4, 6, 1, 3, 6, 3, 2, 4, 4, 1};
  for (int i = 0; i < MCPET_NPHASES; i++) P.sm_cost[i] = p.phase_cost[i] > 0 ? p.phase_cost[i] : default_cost[i];
  int kind = p.kernel == MCPET_KERNEL_AUTO ? MCPET_KERNEL_PERSISTENT : p.kernel;
  if (kind == MCPET_KERNEL_PLAIN && (c->singles_mode || P.positron_frac < 1.0f || P.prompt_frac > 0.0f))
    return set_err(c, MCPET_E_ARG, "mcpet_run: the plain kernel supports the reference's pair tallies only (no singles, "
                                   "positron branching or prompt gammas)");
  int bs = p.block_size > 0 ? p.block_size : 256;
  if (bs % 64 != 0 || bs > 256) return set_err(c, MCPET_E_RANGE, "mcpet_run: block_size must be a multiple of 64, <= 256");
  HIPCHK(c, hipEventRecord(c->ev_z0, c->stream));
  if (p.zero_tallies) {   // tallies and counters in one fill launch
    mcpet_status st = clear_buffers(c, c->stream, true, true);
    if (st != MCPET_OK) return st;
  } else {
    HIPCHK(c, hipMemsetAsync(c->d_counters, 0, sizeof(unsigned long long) * (KC_N + 1), c->stream));
  }
  if (c->singles_mode) {
    HIPCHK(c, hipMemsetAsync(c->d_singles_n, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_coinc, 0, 5 * sizeof(unsigned long long), c->stream));
  }
  HIPCHK(c, hipEventRecord(c->ev_k0, c->stream));
  long long blocks;
  if (kind == MCPET_KERNEL_PERSISTENT) {
    int per_cu = p.blocks_per_cu > 0 ? p.blocks_per_cu : 8;
    blocks = (long long)c->num_cus * per_cu;
    long long need = (P.n_streams + bs - 1) / bs;
    if (blocks > need) blocks = std::max(1LL, need);
  } else {
    blocks = std::max(1LL, (P.n_streams + bs - 1) / bs);
    if (blocks > 0x7fffffffLL) return set_err(c, MCPET_E_RANGE, "mcpet_run: too many streams for one launch");
  }
  // the rebinned tallies hold the sinograms' counts only if every run since the last clear added to them
  c->rbt_ran = P.rb_trues != nullptr && (p.zero_tallies || c->rbt_ran);
  if (P.n_streams > 0) {
    *c->h_params = P;
    HIPCHK(c, hipMemcpyAsync(c->d_params, c->h_params, sizeof(KParams), hipMemcpyHostToDevice, c->stream));
    hipError_t e = mcpet_launch_transport(P, c->d_params, kind, p.count_events, (int)blocks, bs, c->stream);
    if (e != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("transport launch failed: ") + hipGetErrorString(e));
  }
  HIPCHK(c, hipEventRecord(c->ev_k1, c->stream));
  c->ran = true;
  c->timed = true;
  if (!c->singles_mode) {
    HIPCHK(c, hipEventRecord(c->ev_done, c->stream));
    return MCPET_OK;
  }
  // Singles pipeline: the sort needs the number of singles, so wait for the transport.
  HIPCHK(c, hipMemcpyAsync(&c->singles_req, c->d_singles_n, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_singles = (long long)std::min<unsigned long long>(c->singles_req, (unsigned long long)P.singles_cap);
  // detection times are below acq + 2^40 ps (flight times are clamped there)
  const unsigned long long tmax = P.acq_ps + (1ULL << 40);
  int bits = 64;
  while (bits > 1 && !((tmax >> (bits - 1)) & 1ULL)) bits--;
  const int passes = (bits + 7) / 8;
  hipError_t e = mcpet_launch_keys_from_singles(c->d_singles, c->n_singles, c->d_keys[0], c->d_vals[0], c->stream);
  if (e == hipSuccess) e = mcpet_radix_sort(c->n_singles, passes, c->d_keys, c->d_vals, c->d_hist, c->d_scan.data(), &c->sorted, c->stream);
  if (e == hipSuccess) e = mcpet_launch_pairing(P, c->d_keys[c->sorted], c->d_vals[c->sorted], c->n_singles, c->stream);
  if (e != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("coincidence sorting failed: ") + hipGetErrorString(e));
  HIPCHK(c, hipEventRecord(c->ev_c1, c->stream));
  HIPCHK(c, hipEventRecord(c->ev_done, c->stream));
  return MCPET_OK;
}

// ---------------------------------------------------------------------------------- autotuner

// The context's streams in another layout (stream_decays 0: the reference's threads_per_block
// streams per emitting voxel; > 0: the balanced layout). The emitting voxels' device order and
// activities are read back from the device; balanced arrays stay allocated once made.
static mcpet_status apply_stream_layout(mcpet_ctx* c, long long stream_decays)
{
  KParams& P = c->P;
  const long long n = c->n_active;
  const long long tpb = c->cfg.threads_per_block;
  if (stream_decays <= 0 || n <= 0) {
    P.stream_start = nullptr;
    P.stream_id = nullptr;
    P.stream_hint = nullptr;
    P.n_streams = n * tpb;
    P.last_thread_id = (c->nvox - 1) * tpb + (tpb - 1);
    c->cfg.stream_decays = 0;
    return MCPET_OK;
  }
  std::vector<int> vox((size_t)n);
  std::vector<float> act((size_t)n);
  HIPCHK(c, hipMemcpy(vox.data(), c->d_active_vox, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(act.data(), c->d_active_act, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
  std::vector<size_t> order((size_t)n), src_of((size_t)n);
  for (size_t k = 0; k < order.size(); k++) order[k] = k;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return vox[a] < vox[b]; });
  for (size_t j = 0; j < order.size(); j++) src_of[order[j]] = j;   // voxel order = ascending voxel index
  std::vector<long long> start, id;
  const long long total = balanced_layout(act, src_of, effective_time(&c->cfg), stream_decays, start, id);
  if (c->singles_mode && total >= (1LL << 32))
    return set_err(c, MCPET_E_RANGE, "mcpet_autotune: more than 2^32 decay streams (singles event ids)");
  std::vector<int> hint;
  const int shift = stream_hints(start, hint);
  if (!c->d_stream_start) {
    mcpet_status st = dalloc(c, &c->d_stream_start, start.size());
    if (st == MCPET_OK) st = dalloc(c, &c->d_stream_id, id.size());
    if (st != MCPET_OK) return st;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (hint.size() > c->hint_cap) {   // at most n + 2 entries: allocated once at that size
    if (c->d_stream_hint) {
      (void)hipFree(c->d_stream_hint);
      c->dev_bytes -= c->hint_cap * sizeof(int);
      c->d_stream_hint = nullptr;
      c->hint_cap = 0;
    }
    mcpet_status st = dalloc(c, &c->d_stream_hint, (size_t)n + 2);
    if (st != MCPET_OK) return st;
    c->hint_cap = (size_t)n + 2;
  }
  HIPCHK(c, hipMemcpy(c->d_stream_start, start.data(), sizeof(long long) * start.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_stream_id, id.data(), sizeof(long long) * id.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_stream_hint, hint.data(), sizeof(int) * hint.size(), hipMemcpyHostToDevice));
  P.stream_start = c->d_stream_start;
  P.stream_id = c->d_stream_id;
  P.stream_hint = c->d_stream_hint;
  P.hint_shift = shift;
  P.n_streams = total;
  P.last_thread_id = total - 1;
  c->cfg.stream_decays = (int32_t)stream_decays;
  return MCPET_OK;
}

// One timed sample: every stride-th decay stream over the whole acquisition, with the given
// launch shape, phase policy and LDS layout; decays per millisecond of kernel time.
static mcpet_status tune_sample(mcpet_ctx* c, long long stride, int bs, int per_cu, int policy, int lds, mcpet_tune_result* r)
{
  mcpet_run_params p{};
  p.kernel = MCPET_KERNEL_PERSISTENT;
  p.zero_tallies = 1;
  p.block_size = bs;
  p.blocks_per_cu = per_cu;
  p.phase_policy = policy;
  const int saved_lds = c->tuned_lds;
  c->tuned_lds = lds;
  c->P.stream_stride = stride;
  mcpet_status st = mcpet_run(c, &p);
  c->P.stream_stride = 1;
  c->tuned_lds = saved_lds;
  mcpet_stats s{};
  if (st == MCPET_OK) st = mcpet_get_stats(c, &s);
  if (st != MCPET_OK) return st;
  r->block_size = bs;
  r->blocks_per_cu = per_cu;
  r->phase_policy = policy;
  r->ms = s.kernel_ms;
  r->decays_per_ms = s.kernel_ms > 0.0 ? (double)s.decays / s.kernel_ms : 0.0;
  r->lds_layout = lds;
  r->stream_decays = c->cfg.stream_decays;
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_autotune(mcpet_ctx* c, const mcpet_tune_options* opt, mcpet_tune_result* results,
                                       int32_t max_results, int32_t* n_results)
{
  if (!c || !opt || !(opt->stream_fraction > 0.0 && opt->stream_fraction <= 1.0) || opt->n_stream_decays < 0 ||
      opt->n_stream_decays > 8)
    return set_err(c, MCPET_E_ARG, "mcpet_autotune: bad arguments");
  for (int i = 0; i < opt->n_stream_decays; i++)
    if (opt->stream_decays[i] < 0 || opt->stream_decays[i] > 0x7fffffffLL)
      return set_err(c, MCPET_E_ARG, "mcpet_autotune: decays per stream outside [0, 2^31)");
  if (c->P.n_streams <= 0) return set_err(c, MCPET_E_STATE, "mcpet_autotune: the phantom has no emitting voxel");
  const long long stride = std::max(1LL, (long long)std::llround(1.0 / opt->stream_fraction));
  int written = 0;
  auto keep = [&](const mcpet_tune_result& r) {
    if (results && written < max_results) results[written] = r;
    written++;
  };
  // the layout's default: everything in LDS that applies to the phantom
  int lds = MCPET_LDS_XS511 | (c->P.vox_fmt == 1 ? MCPET_LDS_PALETTE : 0) | (c->P.lds_spectrum ? MCPET_LDS_SPECTRUM : 0);
  mcpet_tune_result r{}, best{};
  best.decays_per_ms = -1.0;
  mcpet_status st = tune_sample(c, stride, 256, 4, 1, lds, &r);   // warm-up (code load, caches)
  if (st != MCPET_OK) return st;
  // 1. launch shape x phase policy
  static const int shapes[][2] = {{256, 4}, {256, 8}, {128, 8}};
  for (const auto& sh : shapes)
    for (int pol = 1; pol <= 2; pol++) {
      if ((st = tune_sample(c, stride, sh[0], sh[1], pol, lds, &r)) != MCPET_OK) return st;
      keep(r);
      if (r.decays_per_ms > best.decays_per_ms) best = r;
    }
  // 2. LDS layout: each placement moved to global memory in turn, kept when faster
  if (opt->tune_layout) {
    for (int bit : {MCPET_LDS_PALETTE, MCPET_LDS_SPECTRUM, MCPET_LDS_XS511}) {
      if (!(lds & bit)) continue;
      if ((st = tune_sample(c, stride, best.block_size, best.blocks_per_cu, best.phase_policy, best.lds_layout & ~bit, &r)) !=
          MCPET_OK)
        return st;
      keep(r);
      if (r.decays_per_ms > best.decays_per_ms) best = r;
    }
  }
  // 3. decays per stream (the reference's histories per thread): opt-in, it changes the
  //    random-number streams
  if (opt->tune_streams && opt->n_stream_decays > 0) {
    for (int i = 0; i < opt->n_stream_decays; i++) {
      if ((st = apply_stream_layout(c, opt->stream_decays[i])) != MCPET_OK) return st;
      if ((st = tune_sample(c, stride, best.block_size, best.blocks_per_cu, best.phase_policy, best.lds_layout, &r)) != MCPET_OK)
        return st;
      keep(r);
      if (r.decays_per_ms > best.decays_per_ms) best = r;
    }
    if ((st = apply_stream_layout(c, best.stream_decays)) != MCPET_OK) return st;   // the chosen layout
  }
  c->tuned_bs = best.block_size;
  c->tuned_per_cu = best.blocks_per_cu;
  c->tuned_policy = best.phase_policy;
  c->tuned_lds = best.lds_layout;
  c->ran = false;   // the samples' tallies are not a run's
  keep(best);       // the choice, after the timings
  if (n_results) *n_results = std::min(written, std::max(0, max_results));
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_sync(mcpet_ctx* c)
{
  if (!c) return set_err(nullptr, MCPET_E_ARG, "mcpet_sync: null context");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_fetch(mcpet_ctx* c, mcpet_tallies* t)
{
  if (!c || !t) return set_err(c, MCPET_E_ARG, "mcpet_fetch: null argument");
  if (!c->ran) return set_err(c, MCPET_E_STATE, "mcpet_fetch: no run to fetch");
  mcpet_status st = mcpet_sync(c);
  if (st != MCPET_OK) return st;
  const size_t nb = (size_t)c->cfg.nbins, nv = (size_t)c->nvox;
  if (t->trues) HIPCHK(c, hipMemcpy(t->trues, c->d_trues, 4 * nb, hipMemcpyDeviceToHost));
  if (t->scatter) HIPCHK(c, hipMemcpy(t->scatter, c->d_scatter, 4 * nb, hipMemcpyDeviceToHost));
  if (t->image_trues) HIPCHK(c, hipMemcpy(t->image_trues, c->d_img_t, 4 * nv, hipMemcpyDeviceToHost));
  if (t->image_scatter) HIPCHK(c, hipMemcpy(t->image_scatter, c->d_img_sc, 4 * nv, hipMemcpyDeviceToHost));
  if (t->spectrum && c->cfg.ne > 0) HIPCHK(c, hipMemcpy(t->spectrum, c->d_spectrum, 4 * (size_t)c->cfg.ne, hipMemcpyDeviceToHost));
  int idx = 0;
  HIPCHK(c, hipMemcpy(&idx, c->d_psf_index, sizeof(int), hipMemcpyDeviceToHost));
  t->psf_requested = idx;
  long long n = std::min<long long>(idx, c->P.psf_size);
  if (n < 0) n = 0;
  t->psf_count = n;
  if (t->psf && n > 0) {
    if (t->psf_capacity < n) return set_err(c, MCPET_E_CAPACITY, "mcpet_fetch: psf buffer smaller than the records written");
    HIPCHK(c, hipMemcpy(t->psf, c->d_psf, (size_t)n * sizeof(mcpet_psf_record), hipMemcpyDeviceToHost));
  }
  if (t->voxel_edep && c->d_voxel_edep)
    HIPCHK(c, hipMemcpy(t->voxel_edep, c->d_voxel_edep, 16 * (size_t)c->n_roi, hipMemcpyDeviceToHost));
  if (t->material_edep) HIPCHK(c, hipMemcpy(t->material_edep, c->d_mat_edep, 16 * MCPET_MAX_MATERIALS, hipMemcpyDeviceToHost));
  unsigned long long cnt[KC_N + 1];
  HIPCHK(c, hipMemcpy(cnt, c->d_counters, sizeof cnt, hipMemcpyDeviceToHost));
  t->total_histories = 2ull * cnt[KC_DECAYS];  // two photon histories per decay (K:282)
  HIPCHK(c, hipMemcpy(&t->final_seed, c->d_final_seed, sizeof(int), hipMemcpyDeviceToHost));
  t->singles_count = c->singles_mode ? c->n_singles : 0;
  t->singles_requested = c->singles_mode ? (long long)c->singles_req : 0;
  for (int k = 0; k < 5; k++) t->coincidences[k] = 0;
  if (c->singles_mode) {
    if (t->randoms) HIPCHK(c, hipMemcpy(t->randoms, c->d_randoms, 4 * nb, hipMemcpyDeviceToHost));
    if (t->spurious) HIPCHK(c, hipMemcpy(t->spurious, c->d_spurious, 4 * nb, hipMemcpyDeviceToHost));
    unsigned long long cc[5];
    HIPCHK(c, hipMemcpy(cc, c->d_coinc, sizeof cc, hipMemcpyDeviceToHost));
    for (int k = 0; k < 5; k++) t->coincidences[k] = (int64_t)cc[k];
    if (t->singles && c->n_singles > 0) {
      if (t->singles_buffer < c->n_singles)
        return set_err(c, MCPET_E_CAPACITY, "mcpet_fetch: singles buffer smaller than the singles stored");
      mcpet_single* tmp = nullptr;
      HIPCHK(c, hipMalloc((void**)&tmp, (size_t)c->n_singles * sizeof(mcpet_single)));
      hipError_t e = mcpet_launch_gather_singles(c->d_singles, c->d_vals[c->sorted], c->n_singles, tmp, c->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(t->singles, tmp, (size_t)c->n_singles * sizeof(mcpet_single), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      (void)hipFree(tmp);
      if (e != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("mcpet_fetch: singles: ") + hipGetErrorString(e));
    }
  }
  // Sums over the sinograms (host copies if fetched, else device-side counters).
  if (t->trues && t->scatter) {
    long long a = 0, b = 0;
    for (size_t i = 0; i < nb; i++) {
      a += t->trues[i];
      b += t->scatter[i];
    }
    t->sum_trues = a;
    t->sum_scatter = b;
  } else {
    t->sum_trues = -1;
    t->sum_scatter = -1;
  }
  return MCPET_OK;
}

static mcpet_status read_counts(mcpet_ctx* c, mcpet_stats* s);

extern "C" mcpet_status mcpet_get_stats(mcpet_ctx* c, mcpet_stats* s)
{
  if (!c || !s) return set_err(c, MCPET_E_ARG, "mcpet_get_stats: null argument");
  if (!c->ran) return set_err(c, MCPET_E_STATE, "mcpet_get_stats: no run");
  mcpet_status st = mcpet_sync(c);
  if (st != MCPET_OK) return st;
  std::memset(s, 0, sizeof *s);
  float ms = 0.0f, zms = 0.0f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev_k0, c->ev_k1));
  HIPCHK(c, hipEventElapsedTime(&zms, c->ev_z0, c->ev_k0));
  s->kernel_ms = ms;
  s->zero_ms = zms;
  return read_counts(c, s);
}

static mcpet_status read_counts(mcpet_ctx* c, mcpet_stats* s)
{
  unsigned long long cnt[KC_N + 1];
  HIPCHK(c, hipMemcpy(cnt, c->d_counters, sizeof cnt, hipMemcpyDeviceToHost));
  s->decays = cnt[KC_DECAYS];
  s->streams = cnt[KC_STREAMS];
  s->woodcock_steps = cnt[KC_WOODCOCK];
  s->material_changes = cnt[KC_MATCHANGE];
  s->compton = cnt[KC_COMPTON];
  s->compton_shell_evals = cnt[KC_SHELLEVAL];
  s->rayleigh = cnt[KC_RAYLEIGH];
  s->photoelectric = cnt[KC_PHOTO];
  s->pairs_tallied = cnt[KC_PAIRS];
  s->coincidences = cnt[KC_COINC];
  s->rn_draws = cnt[KC_RNDRAWS];
  s->pairs_emitted = cnt[KC_EMIT_PAIR];
  s->prompts_emitted = cnt[KC_EMIT_PROMPT];
  if (c->singles_mode) {
    float cms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&cms, c->ev_k1, c->ev_c1));
    s->coinc_ms = cms;
    s->singles = c->singles_req;
    unsigned long long cc[5];
    HIPCHK(c, hipMemcpy(cc, c->d_coinc, sizeof cc, hipMemcpyDeviceToHost));
    s->coincidences = cc[0] + cc[1];
  }
  for (int i = 0; i < MCPET_NPHASES; i++) {
    s->phase_exec[i] = cnt[KC_PHASE_EXEC + i];
    s->phase_lanes[i] = cnt[KC_PHASE_LANES + i];
    s->phase_cycles[i] = cnt[KC_PHASE_CYCLES + i];
  }
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_get_detector(const mcpet_ctx* c, mcpet_detector* out)
{
  if (!c || !out) return MCPET_E_ARG;
  *out = c->det;
  return MCPET_OK;
}

// Single-slice rebinning of a sinogram tally onto a D x H x W grid (plane groups of
// ceil(NZS / D) ring sums, angle groups of ceil(NANGLES / H), radial groups of ceil(NRAD / W)).
// rebin_setup builds (once per grid) the plane-group lists and the context's output buffer.
// wait_run: the host waits for the run (and any earlier rebinning) first; without it the
// caller orders its launches after the run itself, and the host waits only when the grid's
// plane lists have to be (re)built.
static mcpet_status rebin_setup(mcpet_ctx* c, int D, int H, int W, bool wait_run = true)
{
  if (!c || D < 1 || H < 1 || W < 1 || W > 256) return set_err(c, MCPET_E_ARG, "rebin: bad arguments (W <= 256)");
  if (!c->ran) return set_err(c, MCPET_E_STATE, "rebin: no run");
  HIPCHK(c, hipSetDevice(c->device));
  const bool cached = c->rb_dims[0] == D && c->rb_dims[1] == H && c->rb_dims[2] == W;
  if (wait_run || !cached) HIPCHK(c, hipStreamSynchronize(c->stream));
  if (cached) return MCPET_OK;
  const mcpet_config& cfg = c->cfg;
  const int gz = (cfg.nzs + D - 1) / D;
  std::vector<int> plane(cfg.nsinos);
  if (mcpet_sinogram_planes(&cfg, plane.data()) != MCPET_OK) return set_err(c, MCPET_E_ARG, mcpet_host_last_error());
  std::vector<int> off(D + 1, 0), list;
  for (int d = 0; d < D; d++) {
    for (int izm = 0; izm < cfg.nsinos; izm++)
      if (plane[izm] / gz == d) list.push_back(izm);
    off[d + 1] = (int)list.size();
  }
  for (void* p : {(void*)c->d_rb_off, (void*)c->d_rb_sino, (void*)c->d_rb_out})
    if (p) (void)hipFree(p);
  c->d_rb_off = c->d_rb_sino = nullptr;
  c->d_rb_out = nullptr;
  c->rb_dims[0] = c->rb_dims[1] = c->rb_dims[2] = 0;
  mcpet_status e = dalloc(c, &c->d_rb_off, (size_t)D + 1);
  if (e == MCPET_OK) e = dalloc(c, &c->d_rb_sino, std::max<size_t>(1, list.size()));
  if (e == MCPET_OK) e = dalloc(c, &c->d_rb_out, (size_t)D * H * W);
  if (e != MCPET_OK) return e;
  HIPCHK(c, hipMemcpy(c->d_rb_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
  if (!list.empty()) HIPCHK(c, hipMemcpy(c->d_rb_sino, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice));
  c->rb_dims[0] = D;
  c->rb_dims[1] = H;
  c->rb_dims[2] = W;
  return MCPET_OK;
}

// Enqueue the rebinning of tally `which` into the device buffer d_out on stream st (after
// rebin_setup for this grid).
static const int* rebin_source(const mcpet_ctx* c, int which)
{
  return which == 0 ? c->d_trues : which == 1 ? c->d_scatter : which == 2 ? c->d_randoms : which == 3 ? c->d_spurious : nullptr;
}

static mcpet_status rebin_launch(mcpet_ctx* c, int which, int D, int H, int W, float* d_out, hipStream_t st)
{
  const int* src = rebin_source(c, which);
  if (which < 0 || which > 3) return set_err(c, MCPET_E_ARG, "rebin: which must be 0 (trues), 1 (scatter), 2 (randoms), 3 (spurious)");
  if (!src) return set_err(c, MCPET_E_STATE, "rebin: this tally is not kept in this mode");
  const mcpet_config& cfg = c->cfg;
  const int gh = (cfg.nangles + H - 1) / H, gw = (cfg.nrad + W - 1) / W;
  hipError_t e = mcpet_launch_rebin(src, c->d_rb_off, c->d_rb_sino, cfg.nangles, cfg.nrad, D, H, W, gh, gw, d_out, st);
  if (e != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("rebin launch failed: ") + hipGetErrorString(e));
  return MCPET_OK;
}

static mcpet_status rebin(mcpet_ctx* c, int which, int D, int H, int W, float* out)
{
  if (!out) return set_err(c, MCPET_E_ARG, "rebin: null output");
  mcpet_status st = rebin_setup(c, D, H, W);
  if (st == MCPET_OK) st = rebin_launch(c, which, D, H, W, c->d_rb_out, c->stream);
  if (st != MCPET_OK) return st;
  HIPCHK(c, hipMemcpyAsync(out, c->d_rb_out, sizeof(float) * (size_t)D * H * W, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MCPET_OK;
}

mcpet_status mcpet_ctx_rebin_device(mcpet_ctx* c, int which, int D, int H, int W, float* d_out, hipStream_t st)
{
  if (!d_out) return set_err(c, MCPET_E_ARG, "rebin: null output");
  mcpet_status e = rebin_setup(c, D, H, W);
  if (e == MCPET_OK) e = rebin_launch(c, which, D, H, W, d_out, st);
  return e;
}

mcpet_status mcpet_ctx_rebin2_device(mcpet_ctx* c, int wa, int wb, int D, int H, int W, float* d_a, float* d_b, hipStream_t st)
{
  if (!d_a || !d_b) return set_err(c, MCPET_E_ARG, "rebin: null output");
  if (wa < 0 || wa > 3 || wb < 0 || wb > 3)
    return set_err(c, MCPET_E_ARG, "rebin: which must be 0 (trues), 1 (scatter), 2 (randoms), 3 (spurious)");
  const int *sa = rebin_source(c, wa), *sb = rebin_source(c, wb);
  if (!sa || !sb) return set_err(c, MCPET_E_STATE, "rebin: this tally is not kept in this mode");
  const bool pair = (wa == 0 && wb == 1) || (wa == 1 && wb == 0);
  if (pair && c->rbt_ran && c->rbt_dims[0] == D && c->rbt_dims[1] == H && c->rbt_dims[2] == W) {
    // the run tallied them on this grid: the same integer sums, no sinogram read
    if (!c->ran) return set_err(c, MCPET_E_STATE, "rebin: no run");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamWaitEvent(st, c->ev_done, 0));
    const size_t n = (size_t)D * H * W;
    const int* ra = wa == 0 ? c->d_rbt : c->d_rbt + n;
    const int* rb = wb == 0 ? c->d_rbt : c->d_rbt + n;
    hipError_t he = mcpet_launch_i2f2(ra, rb, d_a, d_b, n, st);
    if (he != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("rebinned tallies: ") + hipGetErrorString(he));
    return MCPET_OK;
  }
  mcpet_status e = rebin_setup(c, D, H, W, false);
  if (e != MCPET_OK) return e;
  // after the context's last run on the device, whichever launch ran it (no host wait)
  HIPCHK(c, hipStreamWaitEvent(st, c->ev_done, 0));
  const mcpet_config& cfg = c->cfg;
  const int gh = (cfg.nangles + H - 1) / H, gw = (cfg.nrad + W - 1) / W;
  hipError_t he = mcpet_launch_rebin2(sa, sb, c->d_rb_off, c->d_rb_sino, cfg.nangles, cfg.nrad, D, H, W, gh, gw, d_a, d_b, st);
  if (he != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("rebin launch failed: ") + hipGetErrorString(he));
  return MCPET_OK;
}

const unsigned long long* mcpet_ctx_decay_counter(const mcpet_ctx* c) { return c->d_counters + KC_DECAYS; }

mcpet_status mcpet_ctx_decays(mcpet_ctx* c, long long* decays)
{
  unsigned long long n = 0;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpy(&n, c->d_counters + KC_DECAYS, sizeof n, hipMemcpyDeviceToHost));
  *decays = (long long)n;
  return MCPET_OK;
}

// ---------------------------------------------------------------- line integrals on the grid

// Crystal pairs (k1 < k2) of every transaxial group of an H x W grid in CSR form (group
// h W + w), grouped the way the rebinning groups sinogram bins: angle / ceil(NANGLES / H),
// radial / ceil(NRAD / W), with the angle and radial bin of a pair from the binning of
// coincidences (K:1776-1786).
static void lor_groups_build(const KParams& P, int H, int W, std::vector<int>& off, std::vector<int2>& lors)
{
  const int N = P.ncrystals, gh = (P.nangles + H - 1) / H, gw = (P.nrad + W - 1) / W;
  // group of every pair (-1: outside the sinogram), counted, then placed by a counting sort that
  // keeps the (k1, k2) order inside each group
  std::vector<int> gid((size_t)N * N, -1);
  off.assign((size_t)H * W + 1, 0);
  for (int k1 = 0; k1 < N; k1++)
    for (int k2 = k1 + 1; k2 < N; k2++) {
      const int ith = ((k1 + k2 + P.nangles + 1) % N) / 2;
      int ir = std::abs(k2 - k1 - P.nangles);
      if (k1 < ith || k2 >= ith + P.nangles) ir = -ir;
      ir += P.nrad / 2;
      if (ir < 0 || ir >= P.nrad || ith < 0 || ith >= P.nangles) continue;
      const int g = (ith / gh) * W + ir / gw;
      gid[(size_t)k1 * N + k2] = g;
      off[g + 1]++;
    }
  for (size_t i = 0; i + 1 < off.size(); i++) off[i + 1] += off[i];
  lors.resize((size_t)off.back());
  std::vector<int> next(off.begin(), off.end() - 1);
  for (int k1 = 0; k1 < N; k1++)
    for (int k2 = k1 + 1; k2 < N; k2++) {
      const int g = gid[(size_t)k1 * N + k2];
      if (g >= 0) lors[(size_t)next[g]++] = int2{k1, k2};
    }
}

// lor_groups_build on the device, kept per (device, crystals, angles, radial bins, H, W) until
// the process ends (every context of one scanner geometry on a device shares them).
struct LorLists {
  int2* lors = nullptr;
  int* off = nullptr;
};

static mcpet_status lor_lists(mcpet_ctx* c, int H, int W, LorLists* out)
{
  static std::mutex mu;
  static std::map<std::array<int, 6>, LorLists> cache;
  const KParams& P = c->P;
  const std::array<int, 6> key = {c->device, P.ncrystals, P.nangles, P.nrad, H, W};
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) {
    *out = it->second;
    return MCPET_OK;
  }
  std::vector<int> off;
  std::vector<int2> lors;
  lor_groups_build(P, H, W, off, lors);
  // one allocation and one copy: the pairs, then the offsets
  const size_t lb = sizeof(int2) * lors.size(), ob = sizeof(int) * off.size();
  std::vector<char> host(lb + ob);
  if (lb) std::memcpy(host.data(), lors.data(), lb);
  std::memcpy(host.data() + lb, off.data(), ob);
  char* d = nullptr;
  hipError_t e = hipMalloc((void**)&d, host.size());
  if (e == hipSuccess) e = hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (d) (void)hipFree(d);
    return set_err(c, MCPET_E_NOMEM, std::string("project: crystal-pair lists: ") + hipGetErrorString(e));
  }
  LorLists l;
  l.lors = reinterpret_cast<int2*>(d);
  l.off = reinterpret_cast<int*>(d + lb);
  cache[key] = l;
  *out = l;
  return MCPET_OK;
}

mcpet_status mcpet_ctx_project_device(mcpet_ctx* c, int D, int H, int W, const float** d_maps)
{
  if (!c || D < 1 || D > 256 || H < 1 || W < 1 || !d_maps)
    return set_err(c, MCPET_E_ARG, "project: bad arguments (1 <= D <= 256, H, W >= 1)");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->pj_dims[0] == D && c->pj_dims[1] == H && c->pj_dims[2] == W) {
    *d_maps = c->d_pj_out;
    return MCPET_OK;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  LorLists l;
  mcpet_status e = lor_lists(c, H, W, &l);
  if (e != MCPET_OK) return e;
  if (c->d_pj_out) {   // maps of another grid
    (void)hipFree(c->d_pj_out);
    c->dev_bytes -= sizeof(float) * (3 * (size_t)c->pj_alloc + 1);
  }
  c->d_pj_out = nullptr;
  c->pj_dims[0] = c->pj_dims[1] = c->pj_dims[2] = 0;
  const size_t n = (size_t)D * H * W;
  c->pj_alloc = 0;
  if (!c->d_pj_params) e = dalloc(c, &c->d_pj_params, 1);
  if (e == MCPET_OK) e = dalloc(c, &c->d_pj_out, 3 * n + 1);   // + the maximum's word
  if (e != MCPET_OK) return e;
  c->pj_alloc = n;
  // the geometry, materials and voxels of the create-time parameter block (runs do not change them)
  HIPCHK(c, hipMemcpy(c->d_pj_params, &c->P, sizeof(KParams), hipMemcpyHostToDevice));
  // The (activity, mu) map is needed only while the projection runs: one per device, grown to
  // the largest phantom projected and kept until the process ends; projections take turns on
  // it (the lock is held until the stream has finished with it).
  static std::mutex map_mu;
  static std::map<int, std::pair<float2*, size_t>> maps;
  std::lock_guard<std::mutex> lock(map_mu);
  auto& m = maps[c->device];
  if (m.second < (size_t)c->nvox) {
    if (m.first) (void)hipFree(m.first);
    m = {nullptr, 0};
    HIPCHK(c, hipMalloc((void**)&m.first, sizeof(float2) * (size_t)c->nvox));
    m.second = (size_t)c->nvox;
  }
  const int gz = (c->cfg.nzs + D - 1) / D;
  const float step = 0.5f / std::max(c->P.ivx, c->P.ivy);   // half the smaller transaxial voxel side
  hipError_t he = mcpet_launch_project(c->d_pj_params, c->d_active_vox, c->d_active_act, c->n_active, m.first,
                                       (size_t)c->nvox, l.lors, l.off, D, H, W, gz, step, c->d_pj_out, c->d_pj_out + 2 * n,
                                       c->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(c->d_pj_out + n, c->d_pj_out, sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream);
  if (he == hipSuccess)
    he = mcpet_launch_scale_by_max(c->d_pj_out + n, (int)n, reinterpret_cast<unsigned*>(c->d_pj_out + 3 * n), c->stream);
  const hipError_t se = hipStreamSynchronize(c->stream);   // also on failure: the map is released after this
  if (he == hipSuccess) he = se;
  if (he != hipSuccess) return set_err(c, MCPET_E_HIP, std::string("projection failed: ") + hipGetErrorString(he));
  c->pj_dims[0] = D;
  c->pj_dims[1] = H;
  c->pj_dims[2] = W;
  *d_maps = c->d_pj_out;
  return MCPET_OK;
}

static mcpet_status project_maps(mcpet_ctx* c, int D, int H, int W, float* act, float* att)
{
  const float* d = nullptr;
  mcpet_status st = mcpet_ctx_project_device(c, D, H, W, &d);
  if (st != MCPET_OK) return st;
  const size_t n = (size_t)D * H * W;
  if (act) HIPCHK(c, hipMemcpy(act, d, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (att) HIPCHK(c, hipMemcpy(att, d + 2 * n, sizeof(float) * n, hipMemcpyDeviceToHost));
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_project_maps(mcpet_ctx* c, int32_t D, int32_t H, int32_t W, float* act, float* att)
{
  if (!c) return set_err(c, MCPET_E_ARG, "mcpet_project_maps: null context");
  return project_maps(c, D, H, W, act, att);
}

extern "C" mcpet_status mcpet_tally_rebinned(mcpet_ctx* c, int32_t D, int32_t H, int32_t W)
{
  if (!c) return set_err(nullptr, MCPET_E_ARG, "mcpet_tally_rebinned: null context");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  KParams& P = c->P;
  if (D == 0 && H == 0 && W == 0) {
    P.rb_trues = P.rb_scatter = nullptr;
    P.rb_plane = nullptr;
    return MCPET_OK;
  }
  const mcpet_config& cfg = c->cfg;
  if (D < 1 || H < 1 || W < 1 || (long long)D * H * W > (1LL << 28))
    return set_err(c, MCPET_E_ARG, "mcpet_tally_rebinned: bad grid");
  if (c->singles_mode)
    return set_err(c, MCPET_E_STATE, "mcpet_tally_rebinned: the singles pipeline bins its coincidences after the transport");
  if (c->rbt_dims[0] != D || c->rbt_dims[1] != H || c->rbt_dims[2] != W) {
    // plane group of every sinogram as the rebinning's: plane / ceil(NZS / D)
    std::vector<int> plane((size_t)cfg.nsinos);
    if (mcpet_sinogram_planes(&cfg, plane.data()) != MCPET_OK) return set_err(c, MCPET_E_ARG, mcpet_host_last_error());
    const int gz = (cfg.nzs + D - 1) / D;
    for (int& v : plane) v /= gz;
    for (void* p : {(void*)c->d_rbt, (void*)c->d_rbt_plane})
      if (p) (void)hipFree(p);
    if (c->d_rbt) c->dev_bytes -= 2 * sizeof(int) * (size_t)c->rbt_dims[0] * c->rbt_dims[1] * c->rbt_dims[2];
    if (c->d_rbt_plane) c->dev_bytes -= sizeof(int) * (size_t)cfg.nsinos;
    c->d_rbt = c->d_rbt_plane = nullptr;
    c->rbt_dims[0] = c->rbt_dims[1] = c->rbt_dims[2] = 0;
    P.rb_trues = P.rb_scatter = nullptr;
    P.rb_plane = nullptr;
    const size_t n = (size_t)D * H * W;
    mcpet_status e = dalloc(c, &c->d_rbt, 2 * n);
    if (e == MCPET_OK) e = dalloc(c, &c->d_rbt_plane, plane.size());
    if (e != MCPET_OK) return e;
    HIPCHK(c, hipMemcpy(c->d_rbt_plane, plane.data(), sizeof(int) * plane.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->d_rbt, 0, 2 * sizeof(int) * n));
    c->rbt_dims[0] = D;
    c->rbt_dims[1] = H;
    c->rbt_dims[2] = W;
  }
  const size_t n = (size_t)D * H * W;
  P.rb_trues = c->d_rbt;
  P.rb_scatter = c->d_rbt + n;
  P.rb_plane = c->d_rbt_plane;
  P.rb_H = H;
  P.rb_W = W;
  P.rb_gh = (cfg.nangles + H - 1) / H;
  P.rb_gw = (cfg.nrad + W - 1) / W;
  c->rbt_ran = false;   // the next run fills them
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_rebin_sinogram(mcpet_ctx* c, int32_t which, int32_t D, int32_t H, int32_t W, float* out)
{
  return rebin(c, which, D, H, W, out);
}

extern "C" void mcpet_destroy(mcpet_ctx* c)
{
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
}

extern "C" mcpet_status mcpet_stream_seeds(int32_t seed, const int64_t* ids, int32_t n, int32_t* sx, int32_t* sy)
{
  if (!seed_usable(seed)) return set_err(nullptr, MCPET_E_ARG, seed_msg(seed));
  if (!ids || !sx || !sy || n < 0) return set_err(nullptr, MCPET_E_ARG, "mcpet_stream_seeds: bad args");
  uint32_t j1[MCR_JUMP_BITS], j2[MCR_JUMP_BITS];
  mcr_build_jump_table(MCR_A1, MCR_M1, j1);
  mcr_build_jump_table(MCR_A2, MCR_M2, j2);
  for (int32_t i = 0; i < n; i++) mcr_init_stream(ids[i], seed, j1, j2, &sx[i], &sy[i]);
  return MCPET_OK;
}

// ------------------------------------------------------------------------------ self tests

extern "C" mcpet_status mcpet_selftest_ranecu(int32_t seed, const int64_t* ids, int32_t nstreams, int32_t n, float* out,
                                              double* out_d, int32_t as_double)
{
  if (!ids || nstreams < 1 || n < 1 || (as_double ? !out_d : !out)) return set_err(nullptr, MCPET_E_ARG, "selftest_ranecu: bad args");
  if (!seed_usable(seed)) return set_err(nullptr, MCPET_E_ARG, seed_msg(seed));
  uint32_t j[2][MCR_JUMP_BITS];
  mcr_build_jump_table(MCR_A1, MCR_M1, j[0]);
  mcr_build_jump_table(MCR_A2, MCR_M2, j[1]);
  long long* d_ids = nullptr;
  uint32_t* d_j = nullptr;
  void* d_out = nullptr;
  const size_t elem = as_double ? 8 : 4;
  const size_t bytes = (size_t)nstreams * n * elem;
  mcpet_status rc = MCPET_OK;
  hipError_t e = hipMalloc((void**)&d_ids, sizeof(long long) * nstreams);
  if (e == hipSuccess) e = hipMalloc((void**)&d_j, sizeof j);
  if (e == hipSuccess) e = hipMalloc(&d_out, bytes);
  if (e == hipSuccess) e = hipMemcpy(d_ids, ids, sizeof(long long) * nstreams, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_j, j, sizeof j, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = mcpet_launch_selftest_ranecu(seed, d_ids, nstreams, n, (float*)d_out, (double*)d_out, as_double, d_j, d_j + MCR_JUMP_BITS, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(as_double ? (void*)out_d : (void*)out, d_out, bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = set_err(nullptr, MCPET_E_HIP, std::string("selftest_ranecu: ") + hipGetErrorString(e));
  if (d_ids) (void)hipFree(d_ids);
  if (d_j) (void)hipFree(d_j);
  if (d_out) (void)hipFree(d_out);
  return rc;
}

extern "C" mcpet_status mcpet_selftest_fmath(const double* x, const double* y, int32_t n, double* od, float* of)
{
  if (!x || !y || !od || !of || n < 1) return set_err(nullptr, MCPET_E_ARG, "selftest_fmath: bad args");
  double *dx = nullptr, *dy = nullptr, *dod = nullptr;
  float* dof = nullptr;
  mcpet_status rc = MCPET_OK;
  hipError_t e = hipMalloc((void**)&dx, 8 * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void**)&dy, 8 * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void**)&dod, 7 * 8 * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void**)&dof, 8 * 4 * (size_t)n);
  if (e == hipSuccess) e = hipMemcpy(dx, x, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dy, y, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mcpet_launch_selftest_fmath(dx, dy, n, dod, dof, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(od, dod, 7 * 8 * (size_t)n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(of, dof, 8 * 4 * (size_t)n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = set_err(nullptr, MCPET_E_HIP, std::string("selftest_fmath: ") + hipGetErrorString(e));
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (dod) (void)hipFree(dod);
  if (dof) (void)hipFree(dof);
  return rc;
}

extern "C" mcpet_status mcpet_selftest_sort(const uint64_t* keys, int64_t n, int32_t passes, uint32_t* perm)
{
  if (!keys || !perm || n < 0 || n > 0x7fffffffLL || passes < 1 || passes > 8)
    return set_err(nullptr, MCPET_E_ARG, "selftest_sort: bad args");
  if (n == 0) return MCPET_OK;
  unsigned long long* k[2] = {nullptr, nullptr};
  unsigned* v[2] = {nullptr, nullptr};
  unsigned* hist = nullptr;
  const long long m = 256LL * mcpet_rs_tiles(n);
  std::vector<long long> levels;
  mcpet_scan_levels(m, levels);
  std::vector<unsigned*> scan(levels.size(), nullptr);
  hipError_t e = hipSuccess;
  for (int i = 0; i < 2 && e == hipSuccess; i++) {
    e = hipMalloc((void**)&k[i], 8 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void**)&v[i], 4 * (size_t)n);
  }
  if (e == hipSuccess) e = hipMalloc((void**)&hist, 4 * (size_t)m);
  for (size_t i = 0; i < levels.size() && e == hipSuccess; i++) e = hipMalloc((void**)&scan[i], 4 * (size_t)levels[i]);
  int which = 0;
  if (e == hipSuccess) e = hipMemcpy(k[0], keys, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mcpet_launch_iota(v[0], n, 0);
  if (e == hipSuccess) e = mcpet_radix_sort(n, passes, k, v, hist, scan.data(), &which, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(perm, v[which], 4 * (size_t)n, hipMemcpyDeviceToHost);
  mcpet_status rc = MCPET_OK;
  if (e != hipSuccess) rc = set_err(nullptr, MCPET_E_HIP, std::string("selftest_sort: ") + hipGetErrorString(e));
  for (int i = 0; i < 2; i++) {
    if (k[i]) (void)hipFree(k[i]);
    if (v[i]) (void)hipFree(v[i]);
  }
  if (hist) (void)hipFree(hist);
  for (unsigned* p : scan)
    if (p) (void)hipFree(p);
  return rc;
}

extern "C" mcpet_status mcpet_selftest_copy(size_t nbytes, int32_t iters, double* ms_out)
{
  if (nbytes < 16 || iters < 1) return set_err(nullptr, MCPET_E_ARG, "selftest_copy: bad args");
  size_t n16 = nbytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  mcpet_status rc = MCPET_OK;
  hipError_t e = hipMalloc((void**)&a, n16 * 16);
  if (e == hipSuccess) e = hipMalloc((void**)&b, n16 * 16);
  if (e == hipSuccess) e = hipMemset(a, 1, n16 * 16);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = mcpet_launch_copy(a, b, n16, 0);
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  for (int i = 0; i < iters && e == hipSuccess; i++) e = mcpet_launch_copy(a, b, n16, 0);
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (ms_out) *ms_out = ms / iters;
  if (e != hipSuccess) rc = set_err(nullptr, MCPET_E_HIP, std::string("selftest_copy: ") + hipGetErrorString(e));
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

// ---------------------------------------------------------------------------------- internals for the server

size_t mcpet_ctx_device_bytes(const mcpet_ctx* c) { return c->dev_bytes; }
long long mcpet_ctx_streams(const mcpet_ctx* c) { return c->P.n_streams; }
int mcpet_ctx_device(const mcpet_ctx* c) { return c->device; }
int mcpet_ctx_num_cus(const mcpet_ctx* c) { return c->num_cus; }

mcpet_status mcpet_run_contexts(const CtxRun* runs, int n, hipStream_t stream, KParams* h_params, KParams* d_params,
                                int block_size, int blocks, std::string* err)
{
  auto fail = [&](mcpet_status st, const std::string& m) {
    if (err) *err = m;
    return set_err(nullptr, st, m);
  };
  bool ext = false;
  for (int k = 0; k < n; k++) {
    mcpet_ctx* c = runs[k].c;
    KParams P = c->P;
    if (runs[k].seed != 0) P.seed = runs[k].seed;
    if (!seed_usable(P.seed)) return fail(MCPET_E_ARG, seed_msg(P.seed));
    P.stream_base = runs[k].begin;
    P.n_streams = runs[k].end;
    P.sm_policy = 0;
    P.lds_spectrum = 0;
    P.lds_palette = 0;
    ext = ext || P.prompt_frac > 0.0f;
    h_params[k] = P;
    c->rbt_ran = P.rb_trues != nullptr && (runs[k].clear != 0 || c->rbt_ran);
    mcpet_status st = clear_buffers(c, stream, runs[k].clear != 0, runs[k].clear != 0);
    if (st != MCPET_OK) return fail(st, c->err);
    hipError_t e = hipMemsetAsync(c->d_counters + KC_N, 0, sizeof(unsigned long long), stream);   // stream queue
    if (e != hipSuccess) return fail(MCPET_E_HIP, std::string("queue reset: ") + hipGetErrorString(e));
    c->ran = true;
  }
  hipError_t e = hipMemcpyAsync(d_params, h_params, sizeof(KParams) * (size_t)n, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = mcpet_launch_transport_batch(d_params, n, ext, 0, blocks, block_size, stream);
  for (int k = 0; k < n && e == hipSuccess; k++) e = hipEventRecord(runs[k].c->ev_done, stream);
  if (e != hipSuccess) return fail(MCPET_E_HIP, std::string("batched launch: ") + hipGetErrorString(e));
  return MCPET_OK;
}

// ---------------------------------------------------------------------------------- batches

struct mcpet_batch {
  std::vector<mcpet_ctx*> ctx;
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  KParams* d_params = nullptr;
  KParams* h_params = nullptr;   // pinned, n entries
  int tuned_bs = 0, tuned_per_cu = 0, tuned_policy = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_key[3] = {-1, -1, -1};  // block size, blocks, count_events of the captured launch
  bool ran = false;
  std::string err;
};

namespace {

mcpet_status bset_err(mcpet_batch* b, mcpet_status st, const std::string& m)
{
  if (b) b->err = m;
  return set_err(nullptr, st, m);
}

void batch_free(mcpet_batch* b)
{
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (mcpet_ctx* c : b->ctx) mcpet_destroy(c);
  if (b->d_params) (void)hipFree(b->d_params);
  if (b->h_params) (void)hipHostFree(b->h_params);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

#define BHIP(b, call)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess) return bset_err(b, MCPET_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" mcpet_status mcpet_batch_create(const mcpet_config* cfg, const mcpet_phantom* phantoms, int32_t n,
                                           const mcpet_tables* tab, mcpet_batch** out)
{
  if (!out) return set_err(nullptr, MCPET_E_ARG, "mcpet_batch_create: null output pointer");
  *out = nullptr;
  if (!cfg || !phantoms || !tab || n < 1) return set_err(nullptr, MCPET_E_ARG, "mcpet_batch_create: bad arguments");
  if (cfg->detector.tally_psf_sinogram == MCPET_TALLY_SINGLES)
    return set_err(nullptr, MCPET_E_ARG, "mcpet_batch_create: batches support the pair tallies (modes 0-2) only");
  mcpet_batch* b = new mcpet_batch();
  for (int32_t k = 0; k < n; k++) {
    mcpet_ctx* c = nullptr;
    mcpet_status st = mcpet_create(cfg, &phantoms[k], tab, &c);
    if (st != MCPET_OK) {
      std::string m = std::string("mcpet_batch_create: phantom ") + std::to_string(k) + ": " + mcpet_last_error(nullptr);
      batch_free(b);
      delete b;
      return set_err(nullptr, st, m);
    }
    b->ctx.push_back(c);
  }
  b->device = b->ctx[0]->device;
  b->num_cus = b->ctx[0]->num_cus;
  hipError_t e = hipSetDevice(b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_params, sizeof(KParams) * (size_t)n);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_params, sizeof(KParams) * (size_t)n, hipHostMallocDefault);
  if (e != hipSuccess) {
    std::string m = std::string("mcpet_batch_create: ") + hipGetErrorString(e);
    batch_free(b);
    delete b;
    return set_err(nullptr, MCPET_E_HIP, m);
  }
  *out = b;
  return MCPET_OK;
}

static mcpet_status batch_run(mcpet_batch* b, const mcpet_run_params* rp, double stream_fraction);

extern "C" mcpet_status mcpet_batch_run(mcpet_batch* b, const mcpet_run_params* rp) { return batch_run(b, rp, 1.0); }

// stream_fraction < 1: simulate only the first streams of every phantom (the autotuner's
// sample: full acquisition per stream, so the per-stream work is representative).
static mcpet_status batch_run(mcpet_batch* b, const mcpet_run_params* rp, double stream_fraction)
{
  if (!b) return set_err(nullptr, MCPET_E_ARG, "mcpet_batch_run: null batch");
  mcpet_run_params p{};
  p.zero_tallies = 1;
  if (rp) p = *rp;
  BHIP(b, hipSetDevice(b->device));
  BHIP(b, hipStreamSynchronize(b->stream));   // the staging copy of the parameters is reused
  const int n = (int)b->ctx.size();
  bool ext = false;
  for (int k = 0; k < n; k++) {
    KParams P = b->ctx[k]->P;
    if (p.seed != 0) P.seed = p.seed;
    if (!seed_usable(P.seed)) return bset_err(b, MCPET_E_ARG, seed_msg(P.seed));
    if (p.acquisition_time_ps != 0) P.acq_ps = p.acquisition_time_ps;
    if (stream_fraction < 1.0) P.n_streams = (long long)std::ceil((double)P.n_streams * stream_fraction);
    const int policy = p.phase_policy > 0 ? p.phase_policy : (b->tuned_policy > 0 ? b->tuned_policy : 1);
    P.sm_policy = (policy == 2) ? 1 : 0;
    static const int default_cost[MCPET_NPHASES] = {4, 6, 1, 3, 6, 3, 2, 4, 4, 1};
    for (int i = 0; i < MCPET_NPHASES; i++) P.sm_cost[i] = p.phase_cost[i] > 0 ? p.phase_cost[i] : default_cost[i];
    P.lds_spectrum = 0;   // per phantom: global atomics
    P.lds_palette = 0;
    ext = ext || P.prompt_frac > 0.0f;
    b->h_params[k] = P;
  }
  const int bs = p.block_size > 0 ? p.block_size : (b->tuned_bs > 0 ? b->tuned_bs : 256);
  if (bs % 64 != 0 || bs > 256) return bset_err(b, MCPET_E_RANGE, "mcpet_batch_run: block_size must be a multiple of 64, <= 256");
  const int per_cu = p.blocks_per_cu > 0 ? p.blocks_per_cu : (b->tuned_per_cu > 0 ? b->tuned_per_cu : 8);
  const int blocks = b->num_cus * per_cu;
  BHIP(b, hipMemcpyAsync(b->d_params, b->h_params, sizeof(KParams) * (size_t)n, hipMemcpyHostToDevice, b->stream));
  if (p.use_graph) {
    const int key[3] = {bs, blocks, p.count_events ? 1 : 0};
    if (!b->gexec || std::memcmp(key, b->graph_key, sizeof key) != 0 || !p.zero_tallies) {
      if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->gexec = nullptr;
      b->graph = nullptr;
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      mcpet_status st = MCPET_OK;
      for (int k = 0; k < n && st == MCPET_OK; k++) st = clear_buffers(b->ctx[k], b->stream, p.zero_tallies != 0, true);
      hipError_t e = hipSuccess;
      if (st == MCPET_OK) e = mcpet_launch_transport_batch(b->d_params, n, ext, p.count_events, blocks, bs, b->stream);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(b->stream, &g);
      if ((st != MCPET_OK || e != hipSuccess || e2 != hipSuccess) && g) (void)hipGraphDestroy(g);
      if (st != MCPET_OK) return st;
      if (e != hipSuccess || e2 != hipSuccess)
        return bset_err(b, MCPET_E_HIP, std::string("mcpet_batch_run: graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
      b->graph = g;
      BHIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      std::memcpy(b->graph_key, key, sizeof key);
    }
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    BHIP(b, hipGraphLaunch(b->gexec, b->stream));
    BHIP(b, hipEventRecord(b->ev1, b->stream));
  } else {
    for (int k = 0; k < n; k++) {
      mcpet_status st = clear_buffers(b->ctx[k], b->stream, p.zero_tallies != 0, true);
      if (st != MCPET_OK) return st;
    }
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    hipError_t e = mcpet_launch_transport_batch(b->d_params, n, ext, p.count_events, blocks, bs, b->stream);
    if (e != hipSuccess) return bset_err(b, MCPET_E_HIP, std::string("batch launch failed: ") + hipGetErrorString(e));
    BHIP(b, hipEventRecord(b->ev1, b->stream));
  }
  for (mcpet_ctx* c : b->ctx) {
    BHIP(b, hipEventRecord(c->ev_done, b->stream));
    c->ran = true;
    c->rbt_ran = c->P.rb_trues != nullptr && (p.zero_tallies || c->rbt_ran);
  }
  b->ran = true;
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_batch_sync(mcpet_batch* b)
{
  if (!b) return set_err(nullptr, MCPET_E_ARG, "mcpet_batch_sync: null batch");
  BHIP(b, hipSetDevice(b->device));
  BHIP(b, hipStreamSynchronize(b->stream));
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_batch_fetch(mcpet_batch* b, int32_t k, mcpet_tallies* out)
{
  if (!b || k < 0 || k >= (int32_t)b->ctx.size()) return bset_err(b, MCPET_E_ARG, "mcpet_batch_fetch: bad phantom index");
  if (!b->ran) return bset_err(b, MCPET_E_STATE, "mcpet_batch_fetch: no run");
  mcpet_status st = mcpet_batch_sync(b);
  if (st != MCPET_OK) return st;
  st = mcpet_fetch(b->ctx[k], out);
  if (st != MCPET_OK) b->err = b->ctx[k]->err;
  return st;
}

extern "C" mcpet_status mcpet_batch_rebin_sinogram(mcpet_batch* b, int32_t k, int32_t which, int32_t D, int32_t H, int32_t W,
                                                   float* out)
{
  if (!b || k < 0 || k >= (int32_t)b->ctx.size()) return bset_err(b, MCPET_E_ARG, "mcpet_batch_rebin_sinogram: bad phantom index");
  if (!b->ran) return bset_err(b, MCPET_E_STATE, "mcpet_batch_rebin_sinogram: no run");
  mcpet_status st = mcpet_batch_sync(b);
  if (st == MCPET_OK) st = rebin(b->ctx[k], which, D, H, W, out);
  if (st != MCPET_OK) b->err = b->ctx[k]->err;
  return st;
}

extern "C" mcpet_status mcpet_batch_project_maps(mcpet_batch* b, int32_t k, int32_t D, int32_t H, int32_t W, float* act,
                                                 float* att)
{
  if (!b || k < 0 || k >= (int32_t)b->ctx.size()) return bset_err(b, MCPET_E_ARG, "mcpet_batch_project_maps: bad phantom index");
  mcpet_status st = mcpet_batch_sync(b);
  if (st == MCPET_OK) st = project_maps(b->ctx[k], D, H, W, act, att);
  if (st != MCPET_OK) b->err = b->ctx[k]->err;
  return st;
}

extern "C" mcpet_status mcpet_batch_get_stats(mcpet_batch* b, int32_t k, mcpet_stats* s)
{
  if (!b || !s || k < 0 || k >= (int32_t)b->ctx.size()) return bset_err(b, MCPET_E_ARG, "mcpet_batch_get_stats: bad arguments");
  if (!b->ran) return bset_err(b, MCPET_E_STATE, "mcpet_batch_get_stats: no run");
  mcpet_status st = mcpet_batch_sync(b);
  if (st != MCPET_OK) return st;
  std::memset(s, 0, sizeof *s);
  float ms = 0.0f;
  BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  s->kernel_ms = ms;
  return read_counts(b->ctx[k], s);
}

extern "C" mcpet_status mcpet_batch_autotune(mcpet_batch* b, double fraction, mcpet_tune_result* results, int32_t max_results,
                                             int32_t* n_results)
{
  if (!b || !(fraction > 0.0 && fraction <= 1.0)) return bset_err(b, MCPET_E_ARG, "mcpet_batch_autotune: bad arguments");
  // launch shape x phase-selection policy (the tallies do not depend on either)
  static const int shapes[][2] = {{256, 3}, {256, 4}, {256, 6}, {256, 8}, {128, 6}, {128, 8}, {128, 12}, {64, 12}, {64, 16}};
  const int ns = (int)(sizeof shapes / sizeof shapes[0]);
  double best = -1.0;
  int best_i = 0, best_pol = 1, written = 0;
  mcpet_run_params p{};
  p.zero_tallies = 1;
  for (int t = 0; t < 2 * ns; t++) {
    const int i = t % ns, pol = 1 + t / ns;
    p.block_size = shapes[i][0];
    p.blocks_per_cu = shapes[i][1];
    p.phase_policy = pol;
    mcpet_status st = batch_run(b, &p, fraction);
    if (st == MCPET_OK) st = mcpet_batch_sync(b);
    if (st != MCPET_OK) return st;
    float ms = 0.0f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    unsigned long long decays = 0;
    for (mcpet_ctx* c : b->ctx) {
      unsigned long long d = 0;
      BHIP(b, hipMemcpy(&d, c->d_counters + KC_DECAYS, sizeof d, hipMemcpyDeviceToHost));
      decays += d;
    }
    const double rate = ms > 0.0f ? (double)decays / ms : 0.0;
    if (results && written < max_results)
      results[written++] = {shapes[i][0], shapes[i][1], pol, (double)ms, rate, MCPET_LDS_XS511, b->ctx[0]->cfg.stream_decays};
    if (rate > best) {
      best = rate;
      best_i = i;
      best_pol = pol;
    }
  }
  b->tuned_bs = shapes[best_i][0];
  b->tuned_per_cu = shapes[best_i][1];
  b->tuned_policy = best_pol;
  if (n_results) *n_results = written;
  return MCPET_OK;
}

extern "C" void mcpet_batch_destroy(mcpet_batch* b)
{
  if (!b) return;
  (void)hipSetDevice(b->device);
  batch_free(b);
  delete b;
}

extern "C" const char* mcpet_batch_last_error(const mcpet_batch* b)
{
  if (b && !b->err.empty()) return b->err.c_str();
  return mcpet_last_error(nullptr);
}
