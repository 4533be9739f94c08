// server.cpp — serving loop over the batched transport: continuous batching of simulation
// requests under an HBM budget (no reference counterpart; the reference runs one
// simulation per process).
//
// A request's decay streams are simulated in chunks: each step runs, in one batched launch,
// the next `streams_per_step` streams of every admitted request. Because every stream has a
// fixed id (its skip-ahead RANECU sequence) and tallies are integer sums, the chunking and
// the company a request keeps do not change its result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "host_internal.h"
#include "runtime_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  mcpet_config cfg{};
  mcpet_phantom ph{};
  std::vector<float> vmd;            // the request's own copy of the voxels
  int32_t seed = 0;
  int state = MCPET_REQ_QUEUED;
  mcpet_status fail_status = MCPET_OK;   // a rejected or failed request: why
  std::string fail_msg;
  size_t footprint = 0;              // device bytes reserved at admission
  mcpet_ctx* ctx = nullptr;
  long long next = 0, streams = 0;
  Clock::time_point submitted, done;
  // results (host) once done
  std::vector<int32_t> trues, scatter, img_t, img_sc, spectrum;
  std::vector<uint64_t> mat_edep, voxel_edep;
  std::vector<mcpet_psf_record> psf;
  mcpet_tallies t{};
  int grid[3] = {0, 0, 0};           // rebinned request: keep only these grids
  int maps = 0;                      // and the line integrals of its phantom on that grid
  std::vector<float> rb_trues, rb_scatter, rb_act, rb_att;
  mcpet_surrogate* net = nullptr;    // estimate request: the network that answers it
  std::vector<float> est;            // its scatter estimate per decay (rb_scatter: the counts)
  int64_t est_decays = 0;
};

// Upper bound of the device memory mcpet_create allocates for (cfg, phantom, tables).
size_t footprint(const mcpet_config& cfg, const mcpet_phantom& ph, const mcpet_tables& tab)
{
  const size_t nvox = (size_t)ph.nx * ph.ny * ph.nz;
  const size_t N = (size_t)tab.num_values;
  size_t b = 0;
  b += 2 * 4 * (size_t)cfg.nbins;                       // trues, scatter
  b += 2 * 4 * nvox;                                    // emission images
  b += 8 * nvox + 512 * 1024;                           // voxel words (or indices + palette)
  b += 8 * nvox;                                        // emitting voxels and activities
  b += 20 * nvox + 32;                                  // balanced stream layout: offsets, ids, owner hints
  b += N * 8 + N * 15 * 8 + N * 15 * 16;                // Woodcock, total, Compton/Rayleigh tables
  b += sizeof(mcpet_rayleigh_table) + sizeof(mcpet_compton_table);
  b += 4 * (size_t)std::max(1, cfg.ne);
  const int mode = cfg.detector.tally_psf_sinogram;
  if ((mode == 0 || mode == 1) && cfg.detector.psf_size > 0) b += sizeof(mcpet_psf_record) * (size_t)cfg.detector.psf_size;
  if (cfg.voxel_dose && cfg.roi_x_max > -1)
    b += 16 * (size_t)(cfg.roi_x_max - cfg.roi_x_min + 1) * (cfg.roi_y_max - cfg.roi_y_min + 1) *
         (cfg.roi_z_max - cfg.roi_z_min + 1);
  b += 64 * 1024;                                       // counters, parameter block, small tallies
  return b;
}

double percentile(std::vector<double> v, double q)
{
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

struct mcpet_server {
  const mcpet_tables* tab = nullptr;
  size_t budget = 0;
  int max_running = 0;
  long long per_step = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  KParams* h_params = nullptr;
  KParams* d_params = nullptr;
  int64_t next_id = 1;
  std::map<int64_t, std::unique_ptr<Request>> req;
  std::deque<int64_t> queue;
  std::vector<int64_t> running;
  size_t admitted = 0;
  mcpet_server_stats st{};
  std::vector<double> latencies;
  std::string err;
};

static mcpet_status sfail(mcpet_server* s, mcpet_status st, const std::string& m)
{
  if (s) s->err = m;
  mcpet_host_set_error(m.c_str());
  return st;
}

extern "C" mcpet_status mcpet_server_create(const mcpet_tables* tab, size_t budget, int32_t max_running, int64_t per_step,
                                            mcpet_server** out)
{
  if (!out) return sfail(nullptr, MCPET_E_ARG, "mcpet_server_create: null output pointer");
  *out = nullptr;
  if (!tab || budget == 0 || max_running < 1 || per_step < 1)
    return sfail(nullptr, MCPET_E_ARG, "mcpet_server_create: bad arguments");
  mcpet_server* s = new mcpet_server();
  s->tab = tab;
  s->budget = budget;
  s->max_running = max_running;
  s->per_step = per_step;
  s->st.budget_bytes = budget;
  hipError_t e = hipGetDevice(&s->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_params, sizeof(KParams) * (size_t)max_running, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_params, sizeof(KParams) * (size_t)max_running);
  if (e != hipSuccess) {
    std::string m = std::string("mcpet_server_create: ") + hipGetErrorString(e);
    mcpet_server_destroy(s);
    return sfail(nullptr, MCPET_E_HIP, m);
  }
  *out = s;
  return MCPET_OK;
}

// grid: NULL (a full request) or {D, H, W, maps} of a rebinned request; net: the network that
// answers an estimate request (rebinned on its grid).
static mcpet_status submit(mcpet_server* s, const mcpet_config* cfg, const mcpet_phantom* ph, int32_t seed, const int* grid,
                           int64_t* request_id, mcpet_surrogate* net = nullptr)
{
  if (!s || !cfg || !ph || !ph->voxel_mat_dens) return sfail(s, MCPET_E_ARG, "mcpet_server_submit: null argument");
  if (cfg->detector.tally_psf_sinogram == MCPET_TALLY_SINGLES)
    return sfail(s, MCPET_E_ARG, "mcpet_server_submit: requests use the pair tallies (modes 0-2)");
  std::unique_ptr<Request> r(new Request());
  r->id = s->next_id++;
  r->cfg = *cfg;
  r->cfg.gpu_id = s->device;
  r->ph = *ph;
  const size_t nvox = (size_t)ph->nx * ph->ny * ph->nz;
  r->vmd.assign(ph->voxel_mat_dens, ph->voxel_mat_dens + 3 * nvox);
  r->ph.voxel_mat_dens = r->vmd.data();
  r->ph.owned = 0;
  r->seed = seed;
  r->net = net;
  // what mcpet_create and the run would refuse is refused now, so the request never
  // reaches the head of the queue (where a failing creation would be retried every step)
  mcpet_status vst = mcpet_check_create_args(&r->cfg, &r->ph, s->tab, true);
  std::string vmsg = vst != MCPET_OK ? std::string(mcpet_last_error(nullptr)) : std::string();
  if (vst == MCPET_OK && !mcpet_seed_usable(seed != 0 ? seed : r->cfg.seed)) {
    vst = MCPET_E_ARG;
    vmsg = "mcpet_server_submit: unusable seed (0 or below -2147483399)";
  }
  if (grid) {
    std::memcpy(r->grid, grid, sizeof r->grid);
    r->maps = grid[3];
    const size_t n = (size_t)grid[0] * grid[1] * grid[2];
    r->footprint = footprint(r->cfg, r->ph, *s->tab) + 4 * (n + grid[0] + 1) +
                   4 * (size_t)r->cfg.nsinos;   // the rebinning's plane lists and output
    if (r->maps)   // the device's (activity, mu) map, crystal-pair lists, the maps, parameter block
      r->footprint += 8 * nvox + 4 * (size_t)r->cfg.ncrystals * r->cfg.ncrystals + 4 * ((size_t)grid[1] * grid[2] + 1) +
                      12 * n + 4 + 64 * 1024;
  } else {
    r->footprint = footprint(r->cfg, r->ph, *s->tab);
  }
  r->submitted = Clock::now();
  s->st.submitted++;
  if (request_id) *request_id = r->id;
  if (vst == MCPET_OK && r->footprint > s->budget) {
    vst = MCPET_E_CAPACITY;
    vmsg = "mcpet_server_submit: the request needs more device memory than the budget";
  }
  if (vst != MCPET_OK) {
    r->state = MCPET_REQ_REJECTED;
    r->fail_status = vst;
    r->fail_msg = vmsg;
    s->st.rejected++;
    s->req[r->id] = std::move(r);
    return sfail(s, vst, vmsg);
  }
  s->queue.push_back(r->id);
  s->req[r->id] = std::move(r);
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_submit(mcpet_server* s, const mcpet_config* cfg, const mcpet_phantom* ph, int32_t seed,
                                            int64_t* request_id)
{
  return submit(s, cfg, ph, seed, nullptr, request_id);
}

extern "C" mcpet_status mcpet_server_submit_rebinned(mcpet_server* s, const mcpet_config* cfg, const mcpet_phantom* ph,
                                                     int32_t seed, int32_t D, int32_t H, int32_t W, int32_t with_maps,
                                                     int64_t* request_id)
{
  if (D < 1 || H < 1 || W < 1 || W > 256 || (with_maps && D > 256))
    return sfail(s, MCPET_E_ARG, "mcpet_server_submit_rebinned: bad grid (W <= 256; D <= 256 with maps)");
  const int grid[4] = {D, H, W, with_maps ? 1 : 0};
  return submit(s, cfg, ph, seed, grid, request_id);
}

extern "C" mcpet_status mcpet_server_submit_estimate(mcpet_server* s, const mcpet_config* cfg, const mcpet_phantom* ph,
                                                     int32_t seed, mcpet_surrogate* net, int64_t* request_id)
{
  if (!net) return sfail(s, MCPET_E_ARG, "mcpet_server_submit_estimate: null network");
  int D = 0, H = 0, W = 0, cin = 0;
  mcpet_surrogate_dims(net, &D, &H, &W, &cin);
  if (cin != 2 && cin != 4) return sfail(s, MCPET_E_ARG, "mcpet_server_submit_estimate: the network must take 2 or 4 input channels");
  if (W > 256 || D > 256) return sfail(s, MCPET_E_ARG, "mcpet_server_submit_estimate: network grid too large (D, W <= 256)");
  const int grid[4] = {D, H, W, cin == 4 ? 1 : 0};
  return submit(s, cfg, ph, seed, grid, request_id, net);
}

static void complete(mcpet_server* s, Request& r)
{
  mcpet_destroy(r.ctx);
  r.ctx = nullptr;
  s->admitted -= r.footprint;
  r.state = MCPET_REQ_DONE;
  r.done = Clock::now();
  s->latencies.push_back(std::chrono::duration<double, std::milli>(r.done - r.submitted).count());
  s->st.completed++;
}

// A request that could not be created or finished: its context and device memory are released,
// it leaves the queue or the running set, and its fetch reports why.
static void fail_request(mcpet_server* s, Request& r, mcpet_status st, const std::string& why)
{
  if (r.ctx) mcpet_destroy(r.ctx);
  r.ctx = nullptr;
  if (r.state == MCPET_REQ_RUNNING) s->admitted -= r.footprint;
  r.state = MCPET_REQ_FAILED;
  r.fail_status = st;
  r.fail_msg = why;
  r.done = Clock::now();
  s->st.failed++;
  s->err = why;
}

// fetch of a rejected or failed request: its status and message; the request is forgotten.
static bool fetch_refused(mcpet_server* s, std::map<int64_t, std::unique_ptr<Request>>::iterator it, mcpet_status* st)
{
  if (it == s->req.end()) return false;
  const Request& r = *it->second;
  if (r.state != MCPET_REQ_REJECTED && r.state != MCPET_REQ_FAILED) return false;
  const std::string m = "request " + std::to_string(r.id) + (r.state == MCPET_REQ_REJECTED ? " rejected: " : " failed: ") +
                        r.fail_msg;
  *st = sfail(s, r.fail_status, m);
  s->req.erase(it);
  return true;
}

static mcpet_status finish_rebinned(mcpet_server* s, Request& r)
{
  const size_t n = (size_t)r.grid[0] * r.grid[1] * r.grid[2];
  r.rb_trues.resize(n);
  r.rb_scatter.resize(n);
  mcpet_tallies& t = r.t;
  std::memset(&t, 0, sizeof t);
  mcpet_status st = mcpet_fetch(r.ctx, &t);   // counters only (no tally buffers given)
  if (st == MCPET_OK) st = mcpet_rebin_sinogram(r.ctx, 0, r.grid[0], r.grid[1], r.grid[2], r.rb_trues.data());
  if (st == MCPET_OK) st = mcpet_rebin_sinogram(r.ctx, 1, r.grid[0], r.grid[1], r.grid[2], r.rb_scatter.data());
  if (st == MCPET_OK && r.maps) {
    r.rb_act.resize(n);
    r.rb_att.resize(n);
    st = mcpet_project_maps(r.ctx, r.grid[0], r.grid[1], r.grid[2], r.rb_act.data(), r.rb_att.data());
  }
  if (st != MCPET_OK) return sfail(s, st, std::string("request ") + std::to_string(r.id) + ": " + mcpet_last_error(r.ctx));
  complete(s, r);
  return MCPET_OK;
}

// An estimate request: the network's fused device estimate on the request's context.
static mcpet_status finish_estimate(mcpet_server* s, Request& r)
{
  const size_t n = (size_t)r.grid[0] * r.grid[1] * r.grid[2];
  r.est.resize(n);
  r.rb_scatter.resize(n);
  mcpet_tallies& t = r.t;
  std::memset(&t, 0, sizeof t);
  mcpet_status st = mcpet_fetch(r.ctx, &t);   // counters only
  if (st != MCPET_OK) return sfail(s, st, std::string("request ") + std::to_string(r.id) + ": " + mcpet_last_error(r.ctx));
  st = mcpet_surrogate_estimate(r.net, r.ctx, r.est.data(), r.rb_scatter.data(), &r.est_decays);
  if (st != MCPET_OK)
    return sfail(s, st, std::string("request ") + std::to_string(r.id) + ": " + mcpet_surrogate_last_error(r.net));
  complete(s, r);
  return MCPET_OK;
}

static mcpet_status finish(mcpet_server* s, Request& r)
{
  if (r.net) return finish_estimate(s, r);
  if (r.grid[0] > 0) return finish_rebinned(s, r);
  const size_t nvox = (size_t)r.ph.nx * r.ph.ny * r.ph.nz;
  r.trues.resize((size_t)r.cfg.nbins);
  r.scatter.resize((size_t)r.cfg.nbins);
  r.img_t.resize(nvox);
  r.img_sc.resize(nvox);
  r.spectrum.resize((size_t)std::max(1, r.cfg.ne));
  r.mat_edep.resize(2 * MCPET_MAX_MATERIALS);
  mcpet_tallies& t = r.t;
  std::memset(&t, 0, sizeof t);
  t.trues = r.trues.data();
  t.scatter = r.scatter.data();
  t.image_trues = r.img_t.data();
  t.image_scatter = r.img_sc.data();
  t.spectrum = r.spectrum.data();
  t.material_edep = r.mat_edep.data();
  if (r.cfg.voxel_dose && r.cfg.roi_x_max > -1) {
    r.voxel_edep.resize(2 * (size_t)(r.cfg.roi_x_max - r.cfg.roi_x_min + 1) * (r.cfg.roi_y_max - r.cfg.roi_y_min + 1) *
                        (r.cfg.roi_z_max - r.cfg.roi_z_min + 1));
    t.voxel_edep = r.voxel_edep.data();
  }
  const int mode = r.cfg.detector.tally_psf_sinogram;
  if ((mode == 0 || mode == 1) && r.cfg.detector.psf_size > 0) {
    r.psf.resize((size_t)r.cfg.detector.psf_size);
    t.psf = r.psf.data();
    t.psf_capacity = r.cfg.detector.psf_size;
  }
  mcpet_status st = mcpet_fetch(r.ctx, &t);
  if (st != MCPET_OK) return sfail(s, st, std::string("request ") + std::to_string(r.id) + ": " + mcpet_last_error(r.ctx));
  r.psf.resize((size_t)t.psf_count);
  complete(s, r);
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_step(mcpet_server* s, int32_t* running, int32_t* completed)
{
  if (!s) return sfail(nullptr, MCPET_E_ARG, "mcpet_server_step: null server");
  const Clock::time_point t0 = Clock::now();
  if (hipSetDevice(s->device) != hipSuccess) return sfail(s, MCPET_E_HIP, "mcpet_server_step: hipSetDevice failed");
  // 1. admission, in submission order, within the budget. The admitted requests' contexts are
  //    created concurrently (a context's creation is host work: voxel packing, stream layout,
  //    uploads), then entered in submission order.
  std::vector<int64_t> adm;
  {
    size_t used = s->admitted;
    for (int64_t id : s->queue) {
      if ((int)(s->running.size() + adm.size()) >= s->max_running) break;
      const Request& r = *s->req[id];
      if (used + r.footprint > s->budget) break;
      used += r.footprint;
      adm.push_back(id);
    }
  }
  std::vector<mcpet_status> cst(adm.size(), MCPET_OK);
  std::vector<std::string> cerr(adm.size());
  {
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (size_t i = next++; i < adm.size(); i = next++) {
        Request& r = *s->req[adm[i]];
        cst[i] = mcpet_create(&r.cfg, &r.ph, s->tab, &r.ctx);
        if (cst[i] != MCPET_OK) cerr[i] = mcpet_last_error(nullptr);
      }
    };
    const size_t nt = std::min<size_t>(adm.size(), 8);
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(worker);
    if (nt > 0) worker();
    for (auto& t : pool) t.join();
  }
  for (size_t i = 0; i < adm.size(); i++) {   // adm is the queue's head, in order
    Request& r = *s->req[adm[i]];
    s->queue.pop_front();
    std::string why;
    mcpet_status st = cst[i];
    if (st != MCPET_OK) {
      why = std::string("request ") + std::to_string(r.id) + ": " + cerr[i];
    } else if (mcpet_ctx_device_bytes(r.ctx) > r.footprint) {
      st = MCPET_E_STATE;
      why = "mcpet_server_step: device footprint above its estimate";
    }
    if (st != MCPET_OK) {   // only this request fails; the ones behind it are admitted as usual
      fail_request(s, r, st, why);
      continue;
    }
    r.streams = mcpet_ctx_streams(r.ctx);
    r.next = 0;
    r.state = MCPET_REQ_RUNNING;
    s->admitted += r.footprint;
    s->running.push_back(r.id);
  }
  s->st.peak_admitted_bytes = std::max(s->st.peak_admitted_bytes, s->admitted);
  s->st.peak_running = std::max(s->st.peak_running, (int32_t)s->running.size());
  int32_t ndone = 0;
  if (!s->running.empty()) {
    // 2. one batched launch: the next chunk of streams of every running request
    std::vector<CtxRun> runs;
    for (int64_t id : s->running) {
      Request& r = *s->req[id];
      const long long end = std::min(r.streams, r.next + s->per_step);
      runs.push_back({r.ctx, r.next, end, r.next == 0 ? 1 : 0, r.seed});
    }
    std::string e;
    mcpet_status st = mcpet_run_contexts(runs.data(), (int)runs.size(), s->stream, s->h_params, s->d_params, 256,
                                         mcpet_ctx_num_cus(runs[0].c) * 8, &e);
    if (st != MCPET_OK) return sfail(s, st, e);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return sfail(s, MCPET_E_HIP, "mcpet_server_step: launch failed");
    s->st.steps++;
    // 3. completions free their device memory before the next admission; a request whose
    //    results cannot be copied or computed fails alone
    std::vector<int64_t> still;
    for (size_t i = 0; i < s->running.size(); i++) {
      Request& r = *s->req[s->running[i]];
      r.next = runs[i].end;
      if (r.next >= r.streams) {
        const mcpet_status fst = finish(s, r);
        if (fst == MCPET_OK)
          ndone++;
        else
          fail_request(s, r, fst, s->err);
      } else {
        still.push_back(r.id);
      }
    }
    s->running.swap(still);
  }
  s->st.busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  if (running) *running = (int32_t)s->running.size();
  if (completed) *completed = ndone;
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_drain(mcpet_server* s)
{
  if (!s) return sfail(nullptr, MCPET_E_ARG, "mcpet_server_drain: null server");
  while (!s->queue.empty() || !s->running.empty()) {
    mcpet_status st = mcpet_server_step(s, nullptr, nullptr);
    if (st != MCPET_OK) return st;
  }
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_poll(mcpet_server* s, int64_t id, int32_t* state)
{
  if (!s || !state) return sfail(s, MCPET_E_ARG, "mcpet_server_poll: null argument");
  auto it = s->req.find(id);
  *state = it == s->req.end() ? MCPET_REQ_UNKNOWN : it->second->state;
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_fetch(mcpet_server* s, int64_t id, mcpet_tallies* out)
{
  if (!s || !out) return sfail(s, MCPET_E_ARG, "mcpet_server_fetch: null argument");
  auto it = s->req.find(id);
  mcpet_status fst;
  if (fetch_refused(s, it, &fst)) return fst;
  if (it == s->req.end() || it->second->state != MCPET_REQ_DONE)
    return sfail(s, MCPET_E_STATE, "mcpet_server_fetch: request unknown or not finished");
  Request& r = *it->second;
  if (r.grid[0] > 0) return sfail(s, MCPET_E_STATE, "mcpet_server_fetch: a rebinned request (use mcpet_server_fetch_rebinned)");
  const size_t nb = (size_t)r.cfg.nbins, nv = r.img_t.size();
  if (out->trues) std::memcpy(out->trues, r.trues.data(), 4 * nb);
  if (out->scatter) std::memcpy(out->scatter, r.scatter.data(), 4 * nb);
  if (out->image_trues) std::memcpy(out->image_trues, r.img_t.data(), 4 * nv);
  if (out->image_scatter) std::memcpy(out->image_scatter, r.img_sc.data(), 4 * nv);
  if (out->spectrum && r.cfg.ne > 0) std::memcpy(out->spectrum, r.spectrum.data(), 4 * (size_t)r.cfg.ne);
  if (out->material_edep) std::memcpy(out->material_edep, r.mat_edep.data(), 16 * MCPET_MAX_MATERIALS);
  if (out->voxel_edep && !r.voxel_edep.empty()) std::memcpy(out->voxel_edep, r.voxel_edep.data(), 8 * r.voxel_edep.size());
  if (out->psf && !r.psf.empty()) {
    if (out->psf_capacity < (int64_t)r.psf.size()) return sfail(s, MCPET_E_CAPACITY, "mcpet_server_fetch: psf buffer too small");
    std::memcpy(out->psf, r.psf.data(), sizeof(mcpet_psf_record) * r.psf.size());
  }
  out->total_histories = r.t.total_histories;
  out->psf_count = r.t.psf_count;
  out->psf_requested = r.t.psf_requested;
  out->sum_trues = r.t.sum_trues;
  out->sum_scatter = r.t.sum_scatter;
  out->final_seed = r.t.final_seed;
  s->req.erase(it);
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_fetch_rebinned(mcpet_server* s, int64_t id, float* trues, float* scatter, float* act,
                                                    float* att, int64_t* decays)
{
  if (!s) return sfail(s, MCPET_E_ARG, "mcpet_server_fetch_rebinned: null server");
  auto it = s->req.find(id);
  mcpet_status fst;
  if (fetch_refused(s, it, &fst)) return fst;
  if (it == s->req.end() || it->second->state != MCPET_REQ_DONE || it->second->grid[0] == 0 || it->second->net)
    return sfail(s, MCPET_E_STATE, "mcpet_server_fetch_rebinned: request unknown, not finished or not rebinned");
  Request& r = *it->second;
  if ((act || att) && !r.maps)
    return sfail(s, MCPET_E_STATE, "mcpet_server_fetch_rebinned: the request was submitted without maps");
  if (act) std::memcpy(act, r.rb_act.data(), 4 * r.rb_act.size());
  if (att) std::memcpy(att, r.rb_att.data(), 4 * r.rb_att.size());
  if (trues) std::memcpy(trues, r.rb_trues.data(), 4 * r.rb_trues.size());
  if (scatter) std::memcpy(scatter, r.rb_scatter.data(), 4 * r.rb_scatter.size());
  if (decays) *decays = (int64_t)(r.t.total_histories / 2);
  s->req.erase(it);
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_fetch_estimate(mcpet_server* s, int64_t id, float* estimate, float* scatter_counts,
                                                    int64_t* decays)
{
  if (!s) return sfail(s, MCPET_E_ARG, "mcpet_server_fetch_estimate: null server");
  auto it = s->req.find(id);
  mcpet_status fst;
  if (fetch_refused(s, it, &fst)) return fst;
  if (it == s->req.end() || it->second->state != MCPET_REQ_DONE || !it->second->net)
    return sfail(s, MCPET_E_STATE, "mcpet_server_fetch_estimate: request unknown, not finished or not an estimate");
  Request& r = *it->second;
  if (estimate) std::memcpy(estimate, r.est.data(), 4 * r.est.size());
  if (scatter_counts) std::memcpy(scatter_counts, r.rb_scatter.data(), 4 * r.rb_scatter.size());
  if (decays) *decays = r.est_decays;
  s->req.erase(it);
  return MCPET_OK;
}

extern "C" mcpet_status mcpet_server_get_stats(mcpet_server* s, mcpet_server_stats* out)
{
  if (!s || !out) return sfail(s, MCPET_E_ARG, "mcpet_server_get_stats: null argument");
  s->st.queued = (int32_t)s->queue.size();
  s->st.running = (int32_t)s->running.size();
  s->st.latency_p50_ms = percentile(s->latencies, 0.50);
  s->st.latency_p99_ms = percentile(s->latencies, 0.99);
  s->st.latency_max_ms = s->latencies.empty() ? 0.0 : *std::max_element(s->latencies.begin(), s->latencies.end());
  *out = s->st;
  return MCPET_OK;
}

extern "C" void mcpet_server_destroy(mcpet_server* s)
{
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& kv : s->req)
    if (kv.second->ctx) mcpet_destroy(kv.second->ctx);
  if (s->d_params) (void)hipFree(s->d_params);
  if (s->h_params) (void)hipHostFree(s->h_params);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

extern "C" const char* mcpet_server_last_error(const mcpet_server* s)
{
  if (s && !s->err.empty()) return s->err.c_str();
  return mcpet_last_error(nullptr);
}
