// runtime_internal.h — context internals shared by runtime.cpp and server.cpp.
#ifndef MCPET_RUNTIME_INTERNAL_H
#define MCPET_RUNTIME_INTERNAL_H

#include <hip/hip_runtime.h>

#include <string>

#include "kparams.h"
#include "mcpet.h"

// One context's share of a batched launch: streams [begin, end) of its phantom.
struct CtxRun {
  mcpet_ctx* c;
  long long begin, end;
  int clear;       // 1: clear the context's tallies and counters first
  int seed;        // 0: the context's seed
};

// mcpet_create's checks that need neither the device nor an allocation (scan_voxels: also
// every voxel's material slot); on failure the status and the library's last error are set.
mcpet_status mcpet_check_create_args(const mcpet_config* cfg, const mcpet_phantom* ph, const mcpet_tables* tab,
                                     bool scan_voxels);
// Whether a run can use this seed (abMODm divides by zero for 0 and for values below -m2).
bool mcpet_seed_usable(int seed);

size_t mcpet_ctx_device_bytes(const mcpet_ctx* c);
long long mcpet_ctx_streams(const mcpet_ctx* c);
int mcpet_ctx_device(const mcpet_ctx* c);
int mcpet_ctx_num_cus(const mcpet_ctx* c);
// Enqueue one batched launch over runs[0..n-1] on `stream` (h_params: pinned, n entries;
// d_params: device, n entries). The contexts must share their interaction tables.
mcpet_status mcpet_run_contexts(const CtxRun* runs, int n, hipStream_t stream, KParams* h_params, KParams* d_params,
                                int block_size, int blocks, std::string* err);

// Rebinning of tally `which` of the context's last run into the device buffer d_out (D*H*W
// floats), enqueued on `st` after the run has completed (mcpet_rebin_sinogram's grid).
mcpet_status mcpet_ctx_rebin_device(mcpet_ctx* c, int which, int D, int H, int W, float* d_out, hipStream_t st);
// Tallies wa and wb rebinned together (one pass over both sinograms) into d_a and d_b, enqueued
// on `st` behind the context's last mcpet_run without a host wait (the host waits only when
// the grid's plane lists are built, at the first call for a grid).
mcpet_status mcpet_ctx_rebin2_device(mcpet_ctx* c, int wa, int wb, int D, int H, int W, float* d_a, float* d_b, hipStream_t st);
// Decays of the context's last run (its event counter; waits for nothing: call after the run).
mcpet_status mcpet_ctx_decays(mcpet_ctx* c, long long* decays);
// Device address of the context's decay counter (u64; final once its last run has completed).
const unsigned long long* mcpet_ctx_decay_counter(const mcpet_ctx* c);
// Activity line integrals and attenuation factors on a D x H x W rebinned grid, on the device:
// [3][D][H][W] = activity integrals, the same times float32(1 / their maximum), attenuation
// factors; owned by the context, computed at the first call for this grid.
mcpet_status mcpet_ctx_project_device(mcpet_ctx* c, int D, int H, int W, const float** d_maps);
// The network's grid and input channels (unet.hip).
void mcpet_surrogate_dims(const mcpet_surrogate* s, int* D, int* H, int* W, int* cin);

#endif
