// coinc.hip — singles pipeline after the transport: sort the singles by detection time and
// pair them with a coincidence timing window (no counterpart in the reference, which only
// tallies the two photons of each annihilation, MCGPU-PET_kernel.cu:1623-1836).
//
// Sort: least-significant-digit radix sort of the 64-bit detection times with the record
// index as payload, 8-bit digits, as many passes as the largest time needs. Each pass is
// three kernels over tiles of RS_TILE keys: per-tile digit histograms (LDS atomics), an
// exclusive scan of the digit-major histogram table (global offsets), and a stable scatter
// in which every wave ranks its 64 keys by digit with eight ballots and the workgroup adds
// the counts of earlier waves and earlier rounds of the tile from LDS. Stable at every
// level, so equal times keep the order of the singles list.
//
// Pairing: a coincidence is a cluster of exactly two singles whose times differ by at most
// the window, with gaps larger than the window on both sides; a cluster of three or more
// singles is a multiple and is discarded. Every single decides locally whether it opens such
// a cluster (at most two neighbours read), so the kernel is one parallel pass. A pair is
// true (same decay, both unscattered annihilation photons), scattered (same decay, either
// scattered), spurious (same decay, one photon is the prompt gamma) or random (different
// decays), and is binned with the reference's Michelogram arithmetic (earlier single first;
// singles with equal times in the order of (event, kind)).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "devutil.h"
#include "host_internal.h"
#include "kparams.h"
#include "mcpet.h"

#define RS_THREADS 256
#define RS_ROUNDS 16
#define RS_TILE (RS_THREADS * RS_ROUNDS)
#define SCAN_THREADS 256
#define SCAN_ITEMS 4
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

namespace {

__global__ void k_keys_from_singles(const mcpet_single* __restrict__ s, long long n, unsigned long long* __restrict__ keys,
                                    unsigned* __restrict__ vals)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    keys[i] = s[i].time_ps;
    vals[i] = (unsigned)i;
  }
}

// hist[d * ntiles + tile] = occurrences of digit d in the tile.
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const unsigned long long* __restrict__ keys, long long n, int shift,
                                                        unsigned* __restrict__ hist, int ntiles)
{
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const long long base = (long long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; r++) {
    const long long i = base + (long long)r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile by digit, through LDS: every key's place in the tile ordered by
// digit (stable: by round, wave and lane, the input order) is found with eight ballots per wave
// and the counts of earlier waves and rounds; the keys and indices are written there in LDS, and
// the tile is then written out in that order, so that consecutive lanes store consecutive
// elements of one digit's run (its bucket in the output) instead of up to 64 different buckets.
// offs is the digit-major table of global offsets (exclusive scan of the tile histograms): the
// tile's count of digit d is the next entry minus this one.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const unsigned long long* __restrict__ kin,
                                                           const unsigned* __restrict__ vin, unsigned long long* __restrict__ kout,
                                                           unsigned* __restrict__ vout, long long n, int shift,
                                                           const unsigned* __restrict__ offs, int ntiles)
{
  __shared__ unsigned long long s_key[RS_TILE];
  __shared__ unsigned s_val[RS_TILE];
  __shared__ unsigned s_gbase[256];                 // output offset of digit d for this tile
  __shared__ unsigned s_start[256];                 // start of digit d in the tile's sorted order
  __shared__ unsigned s_run[256];                   // keys of digit d placed so far
  __shared__ unsigned s_wcnt[RS_THREADS / 64][256];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const long long base = (long long)blockIdx.x * RS_TILE;
  const int tile_n = (int)min((long long)RS_TILE, n - base);
  {
    const size_t e = (size_t)tid * ntiles + blockIdx.x;
    const unsigned g = offs[e];
    const unsigned next = (e + 1 < (size_t)256 * ntiles) ? offs[e + 1] : (unsigned)n;
    s_gbase[tid] = g;
    s_start[tid] = next - g;                        // this tile's count of digit tid, scanned below
    s_run[tid] = 0u;
  }
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {         // inclusive scan of the counts over digits
    const unsigned x = tid >= off ? s_start[tid - off] : 0u;
    __syncthreads();
    s_start[tid] += x;
    __syncthreads();
  }
  const unsigned excl = tid > 0 ? s_start[tid - 1] : 0u;   // exclusive: the inclusive scan shifted by one digit
  __syncthreads();
  s_start[tid] = excl;
  __syncthreads();
  for (int r = 0; r < RS_ROUNDS; r++) {
    const int j = r * RS_THREADS + tid;
    const bool valid = j < tile_n;
    const unsigned long long key = valid ? kin[base + j] : 0ull;
    const unsigned val = valid ? vin[base + j] : 0u;
    const unsigned d = (unsigned)(key >> shift) & 255u;
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bb = __ballot(valid && bit);
      m &= bit ? bb : ~bb;
    }
    const unsigned rank = (unsigned)__popcll(m & lt);
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; w++) s_wcnt[w][tid] = 0u;
    __syncthreads();
    if (valid && rank == 0u) s_wcnt[wave][d] = (unsigned)__popcll(m);
    __syncthreads();
    if (valid) {
      unsigned pos = s_start[d] + s_run[d] + rank;
      for (int w = 0; w < wave; w++) pos += s_wcnt[w][d];
      s_key[pos] = key;
      s_val[pos] = val;
    }
    __syncthreads();
    unsigned add = 0u;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; w++) add += s_wcnt[w][tid];
    s_run[tid] += add;
    __syncthreads();
  }
  // the tile in digit order: element j of digit d goes to s_gbase[d] + (j - s_start[d])
  for (int j = tid; j < tile_n; j += RS_THREADS) {
    const unsigned long long key = s_key[j];
    const unsigned d = (unsigned)(key >> shift) & 255u;
    const unsigned pos = s_gbase[d] + ((unsigned)j - s_start[d]);
    kout[pos] = key;
    vout[pos] = s_val[j];
  }
}

// Exclusive scan of one tile of SCAN_TILE values; the tile total goes to sums[tile].
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile(const unsigned* __restrict__ in, unsigned* __restrict__ out,
                                                             long long m, unsigned* __restrict__ sums)
{
  __shared__ unsigned s[SCAN_THREADS];
  const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  unsigned v[SCAN_ITEMS], t = 0u;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = (base + k < m) ? in[base + k] : 0u;
    t += v[k];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {   // inclusive Hillis-Steele scan of thread totals
    const unsigned x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  unsigned run = s[threadIdx.x] - t;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < m) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == SCAN_THREADS - 1) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(unsigned* __restrict__ out, long long m, const unsigned* __restrict__ sums)
{
  const long long base = (long long)blockIdx.x * SCAN_TILE;
  const unsigned a = sums[blockIdx.x];
  for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_THREADS)
    if (base + k < m) out[base + k] += a;
}

// One single per lane: does it open a coincidence cluster, and of what kind?
__global__ __launch_bounds__(256) void k_pair(KParams P, const unsigned long long* __restrict__ t, const unsigned* __restrict__ perm,
                                              long long n)
{
  __shared__ unsigned s_cnt[5];
  if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0u;
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long W = P.window_ps;
  if (i < n && (i == 0 || t[i] - t[i - 1] > W) && i + 1 < n && t[i + 1] - t[i] <= W) {
    if (i + 2 < n && t[i + 2] - t[i + 1] <= W) {
      atomicAdd(&s_cnt[4], 1u);                       // three or more singles: multiple
    } else {
      mcpet_single a = P.singles[perm[i]];
      mcpet_single b = P.singles[perm[i + 1]];
      // equal detection times: order by (event, kind), so that the binning (earlier single
      // first) does not depend on the order in which the singles were appended
      if (t[i] == t[i + 1] && (b.event < a.event || (b.event == a.event && b.kind < a.kind))) {
        const mcpet_single x = a;
        a = b;
        b = x;
      }
      int cls;
      if (a.event != b.event) cls = 2;                                  // random
      else if (a.kind == 2 || b.kind == 2) cls = 3;                     // spurious (prompt gamma)
      else cls = (a.scatter == 0 && b.scatter == 0) ? 0 : 1;            // true / scattered
      const int ibin = sino_bin(P, a.phi, a.z, b.phi, b.z);
      if (ibin >= 0) {
        int* sino = cls == 0 ? P.trues : cls == 1 ? P.scatter : cls == 2 ? P.randoms : P.spurious;
        atomicAdd(&sino[ibin], 1);
        atomicAdd(&s_cnt[cls], 1u);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 5 && s_cnt[threadIdx.x]) atomicAdd(&P.coinc_counts[threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
}

__global__ void k_gather_singles(const mcpet_single* __restrict__ s, const unsigned* __restrict__ perm, long long n,
                                 mcpet_single* __restrict__ out)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = s[perm[i]];
}

__global__ void k_iota(unsigned* __restrict__ vals, long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) vals[i] = (unsigned)i;
}

inline int nblocks(long long n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

long long mcpet_rs_tiles(long long n) { return (n + RS_TILE - 1) / RS_TILE; }

// Scratch sizes (in unsigned words) of the digit-offset scan for up to n keys: the table itself
// and the per-level tile sums.
void mcpet_scan_levels(long long m, std::vector<long long>& levels)
{
  levels.clear();
  while (true) {
    const long long t = (m + SCAN_TILE - 1) / SCAN_TILE;
    levels.push_back(t);
    if (t <= 1) break;
    m = t;
  }
}

static hipError_t scan_exclusive(unsigned* data, long long m, unsigned* const* level_bufs, int level, hipStream_t st)
{
  const long long tiles = (m + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, st, data, data, m, level_bufs[level]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || tiles <= 1) return e;
  e = scan_exclusive(level_bufs[level], tiles, level_bufs, level + 1, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, st, data, m, level_bufs[level]);
  return hipGetLastError();
}

hipError_t mcpet_launch_keys_from_singles(const mcpet_single* singles, long long n, unsigned long long* keys, unsigned* vals,
                                         hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys_from_singles, dim3(nblocks(n, 256)), dim3(256), 0, st, singles, n, keys, vals);
  return hipGetLastError();
}

hipError_t mcpet_launch_iota(unsigned* vals, long long n, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(nblocks(n, 256)), dim3(256), 0, st, vals, n);
  return hipGetLastError();
}

// Radix sort of keys[0] (payload vals[0]) by the low 8*passes bits; the result is in
// keys[*which], vals[*which].
hipError_t mcpet_radix_sort(long long n, int passes, unsigned long long* keys[2], unsigned* vals[2], unsigned* hist,
                            unsigned* const* level_bufs, int* which, hipStream_t st)
{
  *which = 0;
  if (n <= 0) return hipSuccess;
  hipError_t e = hipSuccess;
  const int ntiles = (int)mcpet_rs_tiles(n);
  int cur = 0;
  for (int p = 0; p < passes && e == hipSuccess; p++) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, st, keys[cur], n, shift, hist, ntiles);
    e = hipGetLastError();
    if (e == hipSuccess) e = scan_exclusive(hist, 256ll * ntiles, level_bufs, 0, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, st, keys[cur], vals[cur], keys[cur ^ 1], vals[cur ^ 1], n,
                         shift, hist, ntiles);
      e = hipGetLastError();
    }
    cur ^= 1;
  }
  *which = cur;
  return e;
}

hipError_t mcpet_launch_pairing(const KParams& P, const unsigned long long* keys, const unsigned* perm, long long n, hipStream_t st)
{
  if (n < 2) return hipSuccess;
  hipLaunchKernelGGL(k_pair, dim3(nblocks(n, 256)), dim3(256), 0, st, P, keys, perm, n);
  return hipGetLastError();
}

hipError_t mcpet_launch_gather_singles(const mcpet_single* s, const unsigned* perm, long long n, mcpet_single* out, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_singles, dim3(nblocks(n, 256)), dim3(256), 0, st, s, perm, n, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------- allocation-free sort
// mcpet_selftest_sort_noalloc: the same sort as mcpet_selftest_sort with its device buffers (both
// key/payload buffers, the digit histogram table and the scan's per-level tile sums) carved out
// of one static device array, for callers that cannot allocate device memory. The array is
// loaded with the code object, so every device the library runs on holds it until the process
// ends. Its contents are not cleared between calls: nothing in it needs to start at zero (the
// keys and payload are written first, every histogram entry is written by k_rs_hist, every scan
// sum by k_scan_tile), and two calls that overlap on one device race on it.
#define NA_MAX_KEYS (1LL << 20)
#define NA_HIST_WORDS (256LL * ((NA_MAX_KEYS + RS_TILE - 1) / RS_TILE))
#define NA_LEVEL_WORDS 1024LL
#define NA_BYTES (2 * 8 * NA_MAX_KEYS + 2 * 4 * NA_MAX_KEYS + 4 * NA_HIST_WORDS + 4 * NA_LEVEL_WORDS)

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
void* g_na_ptr[kNaMaxDevices] = {};   // device address of `scratch` on each device

mcpet_status na_fail(mcpet_status st, const std::string& msg)
{
  mcpet_host_set_error(msg.c_str());
  return st;
}
}  // namespace

extern "C" mcpet_status mcpet_selftest_sort_noalloc(const uint64_t* keys, int64_t n, int32_t passes, uint32_t* perm)
{
  if (!keys || !perm || n < 0 || passes < 1 || passes > 8)
    return na_fail(MCPET_E_ARG, "mcpet_selftest_sort_noalloc: bad arguments");
  if (n > NA_MAX_KEYS)
    return na_fail(MCPET_E_CAPACITY, "mcpet_selftest_sort_noalloc: " + std::to_string(n) + " keys exceed the static scratch's " +
                                         std::to_string(NA_MAX_KEYS) + " (use mcpet_selftest_sort)");
  if (n == 0) return MCPET_OK;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return na_fail(MCPET_E_HIP, std::string("mcpet_selftest_sort_noalloc: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kNaMaxDevices) return na_fail(MCPET_E_RANGE, "mcpet_selftest_sort_noalloc: device index out of range");
  unsigned char* base = nullptr;
  {
    // first call on this device (current): the array's address on this device
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return na_fail(MCPET_E_HIP, std::string("mcpet_selftest_sort_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
      g_na_ptr[dev] = p;
    }
    base = static_cast<unsigned char*>(g_na_ptr[dev]);
  }
  const long long m = 256LL * mcpet_rs_tiles(n);
  std::vector<long long> levels;
  mcpet_scan_levels(m, levels);
  long long level_words = 0;
  for (long long l : levels) level_words += l;
  if (m > NA_HIST_WORDS || level_words > NA_LEVEL_WORDS)
    return na_fail(MCPET_E_CAPACITY, "mcpet_selftest_sort_noalloc: scan tables exceed the static scratch");
  unsigned long long* k[2] = {reinterpret_cast<unsigned long long*>(base),
                              reinterpret_cast<unsigned long long*>(base + 8 * NA_MAX_KEYS)};
  unsigned* v[2] = {reinterpret_cast<unsigned*>(base + 16 * NA_MAX_KEYS), reinterpret_cast<unsigned*>(base + 20 * NA_MAX_KEYS)};
  unsigned* hist = reinterpret_cast<unsigned*>(base + 24 * NA_MAX_KEYS);
  std::vector<unsigned*> scan(levels.size());
  unsigned* lp = hist + NA_HIST_WORDS;
  for (size_t i = 0; i < levels.size(); i++) {
    scan[i] = lp;
    lp += levels[i];
  }
  int which = 0;
  e = hipMemcpy(k[0], keys, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mcpet_launch_iota(v[0], n, 0);
  if (e == hipSuccess) e = mcpet_radix_sort(n, passes, k, v, hist, scan.data(), &which, 0);
  if (e == hipSuccess) e = hipStreamSynchronize(0);
  if (e == hipSuccess) e = hipMemcpy(perm, v[which], 4 * (size_t)n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return na_fail(MCPET_E_HIP, std::string("mcpet_selftest_sort_noalloc: ") + hipGetErrorString(e));
  return MCPET_OK;
}
