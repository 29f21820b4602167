// gahhg_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value)
// pairs whose count lives in device memory.
//
// The grouped query sorts the points the start table leaves open; their
// number is known only on the device once the table pass has run.  This
// sort reads it there, so the host neither waits for it nor sizes a launch
// by it, and every step is a plain kernel (no allocation, no memset): the
// whole query pass can be recorded into a HIP graph.
//
// Per 8-bit digit (the last one cut at end_bit), three steps over tiles of
// 4096 pairs (256 threads x 16):
//   1. k_rs_hist: per tile, the count of each digit (one LDS atomic per
//      pair: the digits of a tile are spread, so a wave-level digit match
//      first costs more than the atomics' conflicts), written digit-major:
//      hist[d * tiles + tile];
//   2. the device scan (gahhg_scan.hip) of that table, in place, over the
//      256 * tiles entries in use: the first output slot of (digit, tile);
//   3. k_rs_scatter: each tile ranks its pairs stably (wave order, then
//      round, then lane), stages them in LDS in digit order and writes each
//      digit's run to consecutive addresses.
// HBM per pair and digit: 4 B read (hist) + 8 B read + 8 B written.
#include "gahhg_internal.h"

namespace gahhg {

constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;
constexpr int kRsWaveItems = kRsTile / (kRsThreads / 64);

// Lanes (among `act`) whose 8-bit digit equals this lane's.
__device__ __forceinline__ unsigned long long digit_match(uint32_t d, unsigned long long act) {
  unsigned long long m = act;
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const bool bit = (d >> b) & 1u;
    const unsigned long long s = __ballot(bit);
    m &= bit ? s : ~s;
  }
  return m;
}

__device__ __forceinline__ int64_t rs_tiles(const int32_t* d_m) {
  return ((int64_t)*d_m + kRsTile - 1) / kRsTile;
}

__global__ void __launch_bounds__(kRsThreads) k_rs_hist(const uint32_t* __restrict__ key,
                                                        const int32_t* __restrict__ d_m, int shift, uint32_t mask,
                                                        int32_t* __restrict__ hist, int32_t* __restrict__ d_hn) {
  __shared__ int cnt[256];
  const int64_t m = *d_m;
  const int64_t nt = rs_tiles(d_m);
  if (blockIdx.x == 0 && threadIdx.x == 0) *d_hn = (int32_t)(256 * nt);
  if ((int64_t)blockIdx.x >= nt) return;
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t base = (int64_t)blockIdx.x * kRsTile + (int64_t)w * kRsWaveItems;
#pragma unroll 4
  for (int r = 0; r < kRsItems; r++) {
    const int64_t i = base + r * 64 + lane;
    const bool ok = i < m;
    if (ok) atomicAdd(&cnt[(key[i] >> shift) & mask], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nt + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(kRsThreads) k_rs_scatter(const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           const int32_t* __restrict__ d_m, int shift,
                                                           uint32_t mask, const int32_t* __restrict__ offs,
                                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
  __shared__ int wcnt[kRsThreads / 64][256];  // per wave: its pairs of each digit, then their start
  __shared__ int dstart[256];                 // tile-local first slot of each digit
  __shared__ int gshift[256];                 // output address minus tile-local slot, per digit
  __shared__ int wsum[kRsThreads / 64];
  __shared__ uint32_t sk[kRsTile], sv[kRsTile];
  const int64_t m = *d_m;
  const int64_t nt = rs_tiles(d_m);
  if ((int64_t)blockIdx.x >= nt) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int k = 0; k < kRsThreads / 64; k++) wcnt[k][tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRsTile + (int64_t)w * kRsWaveItems;
  uint32_t kk[kRsItems], vv[kRsItems];
  int rk[kRsItems];
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    const int64_t i = base + r * 64 + lane;
    const bool ok = i < m;
    kk[r] = ok ? kin[i] : 0u;
    vv[r] = ok ? vin[i] : 0u;
  }
  // rank within the wave: pairs before this one with the same digit
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    const bool ok = base + r * 64 + lane < m;
    const uint32_t d = (kk[r] >> shift) & mask;
    const unsigned long long mm = digit_match(d, __ballot(ok));
    const int leader = ok ? __ffsll((long long)mm) - 1 : lane;
    int old = 0;
    if (ok && leader == lane) old = atomicAdd(&wcnt[w][d], __popcll(mm));
    old = __shfl(old, leader, 64);
    rk[r] = ok ? old + __popcll(mm & ((1ull << lane) - 1ull)) : -1;
  }
  __syncthreads();
  // thread tid = digit: start of each wave's run, then the digit's start
  int tot = 0;
#pragma unroll
  for (int k = 0; k < kRsThreads / 64; k++) {
    const int c = wcnt[k][tid];
    wcnt[k][tid] = tot;
    tot += c;
  }
  int inc = tot;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const int u = __shfl_up(inc, s, 64);
    if (lane >= s) inc += u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int ex = inc - tot;
  for (int k = 0; k < w; k++) ex += wsum[k];
  dstart[tid] = ex;
  gshift[tid] = offs[(int64_t)tid * nt + blockIdx.x] - ex;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    if (rk[r] >= 0) {
      const uint32_t d = (kk[r] >> shift) & mask;
      const int s = dstart[d] + wcnt[w][d] + rk[r];
      sk[s] = kk[r];
      sv[s] = vv[r];
    }
  }
  __syncthreads();
  const int cnt = (int)min<int64_t>(kRsTile, m - (int64_t)blockIdx.x * kRsTile);
#pragma unroll
  for (int r = 0; r < kRsItems; r++) {
    const int s = r * kRsThreads + tid;
    if (s < cnt) {
      const uint32_t k = sk[s];
      const int64_t pos = (int64_t)gshift[(k >> shift) & mask] + s;
      kout[pos] = k;
      vout[pos] = sv[s];
    }
  }
}

int64_t radix_hist_entries(int64_t n_max) { return 256 * ((n_max + kRsTile - 1) / kRsTile) + 1; }

int radix_sort_pairs(gahhg_ctx* c, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const int32_t* d_m,
                     int64_t n_max, int end_bit, int32_t* hist, uint32_t** kres, uint32_t** vres,
                     int32_t* scan_scratch) {
  if (n_max > INT32_MAX) return set_err(c, GAHHG_E_LIMIT, "radix sort: more than 2^31 pairs");
  const int64_t tiles = (n_max + kRsTile - 1) / kRsTile;
  int32_t* d_hn = hist + 256 * tiles;
  uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
  for (int shift = 0; shift < end_bit; shift += 8) {
    const uint32_t mask = (1u << std::min(8, end_bit - shift)) - 1u;  // bits past end_bit do not count
    if (tiles > 0) {
      hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)tiles), dim3(kRsThreads), 0, c->stream, (const uint32_t*)ki,
                         d_m, shift, mask, hist, d_hn);
      int rc = hip_check(c, hipGetLastError(), "radix histogram");
      if (rc) return rc;
      rc = scan_i32_with(c, hist, hist, 256 * tiles, d_hn, nullptr, 0, scan_scratch);
      if (rc) return rc;
      hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)tiles), dim3(kRsThreads), 0, c->stream, (const uint32_t*)ki,
                         (const uint32_t*)vi, d_m, shift, mask, (const int32_t*)hist, ko, vo);
      rc = hip_check(c, hipGetLastError(), "radix scatter");
      if (rc) return rc;
    }
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  *kres = ki;
  *vres = vi;
  return GAHHG_OK;
}

// Loads this file's code object onto the current device (context set-up).
void load_code_sort() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_rs_hist));
}

}  // namespace gahhg
