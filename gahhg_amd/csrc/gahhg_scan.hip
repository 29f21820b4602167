// gahhg_scan.hip -- device-wide int32 prefix sum for gfx950.
//
// Replaces the reference's Blelloch scan (source/scan/scan.cu:104-171,
// source/scan/kernels.cu:76-124,186-249), which allocates and frees device
// memory on every call and launches up to four kernels.  Here: reduce-then-
// scan over 4096-element tiles (256 threads x 16 items), with a wave64
// shuffle prefix inside each wave and an LDS exchange of the four wave
// totals.  Tile totals are scanned recursively with the same kernels, so
// any length runs in 2 + 2*levels launches with no allocation in the call
// path (scratch is held by the context).
#include "gahhg_internal.h"

namespace gahhg {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int wave_inclusive(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total
// through *total.
__device__ __forceinline__ int block_exclusive(int v, int* total) {
  __shared__ int wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  int inc = wave_inclusive(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / 64; i++) {
    int s = wsum[i];
    if (i < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

__global__ void __launch_bounds__(kScanThreads) k_tile_reduce(const int32_t* __restrict__ in,
                                                              int64_t n_bound, const int32_t* d_n,
                                                              int32_t* __restrict__ sums) {
  const int64_t n = d_n ? (int64_t)*d_n : n_bound;
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;  // coalesced
    if (i < n) s += in[i];
  }
  int tot;
  (void)block_exclusive(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_tile_scan(const int32_t* __restrict__ in,
                                                            int32_t* __restrict__ out,
                                                            int64_t n_bound, const int32_t* d_n,
                                                            const int32_t* __restrict__ offs,
                                                            int32_t* d_total, int inclusive) {
  __shared__ int tile[kScanTile];
  const int64_t n = d_n ? (int64_t)*d_n : n_bound;
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  // coalesced load into LDS, then each thread owns 16 consecutive items
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int j = k * kScanThreads + threadIdx.x;
    int64_t i = base + j;
    tile[j] = (i < n) ? in[i] : 0;
  }
  __syncthreads();
  int loc[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    loc[k] = tile[threadIdx.x * kScanItems + k];
    s += loc[k];
  }
  int tot;
  int ex = block_exclusive(s, &tot);
  const int off = offs ? offs[blockIdx.x] : 0;
  int run = off + ex;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int v = loc[k];
    tile[threadIdx.x * kScanItems + k] = inclusive ? run + v : run;
    run += v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int j = k * kScanThreads + threadIdx.x;
    int64_t i = base + j;
    if (i < n) out[i] = tile[j];
  }
  if (d_total && threadIdx.x == kScanThreads - 1 && blockIdx.x == gridDim.x - 1) *d_total = off + tot;
}

static int64_t tiles_of(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

// Scratch needed for a scan of n elements: tile sums and their offsets at
// every recursion level.
static int64_t scratch_need(int64_t n) {
  int64_t need = 0;
  int64_t t = tiles_of(n);
  while (t > 1) {
    need += 2 * t;
    t = tiles_of(t);
  }
  return need + 2;
}

static int scan_rec(gahhg_ctx* c, const int32_t* in, int32_t* out, int64_t n_bound,
                    const int32_t* d_n, int32_t* d_total, int inclusive, int32_t* scratch) {
  const int64_t tiles = tiles_of(n_bound);
  if (tiles <= 1) {
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kScanThreads), 0, c->stream, in, out, n_bound,
                       d_n, nullptr, d_total, inclusive);
    return hip_check(c, hipGetLastError(), "scan");
  }
  int32_t* sums = scratch;
  int32_t* offs = scratch + tiles;
  hipLaunchKernelGGL(k_tile_reduce, dim3((unsigned)tiles), dim3(kScanThreads), 0, c->stream, in,
                     n_bound, d_n, sums);
  int rc = scan_rec(c, sums, offs, tiles, nullptr, nullptr, 0, scratch + 2 * tiles);
  if (rc) return rc;
  hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)tiles), dim3(kScanThreads), 0, c->stream, in, out,
                     n_bound, d_n, offs, d_total, inclusive);
  return hip_check(c, hipGetLastError(), "scan");
}

int scan_i32(gahhg_ctx* c, const int32_t* in, int32_t* out, int64_t n_bound, const int32_t* d_n,
             int32_t* d_total, int inclusive) {
  if (n_bound <= 0) {
    if (d_total) return fill_i32(c, d_total, 1, 0);
    return GAHHG_OK;
  }
  const int rc = scan_reserve(c, n_bound);
  if (rc) return rc;
  return scan_rec(c, in, out, n_bound, d_n, d_total, inclusive, c->d_scan);
}

int64_t scan_scratch_entries(int64_t n_bound) { return scratch_need(std::max<int64_t>(n_bound, 1)); }

int scan_i32_with(gahhg_ctx* c, const int32_t* in, int32_t* out, int64_t n_bound, const int32_t* d_n,
                  int32_t* d_total, int inclusive, int32_t* scratch) {
  if (!scratch) return scan_i32(c, in, out, n_bound, d_n, d_total, inclusive);
  if (n_bound <= 0) {
    if (d_total) return fill_i32(c, d_total, 1, 0);
    return GAHHG_OK;
  }
  return scan_rec(c, in, out, n_bound, d_n, d_total, inclusive, scratch);
}

// Chunked scan of up to four columns whose length lives in device memory
// (*d_n, the same for every column).  The grid is fixed -- kChunkBlocks
// blocks per column, block b owning a contiguous chunk of whole tiles --
// so the launch does not depend on any bound of the length and costs the
// same two launches for 10 elements as for 10 million; columns run side by
// side (blockIdx.y).  Pass 1 reduces each chunk; pass 2 has every block sum
// the chunk totals before its own (one per thread) and scan its chunk tile
// by tile.
constexpr int kChunkBlocks = kScanThreads;

__device__ __forceinline__ int64_t chunk_len(int64_t n) {
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  return (tiles + kChunkBlocks - 1) / kChunkBlocks * kScanTile;
}

__global__ void __launch_bounds__(kScanThreads) k_chunk_reduce(ScanCols s, const int32_t* __restrict__ d_n,
                                                               int32_t* __restrict__ sums) {
  const int k = blockIdx.y;
  const int64_t n = *d_n;
  const int64_t ch = chunk_len(n);
  const int64_t lo = (int64_t)blockIdx.x * ch, hi = min(n, lo + ch);
  const int32_t* in = s.in[k];
  int v = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kScanThreads) v += in[i];
  int tot;
  (void)block_exclusive(v, &tot);
  if (threadIdx.x == 0) sums[k * kChunkBlocks + blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_chunk_scan(ScanCols s, const int32_t* __restrict__ d_n,
                                                             const int32_t* __restrict__ sums, int inclusive) {
  __shared__ int tile[kScanTile];
  const int k = blockIdx.y;
  const int t = threadIdx.x;
  const int64_t n = *d_n;
  const int64_t ch = chunk_len(n);
  const int64_t lo = (int64_t)blockIdx.x * ch, hi = min(n, lo + ch);
  int run;
  (void)block_exclusive(t < (int)blockIdx.x ? sums[k * kChunkBlocks + t] : 0, &run);  // this chunk's offset
  const int32_t* in = s.in[k];
  int32_t* out = s.out[k];
  for (int64_t base = lo; base < hi; base += kScanTile) {
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
      const int j = q * kScanThreads + t;
      const int64_t i = base + j;
      tile[j] = i < hi ? in[i] : 0;
    }
    __syncthreads();
    int loc[kScanItems];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
      loc[q] = tile[t * kScanItems + q];
      sum += loc[q];
    }
    int ttot;
    int r = run + block_exclusive(sum, &ttot);
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
      tile[t * kScanItems + q] = inclusive ? r + loc[q] : r;
      r += loc[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
      const int j = q * kScanThreads + t;
      const int64_t i = base + j;
      if (i < hi) out[i] = tile[j];
    }
    __syncthreads();
    run += ttot;
  }
  if (s.total[k] && blockIdx.x == kChunkBlocks - 1 && t == 0) *s.total[k] = run;  // the last chunk ends the column
}

int scan_cols_dev(gahhg_ctx* c, const ScanCols& cols, int ncols, const int32_t* d_n, int inclusive, int32_t* sums) {
  hipLaunchKernelGGL(k_chunk_reduce, dim3(kChunkBlocks, ncols), dim3(kScanThreads), 0, c->stream, cols, d_n, sums);
  hipLaunchKernelGGL(k_chunk_scan, dim3(kChunkBlocks, ncols), dim3(kScanThreads), 0, c->stream, cols, d_n,
                     (const int32_t*)sums, inclusive);
  return hip_check(c, hipGetLastError(), "chunked scan");
}

int scan_reserve(gahhg_ctx* c, int64_t n_bound) {
  const int64_t need = scratch_need(std::max<int64_t>(n_bound, 1));
  if (need <= c->scap) return GAHHG_OK;
  (void)hipStreamSynchronize(c->stream);
  dfree(c->d_scan);
  c->d_scan = nullptr;
  c->scap = 0;
  c->qepoch++;
  const int64_t cap = need * 2;
  if (dalloc(c, (void**)&c->d_scan, cap * (int64_t)sizeof(int32_t))) return GAHHG_E_NOMEM;
  c->scap = cap;
  return GAHHG_OK;
}

// Loads this file's code object onto the current device (context set-up).
void load_code_scan() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_tile_reduce));
}

}  // namespace gahhg
