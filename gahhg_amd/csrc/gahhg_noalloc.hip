// gahhg_noalloc.hip -- the allocation-free query entry point.
//
// gahhg_query_batch_xyz_device grows the context's grouping buffers (sort
// keys and indices, radix histograms, scan partials) to the largest pass it
// has seen.  Callers that must not allocate call
// gahhg_query_batch_xyz_device_noalloc instead: its pass takes that scratch
// from `scratch`, a fixed array in device memory sized for
// GAHHG_NOALLOC_MAX_POINTS points.  Each GPU holds its own copy, at its own
// address, from the load of this code object to the end of the process;
// every context on one GPU shares it, so calls that overlap on one GPU race.
// Every word the pass reads is written earlier in the same pass (block
// counts by the table pass, offsets and totals by the scan, histograms by
// k_rs_hist, keys and indices by the compaction and the scatters), so the
// last call's contents need no clearing.
#include <mutex>

#include "gahhg_internal.h"

namespace gahhg {

constexpr int64_t kNoAllocMax = GAHHG_NOALLOC_MAX_POINTS;
// Upper bound of sort_buffer_bytes(kNoAllocMax): block counts, offsets and
// total (8 B per 256 points + 4, rounded to 256 B), 16 B per point for two
// key and two index arrays, 4 B per radix-histogram entry (256 per tile of
// 4096 points, + 1).
constexpr int64_t kNoAllocSortBytes =
    ((8 * ((kNoAllocMax + 255) / 256) + 4 + 255) / 256) * 256 + 16 * kNoAllocMax +
    4 * (256 * ((kNoAllocMax + 4095) / 4096) + 1);
constexpr int64_t kNoAllocSortBytesAligned = (kNoAllocSortBytes + 255) / 256 * 256;
// Scan partial sums: a scan of L elements uses 2 per tile of 4096 at each
// level of its recursion (+2); 4096 int32 cover L far beyond what one pass
// scans (at most 256 * 1025 histogram entries).
constexpr int64_t kNoAllocScanEntries = 4096;
constexpr int64_t kNoAllocBytes = kNoAllocSortBytesAligned + 4 * kNoAllocScanEntries;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
int g_state[kMaxDevices];  // 0: not resolved, 1: resolved, -1: failed
void* g_ptr[kMaxDevices];  // device address of `scratch` on each GPU
}  // namespace

void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_state[device] != 0) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p) {
    g_ptr[device] = p;
    g_state[device] = 1;
  } else {
    (void)hipGetLastError();  // the failure is reported by the entry point
    g_state[device] = -1;
  }
}

static void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  return g_state[device] == 1 ? g_ptr[device] : nullptr;
}

}  // namespace gahhg

using namespace gahhg;

extern "C" int gahhg_query_batch_xyz_device_noalloc(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly,
                                                    int64_t n, int32_t* d_labels) {
  if (!c) return GAHHG_E_ARG;
  if (n > kNoAllocMax)
    return set_err(c, GAHHG_E_LIMIT, "noalloc query: %lld points exceed the limit of %lld per call", (long long)n,
                   (long long)kNoAllocMax);
  void* base = scratch_ptr(c->device);
  if (!base) return set_err(c, GAHHG_E_HIP, "noalloc query: the scratch array of device %d is not available", c->device);
  if (sort_buffer_bytes(kNoAllocMax) > kNoAllocSortBytesAligned ||
      scan_scratch_entries(std::max<int64_t>((kNoAllocMax + 255) / 256, radix_hist_entries(kNoAllocMax))) >
          kNoAllocScanEntries)
    return set_err(c, GAHHG_E_LIMIT, "noalloc query: scratch layout larger than the array");
  SortScratch s;
  s.buf = base;
  s.cap = kNoAllocMax;
  s.scan = (int32_t*)((char*)base + kNoAllocSortBytesAligned);
  return query_device(c, d_xyz, d_poly, n, d_labels, &s);
}
