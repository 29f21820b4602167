// gahhg_load.hip -- text input parsed on the device.
//
// GPISPTP::loadTPFromCSV (source/GPISPTP.cu:14-77) reads a query file line
// by line with sscanf_s("%lf,%lf") and converts each point with FakeLL2XYZ
// (source/GPISPPolygon.cuh:72,82-92); GPISPPolygon::loadPolyFromTXT
// (source/GPISPPolygon.cu:32-278) reads "%lf %lf" vertex lines.  Here the
// file's bytes are copied to HBM once and the rest runs on the device:
//   1. line split: 16 bytes per lane (one 128-bit load), newline count per
//      block, a device scan of the block counts, then each lane writes the
//      start offsets of the lines that begin in its 16 bytes;
//   2. parse: one lane per line, the number grammar of strtod for decimal
//      input, converted with correct rounding (gahhg_decimal.h: Clinger's
//      exact fast path, else Eisel-Lemire with a 128-bit power-of-five table);
//   3. the point converted to xyz with the correctly rounded sin and cos of
//      gahhg_trig.h (the host conversion uses the same source).
// A line the device does not convert exactly (no digits, hex or inf/nan
// text, a nonzero digit past the 19th significant one, a subnormal or
// overflowing value, a separator other than the format's) is listed; the
// host then parses those lines with sscanf, the reference's own call, and
// reports the first that fails ("wrong format", with its line number) as the
// host loader does.  The values therefore equal the host loader's on every
// line.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "gahhg_decimal.h"
#include "gahhg_internal.h"
#include "gahhg_trig.h"

#pragma clang fp contract(off)

namespace gahhg {
namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int kLoadThreads = 256;
constexpr int kLaneBytes = 16;
constexpr int64_t kBlockBytes = (int64_t)kLoadThreads * kLaneBytes;
constexpr int64_t kStageBytes = 64ll << 20;  // pinned staging buffer for file reads

// Number of bytes equal to '\n' in a 32-bit word (exact, no carries between
// bytes: (b & 0x7f) + 0x7f <= 0xfe).
__device__ __forceinline__ int nl_bytes(uint32_t w) {
  const uint32_t t = w ^ 0x0a0a0a0au;
  const uint32_t y = ~(((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t | 0x7f7f7f7fu);
  return __popc(y);
}

__device__ __forceinline__ int nl_count16(uint4 v) {
  return nl_bytes(v.x) + nl_bytes(v.y) + nl_bytes(v.z) + nl_bytes(v.w);
}

__global__ void __launch_bounds__(kLoadThreads) k_nl_count(const uint8_t* __restrict__ text,
                                                           int32_t* __restrict__ counts) {
  __shared__ int part[kLoadThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kBlockBytes + (int64_t)threadIdx.x * kLaneBytes;
  int c = nl_count16(*(const uint4*)(text + base));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kLoadThreads / 64; w++) s += part[w];
    counts[blockIdx.x] = s;
  }
}

// starts[0] = 0; starts[k + 1] = offset after the k-th newline.
__global__ void __launch_bounds__(kLoadThreads) k_nl_write(const uint8_t* __restrict__ text,
                                                           const int32_t* __restrict__ offs,
                                                           int64_t* __restrict__ starts) {
  __shared__ int part[kLoadThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kBlockBytes + (int64_t)threadIdx.x * kLaneBytes;
  const uint4 v = *(const uint4*)(text + base);
  const int c = nl_count16(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = c;  // inclusive scan within the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) part[wid] = inc;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wid; w++) before += part[w];
  int64_t pos = (int64_t)offs[blockIdx.x] + before + inc - c;
  if (blockIdx.x == 0 && threadIdx.x == 0) starts[0] = 0;
  if (c == 0) return;
  const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 16; k++) {
    if (((w4[k >> 2] >> (8 * (k & 3))) & 0xffu) == 0x0au) starts[++pos] = base + k + 1;
  }
}

using dec::parse_decimal;
using dec::skip_space;

// FakeLLVertex + FakeLL2XYZ (source/GPISPPolygon.cuh:72,82-92) with the
// correctly rounded sin and cos; the host's query conversion in the same
// operation order.
__device__ __forceinline__ V3 fake_ll_to_xyz(double lo, double la) {
  const double fakela = la + 90;
  const double realla = fakela > 180 ? 270 - fakela : fakela - 90;
  const double a = realla * kPi / 180.0;
  const double o = lo * kPi / 180.0;
  double sa, ca, so, co;
  trig::sincos_cr(a, &sa, &ca);
  trig::sincos_cr(o, &so, &co);
  return v3(ca * so, sa, ca * co);
}

__device__ __forceinline__ V3 ll_to_xyz_dev(double lo, double la) {
  const double a = la * kPi / 180.0;
  const double o = lo * kPi / 180.0;
  double sa, ca, so, co;
  trig::sincos_cr(a, &sa, &ca);
  trig::sincos_cr(o, &so, &co);
  return v3(ca * so, sa, ca * co);
}

__device__ __forceinline__ int64_t line_end(const int64_t* __restrict__ starts, int64_t line, int64_t nl,
                                            int64_t bytes) {
  return line < nl ? starts[line + 1] - 1 : bytes;
}

// One lane per query line: "%lf,%lf" (longitude, latitude) -> xyz.
__global__ void __launch_bounds__(kLoadThreads) k_parse_csv(const uint8_t* __restrict__ text, int64_t bytes,
                                                            const int64_t* __restrict__ starts, int64_t nl,
                                                            int64_t n, double* __restrict__ xyz,
                                                            int32_t* __restrict__ flags,
                                                            int32_t* __restrict__ n_flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t end = line_end(starts, t, nl, bytes);
  int64_t i = skip_space(text, starts[t], end);
  double lo = 0.0, la = 0.0;
  bool ok = parse_decimal(text, i, end, lo);
  ok = ok && i < end && text[i] == ',';
  if (ok) {
    i = skip_space(text, i + 1, end);
    ok = parse_decimal(text, i, end, la);
  }
  if (ok) {
    store3(xyz + 3 * t, fake_ll_to_xyz(lo, la));
  } else {
    flags[atomicAdd(n_flag, 1)] = (int32_t)t;
  }
}

// One lane per vertex line of a ring: "%lf %lf" -> lon, lat and xyz
// (LLVertex::LL2XYZ, source/GPISPPolygon.cuh:21-30).  Vertex j of the ring
// is line first + j; out[j] receives it.
__global__ void __launch_bounds__(kLoadThreads) k_parse_vertices(const uint8_t* __restrict__ text,
                                                                 int64_t bytes,
                                                                 const int64_t* __restrict__ starts,
                                                                 int64_t nl, int64_t first, int64_t n,
                                                                 double* __restrict__ lon,
                                                                 double* __restrict__ lat,
                                                                 double* __restrict__ xyz,
                                                                 int32_t* __restrict__ flags,
                                                                 int32_t* __restrict__ n_flag) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t line = first + j;
  const int64_t end = line_end(starts, line, nl, bytes);
  int64_t i = skip_space(text, starts[line], end);
  double lo = 0.0, la = 0.0;
  bool ok = parse_decimal(text, i, end, lo);
  if (ok) {
    i = skip_space(text, i, end);
    ok = parse_decimal(text, i, end, la);
  }
  if (ok) {
    lon[j] = lo;
    lat[j] = la;
    store3(xyz + 3 * j, ll_to_xyz_dev(lo, la));
  } else {
    flags[atomicAdd(n_flag, 1)] = (int32_t)j;
  }
}

// LLVertex::LL2XYZ of n points (lon, lat interleaved per point in ll).
__global__ void __launch_bounds__(kLoadThreads) k_ll_to_xyz(const double* __restrict__ ll, int64_t n,
                                                            double* __restrict__ xyz) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  store3(xyz + 3 * j, ll_to_xyz_dev(ll[2 * j], ll[2 * j + 1]));
}

// Byte ranges of listed lines (index = first + flags[k]).
__global__ void k_flag_ranges(const int32_t* __restrict__ flags, int n, const int64_t* __restrict__ starts,
                              int64_t nl, int64_t bytes, int64_t first, int64_t* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t line = first + flags[k];
  out[2 * k] = starts[line];
  out[2 * k + 1] = line_end(starts, line, nl, bytes);
}

__global__ void k_scatter_xyz(const int32_t* __restrict__ idx, const double* __restrict__ val, int n,
                              double* __restrict__ xyz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = idx[k];
  xyz[3 * i] = val[3 * k];
  xyz[3 * i + 1] = val[3 * k + 1];
  xyz[3 * i + 2] = val[3 * k + 2];
}

template <typename T>
int grow(gahhg_ctx* c, T** p, int64_t* cap, int64_t need) {
  if (need <= *cap) return GAHHG_OK;
  (void)hipStreamSynchronize(c->stream);
  dfree(*p);
  *p = nullptr;
  *cap = 0;
  const int64_t n = std::max<int64_t>(need, 1024);
  if (dalloc(c, (void**)p, n * (int64_t)sizeof(T))) return GAHHG_E_NOMEM;
  *cap = n;
  return GAHHG_OK;
}

double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  return (double)ms;
}

// Text buffer for `bytes` bytes, padded with zeros to whole blocks.
int text_buffer(gahhg_ctx* c, int64_t bytes, int64_t* blocks) {
  TextLoad& L = c->tl;
  *blocks = std::max<int64_t>((bytes + kBlockBytes - 1) / kBlockBytes, 1);
  const int64_t padded = *blocks * kBlockBytes;
  int rc = grow(c, &L.text, &L.text_cap, padded);
  if (rc) return rc;
  return hip_check(c, hipMemsetAsync(L.text + bytes, 0, (size_t)(padded - bytes), c->stream), "text pad");
}

// Splits the device text into lines: L.starts[0..nl] and the line count.
int split_lines(gahhg_ctx* c, int64_t bytes, int64_t blocks, bool last_nl, int64_t* nl, int64_t* lines) {
  TextLoad& L = c->tl;
  if (blocks > INT32_MAX) return set_err(c, GAHHG_E_LIMIT, "text too large");
  int rc = grow(c, &L.counts, &L.counts_cap, 2 * blocks + 2);
  if (rc) return rc;
  hipLaunchKernelGGL(k_nl_count, dim3((unsigned)blocks), dim3(kLoadThreads), 0, c->stream,
                     (const uint8_t*)L.text, L.counts);
  int32_t* offs = L.counts + blocks;
  int32_t* total = L.counts + 2 * blocks;
  rc = scan_i32(c, L.counts, offs, blocks, nullptr, total, 0);
  if (rc) return rc;
  int32_t h_total = 0;
  rc = hip_check(c, hipMemcpyAsync(&h_total, total, 4, hipMemcpyDeviceToHost, c->stream), "line count");
  if (rc) return rc;
  rc = hip_check(c, hipStreamSynchronize(c->stream), "line count");
  if (rc) return rc;
  *nl = h_total;
  *lines = h_total + ((bytes > 0 && !last_nl) ? 1 : 0);
  if (*lines > INT32_MAX - 1) return set_err(c, GAHHG_E_LIMIT, "more than 2^31 lines");
  rc = grow(c, &L.starts, &L.starts_cap, *nl + 2);
  if (rc) return rc;
  hipLaunchKernelGGL(k_nl_write, dim3((unsigned)blocks), dim3(kLoadThreads), 0, c->stream,
                     (const uint8_t*)L.text, (const int32_t*)offs, L.starts);
  return hip_check(c, hipGetLastError(), "line split");
}

// The listed lines' text, in ascending line order (index = first + flag).
int listed_lines(gahhg_ctx* c, int n_flag, int64_t nl, int64_t bytes, int64_t first,
                 std::vector<int32_t>* idx, std::vector<std::string>* text) {
  TextLoad& L = c->tl;
  idx->resize((size_t)n_flag);
  int rc = hip_check(c, hipMemcpyAsync(idx->data(), L.flags, 4 * (size_t)n_flag, hipMemcpyDeviceToHost,
                                       c->stream), "listed lines");
  if (rc) return rc;
  rc = hip_check(c, hipStreamSynchronize(c->stream), "listed lines");
  if (rc) return rc;
  std::sort(idx->begin(), idx->end());
  rc = hip_check(c, hipMemcpyAsync(L.flags, idx->data(), 4 * (size_t)n_flag, hipMemcpyHostToDevice, c->stream),
                 "listed lines");
  if (rc) return rc;
  int64_t* d_rng = nullptr;
  if (dalloc(c, (void**)&d_rng, 16 * (int64_t)n_flag)) return GAHHG_E_NOMEM;
  hipLaunchKernelGGL(k_flag_ranges, dim3((unsigned)((n_flag + 255) / 256)), dim3(256), 0, c->stream,
                     (const int32_t*)L.flags, n_flag, (const int64_t*)L.starts, nl, bytes, first, d_rng);
  std::vector<int64_t> rng(2 * (size_t)n_flag);
  rc = hip_check(c, hipMemcpyAsync(rng.data(), d_rng, 16 * (size_t)n_flag, hipMemcpyDeviceToHost, c->stream),
                 "listed ranges");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "listed ranges");
  dfree(d_rng);
  if (rc) return rc;
  text->resize((size_t)n_flag);
  for (int k = 0; k < n_flag; k++) {
    const int64_t a = rng[2 * k], b = rng[2 * k + 1];
    (*text)[k].assign((size_t)(b - a), '\0');
    if (b > a) {
      rc = hip_check(c, hipMemcpy(&(*text)[k][0], L.text + a, (size_t)(b - a), hipMemcpyDeviceToHost),
                     "listed line text");
      if (rc) return rc;
    }
  }
  return GAHHG_OK;
}

// Reads [off, off + n) of the file into dst with `threads` parallel preads
// (copies out of the page cache run at one core's memory bandwidth each).
bool read_parallel(int fd, uint8_t* dst, int64_t off, int64_t n, int threads) {
  const int64_t slice = std::max<int64_t>((n + threads - 1) / threads, 1 << 20);
  std::vector<std::thread> pool;
  std::vector<char> ok((size_t)threads, 1);
  for (int t = 0; t < threads; t++) {
    const int64_t a = (int64_t)t * slice, b = std::min<int64_t>(n, a + slice);
    if (a >= b) break;
    pool.emplace_back([=, &ok]() {
      int64_t done = a;
      while (done < b) {
        const ssize_t r = pread(fd, dst + done, (size_t)(b - done), (off_t)(off + done));
        if (r <= 0) {
          ok[(size_t)t] = 0;
          return;
        }
        done += r;
      }
    });
  }
  for (std::thread& th : pool) th.join();
  for (char v : ok)
    if (!v) return false;
  return true;
}

// Copies `bytes` bytes of an open file to the device text buffer through
// two pinned staging buffers: while one chunk is copied to the device the
// next is read (by several threads) into the other.  *last is the file's
// last byte (0 if empty).
int upload_file(gahhg_ctx* c, FILE* f, int64_t bytes, uint8_t* last) {
  TextLoad& L = c->tl;
  for (int k = 0; k < 2; k++) {
    if (!L.stage[k]) {
      int rc = hip_check(c, hipHostMalloc(&L.stage[k], (size_t)kStageBytes, hipHostMallocDefault), "staging");
      if (rc) return rc;
      rc = hip_check(c, hipEventCreateWithFlags(&L.stage_ev[k], hipEventDisableTiming), "staging event");
      if (rc) return rc;
    }
  }
  const int fd = fileno(f);
  const int threads = std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
  *last = 0;
  int64_t off = 0;
  int k = 0;
  while (off < bytes) {
    const int64_t n = std::min<int64_t>(kStageBytes, bytes - off);
    int rc = hip_check(c, hipEventSynchronize(L.stage_ev[k]), "staging wait");
    if (rc) return rc;
    if (!read_parallel(fd, (uint8_t*)L.stage[k], off, n, threads)) return set_err(c, GAHHG_E_IO, "short read");
    *last = ((uint8_t*)L.stage[k])[n - 1];
    rc = hip_check(c, hipMemcpyAsync(L.text + off, L.stage[k], (size_t)n, hipMemcpyHostToDevice, c->stream),
                   "text upload");
    if (rc) return rc;
    rc = hip_check(c, hipEventRecord(L.stage_ev[k], c->stream), "staging event");
    if (rc) return rc;
    off += n;
    k ^= 1;
  }
  return GAHHG_OK;
}

int open_sized(gahhg_ctx* c, const char* path, FILE** f, int64_t* bytes) {
  if (!path) return set_err(c, GAHHG_E_ARG, "null path");
  *f = fopen(path, "rb");
  if (!*f) return set_err(c, GAHHG_E_IO, "Can not open %s", path);
  struct stat sb;
  if (fstat(fileno(*f), &sb) != 0) {
    fclose(*f);
    return set_err(c, GAHHG_E_IO, "Can not stat %s", path);
  }
  *bytes = (int64_t)sb.st_size;
  return GAHHG_OK;
}

}  // namespace

// Query points: device text (L.text, `bytes` long) -> L.xyz.
int parse_query_text(gahhg_ctx* c, int64_t bytes, int64_t blocks, bool last_nl, const char* name,
                     int64_t* n_out, gahhg_load_info* info) {
  TextLoad& L = c->tl;
  int64_t nl = 0, n = 0;
  (void)hipEventRecord(c->ev[2], c->stream);
  int rc = split_lines(c, bytes, blocks, last_nl, &nl, &n);
  if (rc) return rc;
  rc = grow(c, &L.xyz, &L.xyz_cap, 3 * n);
  if (!rc) rc = grow(c, &L.flags, &L.flags_cap, n + 1);
  if (!rc) rc = grow(c, &L.misc, &L.misc_cap, 4);
  if (rc) return rc;
  rc = fill_i32(c, L.misc, 1, 0);
  if (rc) return rc;
  if (n > 0)
    hipLaunchKernelGGL(k_parse_csv, dim3((unsigned)((n + kLoadThreads - 1) / kLoadThreads)), dim3(kLoadThreads), 0,
                       c->stream, (const uint8_t*)L.text, bytes, (const int64_t*)L.starts, nl, n, L.xyz, L.flags,
                       L.misc);
  (void)hipEventRecord(c->ev[3], c->stream);
  int32_t n_flag = 0;
  rc = hip_check(c, hipMemcpyAsync(&n_flag, L.misc, 4, hipMemcpyDeviceToHost, c->stream), "parse");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "parse");
  if (rc) return rc;
  const double parse_ms = ev_ms(c->ev[2], c->ev[3]);
  if (n_flag > 0) {
    std::vector<int32_t> idx;
    std::vector<std::string> lines;
    rc = listed_lines(c, n_flag, nl, bytes, 0, &idx, &lines);
    if (rc) return rc;
    std::vector<double> val(3 * (size_t)n_flag);
    for (int k = 0; k < n_flag; k++) {
      double lo = 0.0, la = 0.0;
      if (sscanf(lines[k].c_str(), "%lf,%lf", &lo, &la) != 2)
        return set_err(c, GAHHG_E_FORMAT, "%s: line %d: wrong format", name, idx[k] + 1);
      query_lonlat_to_xyz(&lo, &la, 1, &val[3 * (size_t)k]);
    }
    double* d_val = nullptr;
    if (dalloc(c, (void**)&d_val, 24 * (int64_t)n_flag)) return GAHHG_E_NOMEM;
    rc = hip_check(c, hipMemcpyAsync(d_val, val.data(), val.size() * 8, hipMemcpyHostToDevice, c->stream),
                   "listed values");
    if (!rc) {
      hipLaunchKernelGGL(k_scatter_xyz, dim3((unsigned)((n_flag + 255) / 256)), dim3(256), 0, c->stream,
                         (const int32_t*)L.flags, (const double*)d_val, n_flag, L.xyz);
      rc = hip_check(c, hipStreamSynchronize(c->stream), "listed values");
    }
    dfree(d_val);
    if (rc) return rc;
  }
  *n_out = n;
  if (info) {
    info->lines = n;
    info->bytes = bytes;
    info->host_lines = n_flag;
    info->parse_ms = parse_ms;
  }
  return GAHHG_OK;
}

int load_query_csv_device(gahhg_ctx* c, const char* path, int64_t* n, gahhg_load_info* info) {
  FILE* f = nullptr;
  int64_t bytes = 0;
  const double t0 = now_ms();
  int rc = open_sized(c, path, &f, &bytes);
  if (rc) return rc;
  int64_t blocks = 0;
  rc = text_buffer(c, bytes, &blocks);
  uint8_t last = 0;
  if (!rc) rc = upload_file(c, f, bytes, &last);
  fclose(f);
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "text upload");
  if (rc) return rc;
  const double t1 = now_ms();
  rc = parse_query_text(c, bytes, blocks, last == '\n', path, n, info);
  if (rc) return rc;
  if (info) {
    info->read_ms = t1 - t0;
    info->total_ms = now_ms() - t0;
  }
  return GAHHG_OK;
}

int parse_query_csv_device(gahhg_ctx* c, const char* text, int64_t bytes, int64_t* n, gahhg_load_info* info) {
  const double t0 = now_ms();
  int64_t blocks = 0;
  int rc = text_buffer(c, bytes, &blocks);
  if (!rc && bytes > 0)
    rc = hip_check(c, hipMemcpyAsync(c->tl.text, text, (size_t)bytes, hipMemcpyHostToDevice, c->stream),
                   "text upload");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "text upload");
  if (rc) return rc;
  const double t1 = now_ms();
  rc = parse_query_text(c, bytes, blocks, bytes > 0 && text[bytes - 1] == '\n', "text", n, info);
  if (rc) return rc;
  if (info) {
    info->read_ms = t1 - t0;
    info->total_ms = now_ms() - t0;
  }
  return GAHHG_OK;
}

// Polygon TXT with the vertex lines parsed on the device.  The header and
// ring-count lines (a handful) are read on the host from the device's line
// offsets; the structure checks and messages are the host loader's.
int load_polygon_txt_device(gahhg_ctx* c, const char* path, Polygon* out, gahhg_load_info* info) {
  TextLoad& L = c->tl;
  FILE* f = nullptr;
  int64_t bytes = 0;
  const double t0 = now_ms();
  int rc = open_sized(c, path, &f, &bytes);
  if (rc) return rc;
  int64_t blocks = 0;
  rc = text_buffer(c, bytes, &blocks);
  uint8_t last = 0;
  if (!rc) rc = upload_file(c, f, bytes, &last);
  fclose(f);
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "text upload");
  if (rc) return rc;
  const double t1 = now_ms();
  int64_t nl = 0, nlines = 0;
  rc = split_lines(c, bytes, blocks, last == '\n', &nl, &nlines);
  if (rc) return rc;
  auto fail = [&](int64_t line, const char* what) {
    return set_err(c, GAHHG_E_FORMAT, "%s: line %lld: %s", path, (long long)(line + 1), what);
  };
  // text of one line, read from the device
  auto read_line = [&](int64_t line, std::string* s) -> int {
    int64_t se[2] = {0, 0};
    int r = hip_check(c, hipMemcpyAsync(se, L.starts + line, line < nl ? 16 : 8, hipMemcpyDeviceToHost, c->stream),
                      "header line");
    if (!r) r = hip_check(c, hipStreamSynchronize(c->stream), "header line");
    if (r) return r;
    const int64_t a = se[0], b = line < nl ? se[1] - 1 : bytes;
    s->assign((size_t)(b - a), '\0');
    if (b > a) r = hip_check(c, hipMemcpy(&(*s)[0], L.text + a, (size_t)(b - a), hipMemcpyDeviceToHost), "line");
    return r;
  };
  std::string s;
  int holes = 0;
  if (nlines < 1) return fail(0, "wrong format");
  rc = read_line(0, &s);
  if (rc) return rc;
  if (sscanf(s.c_str(), "POLYGON 1 %d", &holes) != 1 || holes < 0) return fail(0, "wrong format");
  struct RingLines {
    int64_t first;  // line of vertex 0
    int32_t cnt;    // lines, closing vertex included
  };
  std::vector<RingLines> rings;
  int64_t line = 1;
  int64_t total = 0;
  for (int r = 0; r <= holes; r++) {
    if (line >= nlines) return fail(line, "wrong format");
    rc = read_line(line, &s);
    if (rc) return rc;
    int cnt = 0, tmp = 0;
    const bool ok = r == 0 ? sscanf(s.c_str(), "0 %d", &cnt) == 1 : sscanf(s.c_str(), "%d %d", &tmp, &cnt) == 2;
    if (!ok) return fail(line, "wrong format");
    if (cnt < 4) return fail(line, "number of edges should be greater than 3");
    if (line + cnt >= nlines) return fail(nlines - 1, "wrong format");
    rings.push_back(RingLines{line + 1, cnt});
    total += cnt;
    line += 1 + cnt;
  }
  rc = grow(c, &L.xyz, &L.xyz_cap, 5 * total);  // lon, lat, xyz per vertex line
  if (!rc) rc = grow(c, &L.flags, &L.flags_cap, total + 1);
  if (!rc) rc = grow(c, &L.misc, &L.misc_cap, 4);
  if (rc) return rc;
  double* d_lon = L.xyz;
  double* d_lat = L.xyz + total;
  double* d_v = L.xyz + 2 * total;
  std::vector<double> lo((size_t)total), la((size_t)total), xyz(3 * (size_t)total);
  (void)hipEventRecord(c->ev[2], c->stream);
  int64_t off = 0;
  int64_t host_lines = 0;
  for (const RingLines& rl : rings) {
    rc = fill_i32(c, L.misc, 1, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_parse_vertices, dim3((unsigned)((rl.cnt + kLoadThreads - 1) / kLoadThreads)),
                       dim3(kLoadThreads), 0, c->stream, (const uint8_t*)L.text, bytes, (const int64_t*)L.starts, nl,
                       rl.first, (int64_t)rl.cnt, d_lon + off, d_lat + off, d_v + 3 * off, L.flags, L.misc);
    int32_t n_flag = 0;
    rc = hip_check(c, hipMemcpyAsync(&n_flag, L.misc, 4, hipMemcpyDeviceToHost, c->stream), "vertices");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "vertices");
    if (rc) return rc;
    std::vector<int32_t> idx;
    std::vector<std::string> lines;
    if (n_flag > 0) {
      rc = listed_lines(c, n_flag, nl, bytes, rl.first, &idx, &lines);
      if (rc) return rc;
    }
    rc = hip_check(c, hipMemcpyAsync(&lo[off], d_lon + off, 8 * (size_t)rl.cnt, hipMemcpyDeviceToHost, c->stream),
                   "vertices");
    if (!rc) rc = hip_check(c, hipMemcpyAsync(&la[off], d_lat + off, 8 * (size_t)rl.cnt, hipMemcpyDeviceToHost,
                                              c->stream), "vertices");
    if (!rc) rc = hip_check(c, hipMemcpyAsync(&xyz[3 * off], d_v + 3 * off, 24 * (size_t)rl.cnt,
                                              hipMemcpyDeviceToHost, c->stream), "vertices");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "vertices");
    if (rc) return rc;
    for (int k = 0; k < n_flag; k++) {
      double x = 0.0, y = 0.0;
      if (sscanf(lines[k].c_str(), "%lf %lf", &x, &y) != 2) return fail(rl.first + idx[k], "wrong format");
      const int64_t v = off + idx[k];
      lo[v] = x;
      la[v] = y;
      lonlat_to_xyz(&x, &y, 1, &xyz[3 * v]);
    }
    host_lines += n_flag;
    off += rl.cnt;
  }
  (void)hipEventRecord(c->ev[3], c->stream);
  (void)hipStreamSynchronize(c->stream);
  // each ring's closing line is dropped by the assembly, as the host loader does
  std::vector<int64_t> ring_off(1, 0);
  for (const RingLines& rl : rings) ring_off.push_back(ring_off.back() + rl.cnt);
  rc = polygon_from_lonlat_xyz(lo.data(), la.data(), xyz.data(), ring_off.data(), (int32_t)rings.size(), out,
                               &c->err);
  if (rc) return rc;
  if (info) {
    info->lines = nlines;
    info->bytes = bytes;
    info->host_lines = host_lines;
    info->read_ms = t1 - t0;
    info->parse_ms = ev_ms(c->ev[2], c->ev[3]);
    info->total_ms = now_ms() - t0;
  }
  return GAHHG_OK;
}

int lonlat_to_xyz_gpu(gahhg_ctx* c, const double* lon, const double* lat, int64_t n, double* xyz) {
  if (n <= 0) return GAHHG_OK;
  std::vector<double> ll((size_t)(2 * n));
  for (int64_t j = 0; j < n; j++) {
    ll[(size_t)(2 * j)] = lon[j];
    ll[(size_t)(2 * j + 1)] = lat[j];
  }
  double* d = nullptr;
  if (dalloc(c, (void**)&d, n * 40)) return GAHHG_E_NOMEM;
  int rc = hip_check(c, hipMemcpyAsync(d, ll.data(), n * 16, hipMemcpyHostToDevice, c->stream), "vertex upload");
  if (!rc) {
    hipLaunchKernelGGL(k_ll_to_xyz, dim3((unsigned)((n + kLoadThreads - 1) / kLoadThreads)), dim3(kLoadThreads), 0,
                       c->stream, (const double*)d, n, d + 2 * n);
    rc = hip_check(c, hipGetLastError(), "vertex conversion");
  }
  if (!rc) rc = hip_check(c, hipMemcpyAsync(xyz, d + 2 * n, n * 24, hipMemcpyDeviceToHost, c->stream), "vertex D2H");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "vertex conversion");
  dfree(d);
  return rc;
}

void free_text_load(gahhg_ctx* c) {
  TextLoad& L = c->tl;
  dfree(L.text);
  dfree(L.starts);
  dfree(L.xyz);
  dfree(L.flags);
  dfree(L.counts);
  dfree(L.misc);
  for (int k = 0; k < 2; k++) {
    if (L.stage[k]) (void)hipHostFree(L.stage[k]);
    if (L.stage_ev[k]) (void)hipEventDestroy(L.stage_ev[k]);
  }
  L = TextLoad();
}

// Loads this file's code object onto the current device (context set-up).
void load_code_load() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_nl_count));
}

}  // namespace gahhg
