// gahhg_mlp.hip -- the learned refinement model: a 12-64-64-1 perceptron
// with ReLU hidden layers, BF16 weights and FP32 accumulation.  No
// counterpart in the reference, whose only refinement rule is the density
// stop of divGrid (source/GPISPAHHG.cu:1595-1605).
//
// Parameter vector (kMlpParams values, the same order in the FP32 master
// copy and the BF16 weights):
//   W1 [64][16]  hidden-1 weights, input columns 12..15 are padding (zero)
//   b1 [64], W2 [64][64], b2 [64], w3 [64], b3 [1]
//
// Forward: one 64-lane wave evaluates 16 rows with MFMA
// (v_mfma_f32_16x16x16_bf16).  Every layer is computed transposed,
// out^T = W * in^T, so that the accumulator layout of one layer (lane l:
// row l % 16, units 4 (l / 16) .. +3 of each 16-unit tile) is exactly the
// B-operand layout the next layer needs: activations never leave registers.
// Inputs and hidden-1 activations are rounded to BF16 as MFMA operands;
// hidden-2 and the output are FP32.
//
// Backward: the same wave recomputes its forward pass and forms the
// gradients of sum_rows dY[row] * y[row] in FP32 through LDS (the hidden
// units are few; the batch dimension is what is large).  Each wave writes
// its partial gradient; a second kernel sums the partials in wave order, so
// results do not depend on scheduling.
#include <string.h>

#include <algorithm>
#include <vector>

#include "gahhg_internal.h"

namespace gahhg {

// A named namespace (not an anonymous one) so that profilers show the
// kernels' names.
namespace mlp_kernels {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float rbf(float f) { return bf2f(f2bf(f)); }

__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
  s4 r;
  r.x = (short)f2bf(a);
  r.y = (short)f2bf(b);
  r.z = (short)f2bf(c);
  r.w = (short)f2bf(d);
  return r;
}

__device__ __forceinline__ s4 load_bf4(const uint16_t* p) {
  const uint2 u = *(const uint2*)p;  // 4 consecutive BF16, 8-byte aligned
  s4 r;
  r.x = (short)(u.x & 0xffffu);
  r.y = (short)(u.x >> 16);
  r.z = (short)(u.y & 0xffffu);
  r.w = (short)(u.y >> 16);
  return r;
}

// Forward pass of the 16 rows of one wave.  z1/z2: pre-activations without
// bias, lane layout (row l % 16, unit 16 m + 4 (l / 16) + i) for tile m.
struct WaveFwd {
  f4 z1[4];
  f4 z2[4];
  float y;  // valid in every lane of the row
};

__device__ __forceinline__ void wave_forward(const float* __restrict__ X, int64_t rows,
                                             const uint16_t* __restrict__ W, int64_t r0, WaveFwd& o) {
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, kg = lane >> 4;
  const int64_t row = r0 + s;
  const bool valid = row < rows;
  float xin[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int f = 4 * kg + i;
    xin[i] = (valid && f < kMlpIn) ? X[row * kMlpIn + f] : 0.0f;
  }
  const s4 bx = pack4(xin[0], xin[1], xin[2], xin[3]);
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  s4 h1[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const s4 aw = load_bf4(W + kMlpW1 + (16 * m + s) * 16 + 4 * kg);
    o.z1[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aw, bx, zero, 0, 0, 0);
    float h[4];
#pragma unroll
    for (int i = 0; i < 4; i++) h[i] = fmaxf(o.z1[m][i] + bf2f(W[kMlpB1 + 16 * m + 4 * kg + i]), 0.0f);
    h1[m] = pack4(h[0], h[1], h[2], h[3]);
  }
  float part = 0.0f;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    f4 acc = zero;
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const s4 aw = load_bf4(W + kMlpW2 + (16 * m + s) * 64 + 16 * ks + 4 * kg);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aw, h1[ks], acc, 0, 0, 0);
    }
    o.z2[m] = acc;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int h = 16 * m + 4 * kg + i;
      part += fmaxf(acc[i] + bf2f(W[kMlpB2 + h]), 0.0f) * bf2f(W[kMlpW3 + h]);
    }
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  o.y = part + bf2f(W[kMlpB3]);
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ X, int64_t rows,
                                                     const uint16_t* __restrict__ W, float* __restrict__ Y) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t r0 = wave * 16;
  if (r0 >= rows) return;  // wave-uniform
  WaveFwd o;
  wave_forward(X, rows, W, r0, o);
  const int lane = threadIdx.x & 63;
  if (lane < 16 && r0 + lane < rows) Y[r0 + lane] = o.y;
}

// Per-wave partial gradients, part[wave * kMlpParams + k].
__global__ void __launch_bounds__(256) k_mlp_backward(const float* __restrict__ X, int64_t rows,
                                                      const uint16_t* __restrict__ W,
                                                      const float* __restrict__ dY, float* __restrict__ part) {
  __shared__ float lds[4][4][16][64];  // per wave: z1 pre-activation, H1 (BF16 values), dZ2, dZ1
  const int wv = threadIdx.x >> 6;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t r0 = wave * 16;
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, kg = lane >> 4;
  const bool active = r0 < rows;  // wave-uniform
  float (*Z1)[64] = lds[wv][0];
  float (*H1)[64] = lds[wv][1];
  float (*D2)[64] = lds[wv][2];
  float (*D1)[64] = lds[wv][3];
  float* out = part + wave * kMlpParams;
  WaveFwd o;
  float dy = 0.0f;
  if (active) {
    wave_forward(X, rows, W, r0, o);
    dy = (r0 + s < rows) ? dY[r0 + s] : 0.0f;
    float gb3 = dy;
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int h = 16 * m + 4 * kg + i;
        const float p1 = o.z1[m][i] + bf2f(W[kMlpB1 + h]);
        Z1[s][h] = p1;
        H1[s][h] = rbf(fmaxf(p1, 0.0f));
        const float p2 = o.z2[m][i] + bf2f(W[kMlpB2 + h]);
        const float d2 = p2 > 0.0f ? dy * bf2f(W[kMlpW3 + h]) : 0.0f;
        D2[s][h] = d2;
        // w3 and b2 gradients: sums over the 16 rows of the wave
        float gw3 = dy * fmaxf(p2, 0.0f), gb2 = d2;
#pragma unroll
        for (int x = 1; x < 16; x <<= 1) {
          gw3 += __shfl_xor(gw3, x, 64);
          gb2 += __shfl_xor(gb2, x, 64);
        }
        if (s == 0) {
          out[kMlpW3 + h] = gw3;
          out[kMlpB2 + h] = gb2;
        }
      }
    }
#pragma unroll
    for (int x = 1; x < 16; x <<= 1) gb3 += __shfl_xor(gb3, x, 64);
    if (lane == 0) out[kMlpB3] = gb3;
  }
  __syncthreads();
  if (active) {
    // W2 gradient: row h2 = lane, all 64 columns
    for (int h1 = 0; h1 < 64; h1++) {
      float acc = 0.0f;
#pragma unroll
      for (int r = 0; r < 16; r++) acc += D2[r][lane] * H1[r][h1];
      out[kMlpW2 + lane * 64 + h1] = acc;
    }
    // hidden-1 gradient: unit h1 = lane, every row (straight through the
    // BF16 rounding of H1)
    for (int r = 0; r < 16; r++) {
      float acc = 0.0f;
      for (int h2 = 0; h2 < 64; h2++) acc += D2[r][h2] * bf2f(W[kMlpW2 + h2 * 64 + lane]);
      D1[r][lane] = Z1[r][lane] > 0.0f ? acc : 0.0f;
    }
    float gb1 = 0.0f;
    for (int r = 0; r < 16; r++) gb1 += D1[r][lane];
    out[kMlpB1 + lane] = gb1;
    for (int f = 0; f < 16; f++) {
      float acc = 0.0f;
      if (f < kMlpIn)
        for (int r = 0; r < 16; r++) {
          const int64_t row = r0 + r;
          const float xv = row < rows ? rbf(X[row * kMlpIn + f]) : 0.0f;
          acc += D1[r][lane] * xv;
        }
      out[kMlpW1 + lane * 16 + f] = acc;
    }
  }
}

// grad[k] = sum over waves of part[w][k], in wave order.
__global__ void k_mlp_reduce(const float* __restrict__ part, int64_t waves, float* __restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kMlpParams) return;
  float acc = 0.0f;
  for (int64_t w = 0; w < waves; w++) acc += part[w * kMlpParams + k];
  grad[k] = acc;
}

// Adam on the FP32 master copy; writes the BF16 weights.  Padding columns
// of W1 stay zero.
__global__ void k_mlp_adam(float* __restrict__ master, const float* __restrict__ grad, float* __restrict__ m1,
                           float* __restrict__ m2, int step, float lr, float b1, float b2, float eps,
                           uint16_t* __restrict__ W) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kMlpParams) return;
  if (k < kMlpB1 && (k & 15) >= kMlpIn) {
    master[k] = 0.0f;
    W[k] = 0;
    return;
  }
  const float g = grad[k];
  const float a = b1 * m1[k] + (1.0f - b1) * g;
  const float v = b2 * m2[k] + (1.0f - b2) * g * g;
  m1[k] = a;
  m2[k] = v;
  const float ah = a / (1.0f - powf(b1, (float)step));
  const float vh = v / (1.0f - powf(b2, (float)step));
  const float w = master[k] - lr * ah / (sqrtf(vh) + eps);
  master[k] = w;
  W[k] = f2bf(w);
}

}  // namespace mlp_kernels
using namespace mlp_kernels;

int mlp_forward(gahhg_ctx* c, const float* d_x, int64_t rows, const uint16_t* d_w, float* d_y) {
  if (rows <= 0) return GAHHG_OK;
  const int64_t waves = (rows + 15) / 16;
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, c->stream, d_x, rows, d_w,
                     d_y);
  return hip_check(c, hipGetLastError(), "mlp forward");
}

int mlp_backward(gahhg_ctx* c, const float* d_x, int64_t rows, const uint16_t* d_w, const float* d_dy,
                 float* d_grad) {
  if (rows <= 0) return fill_i32(c, (int32_t*)d_grad, kMlpParams, 0);
  const int64_t waves = (rows + 15) / 16;
  const int64_t blocks = (waves + 3) / 4;
  const int64_t need = blocks * 4 * kMlpParams;
  if (c->mlp_part_cap < need) {
    dfree(c->d_mlp_part);
    c->d_mlp_part = nullptr;
    c->mlp_part_cap = 0;
    if (dalloc(c, (void**)&c->d_mlp_part, need * 4)) return GAHHG_E_NOMEM;
    c->mlp_part_cap = need;
  }
  hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)blocks), dim3(256), 0, c->stream, d_x, rows, d_w, d_dy,
                     c->d_mlp_part);
  int rc = hip_check(c, hipGetLastError(), "mlp backward");
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream,
                     (const float*)c->d_mlp_part, waves, d_grad);
  return hip_check(c, hipGetLastError(), "mlp reduce");
}

int mlp_adam(gahhg_ctx* c, float* d_master, const float* d_grad, float* d_m, float* d_v, int step, float lr,
             uint16_t* d_w) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream, d_master, d_grad, d_m,
                     d_v, step, lr, 0.9f, 0.999f, 1e-8f, d_w);
  return hip_check(c, hipGetLastError(), "mlp adam");
}

// Loads this file's code object onto the current device (context set-up).
void load_code_mlp() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&mlp_kernels::k_mlp_forward));
}

}  // namespace gahhg
