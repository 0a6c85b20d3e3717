// Learned preconditioner network: a 4-layer 3x3 convolutional network over
// the time slices of the lattice (input 8 channels = 7 stencil coefficients +
// 0, hidden 32 channels, output 32 channels of which 22 are the correction of
// the incomplete Cholesky factor), BF16 operands with FP32 accumulation on
// the CDNA matrix cores (v_mfma_f32_16x16x16_bf16), and its training
// kernels: backward data, weight and bias gradients, Adam.  No reference
// counterpart (the reference has no iterative solver); the correction it
// produces is applied by spde_ic_correct.
//
// Layouts.  Activations: pixel-major (NHWC), pixel p = x + M (y + N t), the
// channels of a pixel contiguous (FP32).  Master weights per layer: W[k][co]
// FP32 row-major with k = tap * cin + ci, tap = 3 (dy + 1) + (dx + 1), and
// bias[32].  The matrix-core operand of a layer is packed once per update as
// BF16 fragments: frag[((kc * 2 + h) * 64 + lane) * 4 + j] =
// W[16 kc + 4 (lane / 16) + j][16 h + lane % 16] (zero for k >= 9 cin), the B
// operand of the 16x16x16 instruction for k-chunk kc and output half h.
//
// Each wave computes a 16-pixel x 32-channel tile: A rows are pixels (lane
// % 16), A k-columns 4 (lane / 16) + j; D rows 4 (lane / 16) + j, columns lane % 16.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "internal.hpp"

namespace spde {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCh = 32;  // hidden and padded output channels

// float -> bf16, round to nearest even (finite inputs)
__device__ __host__ inline unsigned short to_bf16(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// Convolution of one layer.  in: P x cin; mask (optional, P x cin): the input
// is in * (mask > 0) (backward data through a ReLU); out: P x 32, out = act(D +
// bias) with act ReLU when relu, bias optional.  nk = number of 16-wide k chunks.
__global__ __launch_bounds__(256) void conv_kernel(int M, int N, int64_t P, int cin, int nk,
                                                   const float* __restrict__ in, const float* __restrict__ mask,
                                                   const unsigned short* __restrict__ frag,
                                                   const float* __restrict__ bias, int relu, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t p0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (p0 >= P) return;  // whole wave
  const int64_t MN = (int64_t)M * N;
  const int64_t p = p0 + (lane & 15);
  const int kq = lane >> 4;
  const bool pv = p < P;
  int x = 0, y = 0;
  int64_t tb = 0;
  if (pv) {
    const int64_t t = p / MN, rem = p - t * MN;
    y = (int)(rem / M);
    x = (int)(rem - (int64_t)y * M);
    tb = t * MN;
  }
  f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < nk; ++kc) {
    const int k0 = 16 * kc + 4 * kq;
    const int tap = k0 / cin, ci = k0 - tap * cin;
    bf16x4 a = {0, 0, 0, 0};
    if (pv && tap < 9) {
      const int nx = x + tap % 3 - 1, ny = y + tap / 3 - 1;
      if (nx >= 0 && nx < M && ny >= 0 && ny < N) {
        const int64_t q = (tb + (int64_t)ny * M + nx) * cin + ci;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v = in[q + j];
          if (mask && !(mask[q + j] > 0.f)) v = 0.f;
          a[j] = (short)to_bf16(v);
        }
      }
    }
    const bf16x4 b0 = *(const bf16x4*)(frag + ((int64_t)(kc * 2 + 0) * 64 + lane) * 4);
    const bf16x4 b1 = *(const bf16x4*)(frag + ((int64_t)(kc * 2 + 1) * 64 + lane) * 4);
    acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc1, 0, 0, 0);
  }
  const int co = lane & 15;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t pp = p0 + 4 * kq + j;
    if (pp >= P) continue;
    float v0 = acc0[j] + (bias ? bias[co] : 0.f), v1 = acc1[j] + (bias ? bias[co + 16] : 0.f);
    if (relu) {
      v0 = v0 > 0.f ? v0 : 0.f;
      v1 = v1 > 0.f ? v1 : 0.f;
    }
    out[pp * kCh + co] = v0;
    out[pp * kCh + co + 16] = v1;
  }
}

// Partial weight gradients over a chunk of pixels:
// part[chunk][k][co] = sum_{p in chunk} X[p + off(tap)][ci] dZ[p][co] with
// dZ = G (G * (Y > 0) when mask), both operands rounded to BF16 as in the
// forward pass.  One wave per (chunk, 16 k-rows, 16 output columns).
__global__ __launch_bounds__(64) void conv_wgrad_kernel(int M, int N, int64_t P, int cin, int64_t chunk,
                                                        const float* __restrict__ X, const float* __restrict__ G,
                                                        const float* __restrict__ Y, float* __restrict__ part) {
  const int lane = threadIdx.x;
  const int nkt = (9 * cin + 15) / 16;
  const int tile = blockIdx.y;       // k-tile * 2 + column half
  const int kt = tile >> 1, h = tile & 1;
  const int64_t c0 = (int64_t)blockIdx.x * chunk;
  const int64_t c1 = c0 + chunk < P ? c0 + chunk : P;
  const int64_t MN = (int64_t)M * N;
  const int krow = 16 * kt + (lane & 15);
  const int tap = krow / cin, ci = krow - tap * cin;
  const int co = 16 * h + (lane & 15);
  const int kq = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int64_t p0 = c0; p0 < c1; p0 += 16) {
    bf16x4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t p = p0 + 4 * kq + j;
      if (p >= c1) continue;
      if (tap < 9) {
        const int64_t t = p / MN, rem = p - t * MN;
        const int y = (int)(rem / M), x = (int)(rem - (int64_t)y * M);
        const int nx = x + tap % 3 - 1, ny = y + tap / 3 - 1;
        if (nx >= 0 && nx < M && ny >= 0 && ny < N) a[j] = (short)to_bf16(X[(t * MN + (int64_t)ny * M + nx) * cin + ci]);
      }
      float g = G[p * kCh + co];
      if (Y && !(Y[p * kCh + co] > 0.f)) g = 0.f;
      b[j] = (short)to_bf16(g);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  // D rows (k) 4 kq + j, column lane & 15
  float* dst = part + (int64_t)blockIdx.x * (nkt * 16) * kCh;
#pragma unroll
  for (int j = 0; j < 4; ++j) dst[(int64_t)(16 * kt + 4 * kq + j) * kCh + co] = acc[j];
  (void)nkt;
}

// Bias gradient partials: part[chunk][co] = sum_{p in chunk} dZ[p][co].
__global__ __launch_bounds__(64) void conv_bgrad_kernel(int64_t P, int64_t chunk, const float* __restrict__ G,
                                                        const float* __restrict__ Y, float* __restrict__ part) {
  const int co = threadIdx.x;
  if (co >= kCh) return;
  const int64_t c0 = (int64_t)blockIdx.x * chunk;
  const int64_t c1 = c0 + chunk < P ? c0 + chunk : P;
  float s = 0.f;
  for (int64_t p = c0; p < c1; ++p) {
    float g = G[p * kCh + co];
    if (Y && !(Y[p * kCh + co] > 0.f)) g = 0.f;
    s += g;
  }
  part[(int64_t)blockIdx.x * kCh + co] = s;
}

// grad[e] = sum over chunks of part[c][e] in chunk order (deterministic);
// grad may be part itself (each thread reads its column before writing it).
__global__ void chunk_sum_kernel(int nchunk, int64_t width, const float* part, float* grad) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= width) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * width + e];
  grad[e] = s;
}

// BF16 fragments of the forward operand (transposed = 0) or of the
// backward-data operand (transposed = 1: W'[tap' * 32 + co][ci] =
// W[(8 - tap') * cin + ci][co], output channels ci, padded to 32).
__global__ void pack_kernel(int cin, int transposed, const float* __restrict__ W, unsigned short* __restrict__ frag) {
  const int kin = transposed ? 9 * kCh : 9 * cin;  // k extent of the packed operand
  const int nk = (kin + 15) / 16;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nk * 2 * 64 * 4) return;
  const int j = e & 3, lane = (e >> 2) & 63, h = (e >> 8) & 1, kc = e >> 9;
  const int k = 16 * kc + 4 * (lane >> 4) + j;
  const int col = 16 * h + (lane & 15);
  float v = 0.f;
  if (k < kin) {
    if (!transposed) {
      v = W[(int64_t)k * kCh + col];
    } else {
      const int tp = k / kCh, co = k - tp * kCh;  // packed row = (tap', co); packed column = ci
      if (col < cin) v = W[(int64_t)((8 - tp) * cin + col) * kCh + co];
    }
  }
  frag[e] = to_bf16(v);
}

// Adam on FP32 parameters: m = b1 m + (1 - b1) g; v = b2 v + (1 - b2) g^2;
// theta -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps); entries with
// mask[e] == 0 stay fixed.
__global__ void adam_kernel(int64_t n, const float* __restrict__ g, const unsigned char* __restrict__ mask, float lr,
                            float b1, float b2, float eps, float c1, float c2, float* __restrict__ th,
                            float* __restrict__ m, float* __restrict__ v) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n || (mask && !mask[e])) return;
  const float ge = g[e];
  const float me = b1 * m[e] + (1.f - b1) * ge, ve = b2 * v[e] + (1.f - b2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  th[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
}

// Lane layout check of the BF16 matrix instruction: A[i][k] = 16 i + k,
// B = identity, so D = A; out[lane * 4 + j] = D fragment element j of lane.
__global__ void bf16_probe_kernel(float* out) {
  const int lane = threadIdx.x;
  bf16x4 a, b;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = 4 * (lane >> 4) + j;
    a[j] = (short)to_bf16((float)(16 * (lane & 15) + k));
    b[j] = (short)to_bf16(k == (lane & 15) ? 1.f : 0.f);
  }
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) out[lane * 4 + j] = d[j];
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace spde

using namespace spde;

extern "C" {

int spde_conv_packed_size(int cin, int transposed, int64_t* halves) {
  if ((cin != 8 && cin != 32) || !halves) {
    set_error("spde_conv_packed_size: cin must be 8 or 32");
    return SPDE_ERR_ARG;
  }
  const int kin = transposed ? 9 * kCh : 9 * cin;
  *halves = (int64_t)((kin + 15) / 16) * 2 * 64 * 4;
  return SPDE_OK;
}

int spde_conv_pack(int cin, int transposed, const float* W, unsigned short* frag, void* stream) {
  int64_t sz = 0;
  int rc = spde_conv_packed_size(cin, transposed, &sz);
  if (rc) return rc;
  if (!W || !frag) {
    set_error("spde_conv_pack: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipLaunchKernelGGL(pack_kernel, dim3(nblk(sz, 256)), dim3(256), 0, (hipStream_t)stream, cin, transposed, W, frag);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_conv_forward(int M, int N, int T, int cin, const float* in, const float* mask, const unsigned short* frag,
                      const float* bias, int relu, float* out, void* stream) {
  if (M < 1 || N < 1 || T < 1 || (cin != 8 && cin != 32) || !in || !frag || !out) {
    set_error("spde_conv_forward: invalid argument");
    return SPDE_ERR_ARG;
  }
  const int64_t P = (int64_t)M * N * T;
  const int nk = (9 * cin + 15) / 16;
  const int64_t tiles = (P + 15) / 16;
  hipLaunchKernelGGL(conv_kernel, dim3(nblk(tiles, 4)), dim3(256), 0, (hipStream_t)stream, M, N, P, cin, nk, in, mask,
                     frag, bias, relu, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_conv_wgrad_workspace(int64_t P, int cin, int64_t* floats) {
  if (P < 1 || (cin != 8 && cin != 32) || !floats) {
    set_error("spde_conv_wgrad_workspace: invalid argument");
    return SPDE_ERR_ARG;
  }
  const int64_t nchunk = P / 1024 + 1 < 256 ? P / 1024 + 1 : 256;
  const int nkt = (9 * cin + 15) / 16;
  *floats = nchunk * (int64_t)nkt * 16 * kCh + nchunk * kCh;
  return SPDE_OK;
}

int spde_conv_backward(int M, int N, int T, int cin, const float* X, const float* G, const float* Y, float* dW,
                       float* db, float* work, void* stream) {
  if (M < 1 || N < 1 || T < 1 || (cin != 8 && cin != 32) || !X || !G || !dW || !db || !work) {
    set_error("spde_conv_backward: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t P = (int64_t)M * N * T;
  const int64_t nchunk = P / 1024 + 1 < 256 ? P / 1024 + 1 : 256;
  int64_t chunk = (P + nchunk - 1) / nchunk;
  chunk = (chunk + 15) / 16 * 16;
  const int nkt = (9 * cin + 15) / 16;
  float* wpart = work;
  float* bpart = work + nchunk * (int64_t)nkt * 16 * kCh;
  hipLaunchKernelGGL(conv_wgrad_kernel, dim3((unsigned)nchunk, (unsigned)(2 * nkt)), dim3(64), 0, s, M, N, P, cin,
                     chunk, X, G, Y, wpart);
  hipLaunchKernelGGL(conv_bgrad_kernel, dim3((unsigned)nchunk), dim3(64), 0, s, P, chunk, G, Y, bpart);
  // only k < 9 cin rows are real weights
  hipLaunchKernelGGL(chunk_sum_kernel, dim3(nblk((int64_t)nkt * 16 * kCh, 256)), dim3(256), 0, s, (int)nchunk,
                     (int64_t)nkt * 16 * kCh, wpart, wpart);
  SPDE_HIP(hipMemcpyAsync(dW, wpart, sizeof(float) * 9 * cin * kCh, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(chunk_sum_kernel, dim3(1), dim3(64), 0, s, (int)nchunk, (int64_t)kCh, bpart, db);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_adam(int64_t n, const float* g, const unsigned char* mask, float lr, float b1, float b2, float eps, int step,
              float* theta, float* m, float* v, void* stream) {
  if (n < 1 || !g || !theta || !m || !v || step < 1) {
    set_error("spde_adam: invalid argument");
    return SPDE_ERR_ARG;
  }
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, n, g, mask, lr, b1, b2, eps, c1,
                     c2, theta, m, v);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_debug_bf16_probe(float* out, void* stream) {
  if (!out) {
    set_error("spde_debug_bf16_probe: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipLaunchKernelGGL(bf16_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, out);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

}  // extern "C"
