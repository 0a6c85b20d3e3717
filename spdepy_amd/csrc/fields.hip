// B-spline coefficient fields of the variable-coefficient model on the
// device, and the contraction of the per-cell gradient adjoints back onto the
// basis coefficients.
//
// Fields (var_advection_var_diffusion2D.py:91-99, spat2Dtemp_regular_mesh.py:
// 251-275): kappa = exp(bs c_kappa) at the cell centres; gamma = exp(bsH c_gamma),
// vx = bsH c_vx, vy = bsH c_vy on the faces W,E,S,N; H = gamma I + v v^T per
// face; the advection faces E,N,W,S are bsA c_wx, bsA c_wy, bsA c_wx, bsA c_wy.
// The contraction is the chain rule of the same maps applied to the per-cell
// adjoints of kappa, H and the advection faces.
//
// Bases are dense row-major: bs[Ns][nb], bsH[Ns][4][nb], bsA[Ns][4][nb].
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

// Static scratch of spde_fields_contract_noalloc: the largest workspace the
// contraction can ask for (at most 240 block partial rows of 6 * 64 + 40
// doubles, plus their sum), so no call allocates.  Each device holds its own
// copy until the process ends; all calls on one device share it (calls that
// overlap on one device race).  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
constexpr int64_t kContractScratch = (240 + 1) * (6 * 64 + 40);
static __device__ __align__(256) double scratch[kContractScratch];

namespace spde {

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wave per cell: lanes take basis columns, 17 dot products reduced
// across the wave (fixed order: deterministic).
__global__ __launch_bounds__(256) void fields_kernel(int64_t Ns, int nb, const double* __restrict__ bs,
                                                     const double* __restrict__ bsH, const double* __restrict__ bsA,
                                                     const double* __restrict__ c, double* __restrict__ kappa,
                                                     double* __restrict__ H, double* __restrict__ G,
                                                     double* __restrict__ fgv) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= Ns) return;
  double dk = 0.0, dg[4] = {0, 0, 0, 0}, dvx[4] = {0, 0, 0, 0}, dvy[4] = {0, 0, 0, 0}, dw[4] = {0, 0, 0, 0};
  for (int j = lane; j < nb; j += 64) {
    dk += bs[k * nb + j] * c[j];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const double h = bsH[(k * 4 + f) * nb + j];
      dg[f] += h * c[nb + j];
      dvx[f] += h * c[2 * nb + j];
      dvy[f] += h * c[3 * nb + j];
      dw[f] += bsA[(k * 4 + f) * nb + j] * c[((f & 1) ? 5 : 4) * nb + j];
    }
  }
  dk = wave_sum(dk);
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    dg[f] = wave_sum(dg[f]);
    dvx[f] = wave_sum(dvx[f]);
    dvy[f] = wave_sum(dvy[f]);
    dw[f] = wave_sum(dw[f]);
  }
  if (lane == 0) {
    kappa[k] = exp(dk);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const double g = exp(dg[f]), vx = dvx[f], vy = dvy[f];
      double* h = H + k * 16 + 4 * f;
      h[0] = g + vx * vx;
      h[1] = vx * vy;
      h[2] = vy * vx;
      h[3] = g + vy * vy;
      G[k * 4 + f] = dw[f];
      fgv[k * 12 + 3 * f] = g;
      fgv[k * 12 + 3 * f + 1] = vx;
      fgv[k * 12 + 3 * f + 2] = vy;
    }
  }
}

constexpr int kCells = 40;  // columns of the per-cell adjoints
constexpr int kContractWidth = 6 * 64 + kCells;  // one block's partial row (doubles)

// Block partials of the contraction over a contiguous chunk of cells.
// Thread (part, j): basis column j0 + j, every 4th cell of the chunk.
// partial[b][0..6*64) = six coefficient blocks for columns j0..j0+63,
// partial[b][384..424) = column sums of the cell adjoints.
__global__ __launch_bounds__(256) void contract_kernel(int64_t Ns, int nb, int j0, int64_t chunk,
                                                       const double* __restrict__ bs, const double* __restrict__ bsH,
                                                       const double* __restrict__ bsA, const double* __restrict__ cells,
                                                       const double* __restrict__ kappa, const double* __restrict__ fgv,
                                                       double* __restrict__ partial) {
  __shared__ double red[4][6 * 64 + kCells];
  const int j = threadIdx.x & 63, part = threadIdx.x >> 6;
  const int col = j0 + j;
  const bool on = col < nb;
  const int64_t k0 = (int64_t)blockIdx.x * chunk, k1 = (k0 + chunk < Ns) ? k0 + chunk : Ns;
  double a[6] = {0, 0, 0, 0, 0, 0}, csum = 0.0;
  for (int64_t k = k0 + part; k < k1; k += 4) {
    const double* ck = cells + k * kCells;
    if (j < kCells) csum += ck[j];
    if (!on) continue;
    const double wk = ck[0] * kappa[k];
    a[0] += bs[k * nb + col] * wk;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const double h00 = ck[1 + 4 * f], h01 = ck[2 + 4 * f], h10 = ck[3 + 4 * f], h11 = ck[4 + 4 * f];
      const double g = fgv[k * 12 + 3 * f], vx = fgv[k * 12 + 3 * f + 1], vy = fgv[k * 12 + 3 * f + 2];
      const double cross = h01 + h10;
      const double bh = bsH[(k * 4 + f) * nb + col];
      a[1] += bh * ((h00 + h11) * g);
      a[2] += bh * (2.0 * vx * h00 + vy * cross);
      a[3] += bh * (2.0 * vy * h11 + vx * cross);
    }
    a[4] += bsA[(k * 4 + 0) * nb + col] * ck[17] + bsA[(k * 4 + 2) * nb + col] * ck[19];
    a[5] += bsA[(k * 4 + 1) * nb + col] * ck[22] + bsA[(k * 4 + 3) * nb + col] * ck[24];
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) red[part][q * 64 + j] = a[q];
  if (j < kCells) red[part][6 * 64 + j] = csum;
  __syncthreads();
  for (int t = threadIdx.x; t < 6 * 64 + kCells; t += 256)
    partial[(int64_t)blockIdx.x * (6 * 64 + kCells) + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

// Blocks of the contraction: chunks of at least 64 cells, at most 240 blocks.
inline int64_t contract_blocks(int64_t Ns) { return (Ns + 63) / 64 < 240 ? (Ns + 63) / 64 : 240; }

// out[t] = sum over blocks of partial[b][t] in block order (deterministic).
__global__ void contract_sum_kernel(int nblk, const double* __restrict__ partial, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * 64 + kCells) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * (6 * 64 + kCells) + t];
  out[t] = s;
}


// Contraction shared by both entry points.  work: partials of up to 240
// blocks, then the summed row (kContractWidth doubles each).
int contract_impl(int64_t Ns, int nb, const double* bs, const double* bsH, const double* bsA, const double* cells,
                  const double* kappa, const double* fgv, double* work, double* out, hipStream_t s) {
  const int64_t nblk = contract_blocks(Ns);
  const int64_t chunk = (Ns + nblk - 1) / nblk;
  double* partial = work;
  double* sums = work + nblk * kContractWidth;
  const int width = kContractWidth;
  for (int j0 = 0; j0 < nb; j0 += 64) {
    hipLaunchKernelGGL(contract_kernel, dim3((unsigned)nblk), dim3(256), 0, s, Ns, nb, j0, chunk, bs, bsH, bsA,
                       cells, kappa, fgv, partial);
    hipLaunchKernelGGL(contract_sum_kernel, dim3((width + 255) / 256), dim3(256), 0, s, (int)nblk, partial, sums);
    // out layout: [kappa(nb) | gamma(nb) | vx(nb) | vy(nb) | wx(nb) | wy(nb) | cell column sums(40)]
    const int w = (nb - j0) < 64 ? (nb - j0) : 64;
    for (int q = 0; q < 6; ++q)
      SPDE_HIP(hipMemcpyAsync(out + q * nb + j0, sums + q * 64, sizeof(double) * w, hipMemcpyDeviceToDevice, s));
    if (j0 == 0)
      SPDE_HIP(hipMemcpyAsync(out + 6 * nb, sums + 6 * 64, sizeof(double) * kCells, hipMemcpyDeviceToDevice, s));
  }
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // device address of `scratch`, per device

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int contract_scratch(double** ptr) {
  int dev = 0;
  SPDE_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SPDE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("spde_fields_contract_noalloc: no device address for the scratch array");
      return SPDE_ERR_HIP;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *ptr = g_scratch_ptr[dev];
  return SPDE_OK;
}

}  // namespace spde

using namespace spde;

extern "C" {

int spde_fields_eval(int64_t Ns, int nb, const double* bs, const double* bsH, const double* bsA, const double* coef,
                     double* kappa, double* H, double* G, double* fgv, void* stream) {
  if (Ns <= 0 || nb <= 0 || !bs || !bsH || !bsA || !coef || !kappa || !H || !G || !fgv) {
    set_error("spde_fields_eval: invalid argument");
    return SPDE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = (unsigned)((Ns + 3) / 4);
  hipLaunchKernelGGL(fields_kernel, dim3(blocks), dim3(256), 0, s, Ns, nb, bs, bsH, bsA, coef, kappa, H, G, fgv);
  SPDE_HIP(hipGetLastError());
  return SPDE_OK;
}

int spde_fields_contract_workspace(int64_t Ns, int nb, int64_t* doubles) {
  if (Ns <= 0 || nb <= 0 || !doubles) {
    set_error("spde_fields_contract_workspace: invalid argument");
    return SPDE_ERR_ARG;
  }
  const int64_t nblk = contract_blocks(Ns);
  static_assert(kContractScratch == (240 + 1) * kContractWidth, "scratch holds the largest workspace");
  *doubles = nblk * kContractWidth + kContractWidth;
  return SPDE_OK;
}

int spde_fields_contract(int64_t Ns, int nb, const double* bs, const double* bsH, const double* bsA,
                         const double* cells, const double* kappa, const double* fgv, double* work, double* out,
                         void* stream) {
  if (Ns <= 0 || nb <= 0 || !bs || !bsH || !bsA || !cells || !kappa || !fgv || !work || !out) {
    set_error("spde_fields_contract: invalid argument");
    return SPDE_ERR_ARG;
  }
  return contract_impl(Ns, nb, bs, bsH, bsA, cells, kappa, fgv, work, out, (hipStream_t)stream);
}

int spde_fields_contract_noalloc(int64_t Ns, int nb, const double* bs, const double* bsH, const double* bsA,
                                 const double* cells, const double* kappa, const double* fgv, double* out,
                                 void* stream) {
  if (Ns <= 0 || nb <= 0 || !bs || !bsH || !bsA || !cells || !kappa || !fgv || !out) {
    set_error("spde_fields_contract_noalloc: invalid argument");
    return SPDE_ERR_ARG;
  }
  const int64_t need = contract_blocks(Ns) * kContractWidth + kContractWidth;
  if (need > kContractScratch) {
    set_error("spde_fields_contract_noalloc: the contraction needs " + std::to_string(need) +
              " doubles of scratch, more than the static maximum of " + std::to_string(kContractScratch));
    return SPDE_ERR_ARG;
  }
  double* work = nullptr;
  const int rc = contract_scratch(&work);
  if (rc != SPDE_OK) return rc;
  return contract_impl(Ns, nb, bs, bsH, bsA, cells, kappa, fgv, work, out, (hipStream_t)stream);
}

}  // extern "C"
