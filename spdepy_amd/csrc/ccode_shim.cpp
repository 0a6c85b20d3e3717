// Stencil objects with the reference's C interface, one library per stencil
// kind and boundary condition, so that the reference's own ctypes bindings
// (setClib, advection_idiffusion2D.py:246-300 and the same method in every
// model file) load them unchanged in place of lib_{AcH,AH,Acw,Aw}_2D_b{1,2,3}.so:
//
//   AH_new(numX, numY, H, hx, hy) -> object     (AcH_2D_b1.cpp:170-185, AH_2D_b1.cpp:154-169)
//   AH_Row / AH_Col / AH_Val(object) -> host arrays of numX numY 9 entries
//   AH_delete(object)
//   Aw_new(numX, numY, G, hx, hy, diff[, dG])   (Acw_2D_b1.cpp:117-132, Aw_2D_b1.cpp:136-151)
//   Aw_Row / Aw_Col / Aw_Val (numX numY 5 entries), Aw_delete
//
// The stencil itself is computed on the GPU by spde_stencil_diffusion /
// spde_stencil_advection (same slot layout and sentinel row numX numY); the
// object owns host copies of the three arrays, as the reference's objects do.
// Built once per (SHIM_DIFFUSION, SHIM_VARIABLE, SHIM_BC).  The constant
// diffusion stencil takes a pointer to the 2x2 tensor for every bc, which is
// what the reference's Python side passes (its AcH_2D_b2.cpp declares a
// scalar there, and reads an undefined value); bc = 2 then uses the periodic
// stencil with every face equal to H.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/spdepy_amd.h"

#ifndef SHIM_DIFFUSION
#error "build with -DSHIM_DIFFUSION=0|1 -DSHIM_VARIABLE=0|1 -DSHIM_BC=1|2|3"
#endif

namespace {

struct StencilObj {
  std::vector<int> row, col;
  std::vector<double> val;
};

template <class T>
T* to_device(const T* h, size_t n) {
  T* d = nullptr;
  if (!h || n == 0) return nullptr;
  if (hipMalloc((void**)&d, sizeof(T) * n) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(d);
    return nullptr;
  }
  return d;
}

// Runs `call` on device outputs of `slots` entries per cell and copies them back.
template <class F>
StencilObj* make(int M, int N, int slots, F call) {
  const size_t n = (size_t)M * N * slots;
  int32_t *row = nullptr, *col = nullptr;
  double* val = nullptr;
  StencilObj* o = nullptr;
  if (hipMalloc((void**)&row, sizeof(int32_t) * n) == hipSuccess &&
      hipMalloc((void**)&col, sizeof(int32_t) * n) == hipSuccess &&
      hipMalloc((void**)&val, sizeof(double) * n) == hipSuccess) {
    const int rc = call(row, col, val);
    if (rc == SPDE_OK && hipDeviceSynchronize() == hipSuccess) {
      o = new StencilObj();
      o->row.resize(n);
      o->col.resize(n);
      o->val.resize(n);
      if (hipMemcpy(o->row.data(), row, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(o->col.data(), col, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(o->val.data(), val, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) {
        delete o;
        o = nullptr;
      }
    } else {
      std::fprintf(stderr, "spdepy_amd stencil: %s\n", spde_last_error());
    }
  }
  hipFree(row);
  hipFree(col);
  hipFree(val);
  return o;
}

}  // namespace

extern "C" {

#if SHIM_DIFFUSION
void* AH_new(int numX, int numY, const double* H, double hx, double hy) {
  const size_t nh = SHIM_VARIABLE ? (size_t)numX * numY * 16 : 4;
  double* dH = to_device(H, nh);
  if (!dH) return nullptr;
  StencilObj* o = make(numX, numY, 9, [&](int32_t* r, int32_t* c, double* v) {
    return spde_stencil_diffusion(numX, numY, dH, SHIM_VARIABLE, hx, hy, SHIM_BC, r, c, v, nullptr);
  });
  hipFree(dH);
  return o;
}
int* AH_Row(void* o) { return o ? static_cast<StencilObj*>(o)->row.data() : nullptr; }
int* AH_Col(void* o) { return o ? static_cast<StencilObj*>(o)->col.data() : nullptr; }
double* AH_Val(void* o) { return o ? static_cast<StencilObj*>(o)->val.data() : nullptr; }
void AH_delete(void* o) { delete static_cast<StencilObj*>(o); }
#else
#if SHIM_VARIABLE
void* Aw_new(int numX, int numY, const double* G, double hx, double hy, int diff, const double* dG) {
  const size_t ng = (size_t)numX * numY * 4;
#else
void* Aw_new(int numX, int numY, const double* G, double hx, double hy, int diff) {
  const size_t ng = 2;
  const double* dG = nullptr;
#endif
  double* dGv = to_device(G, ng);
  double* ddG = to_device(dG, ng);
  if (!dGv || (dG && !ddG)) {
    hipFree(dGv);
    hipFree(ddG);
    return nullptr;
  }
  StencilObj* o = make(numX, numY, 5, [&](int32_t* r, int32_t* c, double* v) {
    return spde_stencil_advection(numX, numY, dGv, ddG, SHIM_VARIABLE, hx, hy, diff, SHIM_BC, r, c, v, nullptr);
  });
  hipFree(dGv);
  hipFree(ddG);
  return o;
}
int* Aw_Row(void* o) { return o ? static_cast<StencilObj*>(o)->row.data() : nullptr; }
int* Aw_Col(void* o) { return o ? static_cast<StencilObj*>(o)->col.data() : nullptr; }
double* Aw_Val(void* o) { return o ? static_cast<StencilObj*>(o)->val.data() : nullptr; }
void Aw_delete(void* o) { delete static_cast<StencilObj*>(o); }
#endif

}  // extern "C"
