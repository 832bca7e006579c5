// drd_common.hpp -- shared device/host definitions of the gfx950 solve path.
//
// Fields live in HBM in the reference's Fortran layout (x fastest, ghost
// layers kept), so host arrays cross the C-ABI with plain copies:
//   cell field (-1:nx,-1:ny)  ld = nx+2
//   u field    (-1:nx+1,-1:ny) ld = nx+3
//   v field    (-1:nx,0:ny)    ld = nx+2
// Krylov vectors use the reference's packed orders (TP/matrix_tool.f90).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/drd_gpu.h"

namespace drd {

constexpr int kBlock = 256;  // 4 wave64 per workgroup
constexpr int kWaves = kBlock / 64;

struct Grid {
  int nx, ny;
  __host__ __device__ int ldc() const { return nx + 2; }
  __host__ __device__ int ldu() const { return nx + 3; }
  __host__ __device__ int64_t nc() const { return (int64_t)(nx + 2) * (ny + 2); }
  __host__ __device__ int64_t nuf() const { return (int64_t)(nx + 3) * (ny + 2); }
  __host__ __device__ int64_t nvf() const { return (int64_t)(nx + 2) * (ny + 1); }
  __host__ __device__ int64_t mn() const { return (int64_t)(nx + 1) * ny; }
  __host__ __device__ int64_t nuv() const { return mn() + (int64_t)nx * (ny - 1); }
  __host__ __device__ int64_t ncell() const { return (int64_t)nx * ny; }
  __host__ __device__ int64_t ci(int i, int k) const { return (int64_t)(i + 1) + (int64_t)(nx + 2) * (k + 1); }
  __host__ __device__ int64_t ui(int i, int k) const { return (int64_t)(i + 1) + (int64_t)(nx + 3) * (k + 1); }
  __host__ __device__ int64_t vi(int i, int k) const { return (int64_t)(i + 1) + (int64_t)(nx + 2) * k; }
};

// Scalars of module `global` the kernels read (TP/global.f90:42-52).
struct Scal {
  double dt, re, ca, xld, s, I, dr, dz, dr2, dz2;
  double penalty_k, ggx, ggz, uwT, uwB, adj;
  double coef_bot, coef_top;
  int lr_d;  // lr_bd_type == 'd'
  int phi_mode;
};

// Device-side Krylov state: the reverse-communication scalars of
// LIB/BiCGSTABREVCOM.f:114-119 plus flags the kernels test so that the
// whole iteration runs without host round trips.
struct KState {
  double rho, rho1, alpha, omega, bnrm2, tol, resid, snrm;
  double breaktol;
  double red[8];      // finished reductions of the last phase
  int32_t iter, maxit, info, done, early, xnz;
  uint32_t ticket;    // last-block counter of the running reduction
  int32_t pad;
};

// ---- wave64 / workgroup reductions (fixed order -> deterministic) -------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sums NV values over the workgroup; thread 0 gets the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double lds[NV][kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double t = wave_sum(v[j]);
    if (lane == 0) lds[j][w] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double t = 0.0;
#pragma unroll
      for (int q = 0; q < kWaves; ++q) t += lds[j][q];
      v[j] = t;
    }
  }
  __syncthreads();
}

// Writes this block's NV partials, and returns true in the one block that
// finishes last; that block then holds the grand totals (in fixed block
// order) in `tot` on thread 0.
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double* part, uint32_t* ticket,
                                         double (&tot)[NV]) {
  block_sum<NV>(v);
  const uint32_t nb = gridDim.x * gridDim.y;
  const uint32_t b = blockIdx.x + gridDim.x * blockIdx.y;
  __shared__ int is_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) part[(size_t)b * NV + j] = v[j];
    __threadfence();
    uint32_t t = atomicAdd(ticket, 1u);
    is_last = (t == nb - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0.0;
  for (uint32_t q = threadIdx.x; q < nb; q += blockDim.x) {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      acc[j] += __hip_atomic_load(part + (size_t)q * NV + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<NV>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) tot[j] = acc[j];
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

}  // namespace drd
