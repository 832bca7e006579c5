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

// Exact floor(n / d) for 0 <= n < 2^31 by multiply and shift (Granlund and
// Montgomery, PLDI 1994): l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (mulhi(m, n) + n) >> l.  Replaces the software integer division
// in the index decode of every stencil thread.
struct FastDiv {
  uint32_t m;
  int l;
  __host__ static FastDiv make(uint32_t d) {
    int s = 0;
    while ((1ull << s) < d) ++s;
    return FastDiv{(uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1), s};
  }
  __device__ __forceinline__ int div(int64_t n) const {
    const uint32_t u = (uint32_t)n;
    return (int)((__umulhi(m, u) + u) >> l);
  }
};

// px: periodic in x (lr_bd_type 'p'): a u row has nx unknowns (u(nx) =
// u(0)) instead of nx + 1, and the ghost columns of every field hold the
// periodic copies f(-1) = f(nx-1), f(nx) = f(0) (u: also f(nx+1) = f(1)).
struct Grid {
  int nx, ny;
  FastDiv by_nxu, by_nx;  // division by nxu() (u unknowns per row) and by nx
  int px;
  __host__ static Grid make(int nx, int ny, int px = 0) {
    return Grid{nx, ny, FastDiv::make((uint32_t)(nx + 1 - px)), FastDiv::make((uint32_t)nx), px};
  }
  __host__ __device__ int nxu() const { return nx + 1 - px; }
  __host__ __device__ int ldc() const { return nx + 2; }
  __host__ __device__ int ldu() const { return nx + 3; }
  __host__ __device__ int64_t nc() const { return (int64_t)(nx + 2) * (ny + 2); }
  __host__ __device__ int64_t nuf() const { return (int64_t)(nx + 3) * (ny + 2); }
  __host__ __device__ int64_t nvf() const { return (int64_t)(nx + 2) * (ny + 1); }
  __host__ __device__ int64_t mn() const { return (int64_t)nxu() * ny; }
  // the interior column of a ghost column: wrapped (px) or clamped
  __host__ __device__ int xcol(int i) const {
    return px ? (i < 0 ? i + nx : (i >= nx ? i - nx : i)) : (i < 0 ? 0 : (i >= nx ? nx - 1 : i));
  }
  __host__ __device__ int64_t nuv() const { return mn() + (int64_t)nx * (ny - 1); }
  __host__ __device__ int64_t ncell() const { return (int64_t)nx * ny; }
  __host__ __device__ int64_t ci(int i, int k) const { return (int64_t)(i + 1) + (int64_t)(nx + 2) * (k + 1); }
  __host__ __device__ int64_t ui(int i, int k) const { return (int64_t)(i + 1) + (int64_t)(nx + 3) * (k + 1); }
  __host__ __device__ int64_t vi(int i, int k) const { return (int64_t)(i + 1) + (int64_t)(nx + 2) * k; }
  // 32-bit forms for the hot stencils (every array is < 2^31 elements)
  __device__ __forceinline__ uint32_t cio(int i, int k) const { return (uint32_t)((i + 1) + (nx + 2) * (k + 1)); }
};

// Load / store at a 32-bit element offset: the byte offset stays 32 bits, so
// the access uses the scalar-base + 32-bit vector-offset address form
// instead of 64-bit address arithmetic per neighbour.
template <class T>
__device__ __forceinline__ T ld32(const T* __restrict__ p, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (uint32_t)sizeof(T));
}
// a load of a stored FP32 or FP64 value, widened to double for the arithmetic
template <class T>
__device__ __forceinline__ double ldd(const T* __restrict__ p, uint32_t i) {
  return (double)ld32(p, i);
}
template <class T>
__device__ __forceinline__ void st32(T* __restrict__ p, uint32_t i, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (uint32_t)sizeof(T)) = v;
}

// ---- batched launches ---------------------------------------------------------
// A batch group is several instances of one grid size whose per-instance
// buffers (fields, Krylov vectors, state, partial sums, scratch) share one
// layout: instance i's copy of any of them sits i * bs bytes after instance
// 0's.  Kernels on the time-step path take the instance from blockIdx.z and
// offset their per-instance pointers with bofs (tables shared by the group
// are not offset).  A single instance is the case n = 1, bs = 0.
struct BatchCfg {
  int n = 1;
  int64_t bs = 0;
};
BatchCfg& batch_cfg();  // the calling host thread's current group (drd_api.hip)
inline unsigned bat_n() { return (unsigned)batch_cfg().n; }
inline int64_t bat_bs() { return batch_cfg().bs; }

// The offset is pointer arithmetic on the argument itself (not a round trip
// through an integer), so the compiler still knows the result points to
// global memory and emits global loads and stores rather than flat ones.
template <class T>
__device__ __forceinline__ T* bofs(T* p, int64_t bs) {
  return p ? (T*)((const char*)p + (int64_t)blockIdx.z * bs) : p;
}

// Scalars of module `global` the kernels read (TP/global.f90:42-52).
struct Scal {
  double dt, re, ca, xld, s, I, dr, dz, dr2, dz2;
  double penalty_k, ggx, ggz, uwT, uwB, adj;
  double coef_bot, coef_top;
  int lr_d;  // lr_bd_type == 'd', or the MF variant (whose inflow is always on)
  int phi_mode;
  int mf;    // the MF variant's momentum right-hand side (MF/VeloSolve_2_combined.f90)
  // quotients of the above that the CH stencils use (scal_derive)
  double idr2, idz2, s_I, I_dz2, floor22;
  double idrdz;  // 1 / (dr dz): the FP32 velocity stencil multiplies by it
};

// The stencils' uniform quotients, each the correctly rounded value of the
// expression the stencil would evaluate (1.0 / dr2, s / I, I / dz2, the
// block floor), so that using them changes no result.
__host__ __device__ inline void scal_derive(Scal& s) {
  s.idr2 = 1.0 / s.dr2;
  s.idz2 = 1.0 / s.dz2;
  s.idrdz = 1.0 / (s.dr * s.dz);
  s.s_I = s.s / s.I;
  s.I_dz2 = s.I / s.dz2;
  s.floor22 = 1e-3 * s.dt * s.xld * (2.0 / s.dr2 + 2.0 / s.dz2);
}

// Device-side Krylov state: the reverse-communication scalars of
// LIB/BiCGSTABREVCOM.f:114-119 plus flags the kernels test so that the
// whole iteration runs without host round trips.
struct KState {
  double rho, rho1, alpha, omega, bnrm2, tol, resid, snrm;
  double breaktol;
  double red[8];      // finished reductions of the last phase
  int32_t iter, maxit, info, done, early, xnz;
  uint32_t eticket;  // workgroups of the early-exit update that have finished
  int32_t pad;
};

// ---- canonical reductions ---------------------------------------------------
// Every sum on the path (dot products, norms, the divergence mean) is the
// pairwise sum over the element index, padded with zeros to a power of two:
//   S[lo, lo+2L) = S[lo, lo+L) + S[lo+L, lo+2L).
// The value depends on the data only, not on how many threads or
// workgroups compute it, and a serial restatement reproduces it bit for
// bit.  (The padding width can change only the sign of an all-zero sum.)
// A Krylov vector is 1 segment, or 2 equal segments (the phi|mu blocks of
// the Cahn-Hilliard system); its dot product is the sum of the segment
// sums in segment order.
//
// A workgroup of a fused sweep covers kTile consecutive elements of each
// segment; thread t holds element j*kBlock + t of the tile in slot j, so
// wave w holds the aligned 64-element chunk j*kWaves + w in every slot.
constexpr int kTile = 1024;                 // elements per workgroup and segment
constexpr int kSlots = kTile / kBlock;      // 4 elements per thread
constexpr int kChunks = kSlots * kWaves;    // 16 chunks of 64 per tile
constexpr int64_t kMaxTiles = (int64_t)kBlock * kBlock;  // per segment

// v as seen from another lane of the same row of 16 (DPP lane permutation
// CTRL; the whole wave is active wherever this is used).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Pairwise sum of the wave's 64 values (lane l holds element l), on lane 0.
// Levels 1-4 add aligned groups of 2, 4, 8 and 16 lanes: every lane adds the
// subtotal of a lane in the sibling group, taken by a DPP permutation (quad
// swaps 1 and 2, half-row and row mirrors), so every lane of a group ends
// with the group's sum (x + y and y + x are the same IEEE sum) and lane 0
// forms S[0, 2m) = S[0, m) + S[m, 2m) at each level.  The row sums R0..R3 of
// lanes 0, 16, 32 and 48 are then added as (R0 + R1) + (R2 + R3).  The
// whole wave must be active; only lane 0's result is meaningful.
__device__ __forceinline__ double wave_pairwise(double v) {
  v = v + dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v = v + dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v = v + dpp_f64<0x141>(v);  // row_half_mirror
  v = v + dpp_f64<0x140>(v);  // row_mirror
  const double r1 = readlane_f64(v, 16), r2 = readlane_f64(v, 32), r3 = readlane_f64(v, 48);
  const double lo = v + r1, hi = r2 + r3;
  return lo + hi;
}

template <int N>
__device__ __forceinline__ double pairwise_fixed(const double* a) {
  if constexpr (N == 1)
    return a[0];
  else
    return pairwise_fixed<N / 2>(a) + pairwise_fixed<N / 2>(a + N / 2);
}

// Pairwise sum of kBlock consecutive values (thread t holds value t);
// the result is on thread 0.  Every thread of the workgroup must call it.
__device__ __forceinline__ double block_pairwise(double v) {
  __shared__ double lw[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_pairwise(v);
  if (lane == 0) lw[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = pairwise_fixed<kWaves>(lw);
  __syncthreads();
  return r;
}

// Pairwise sums of K tiles at once; t[k][j] is slot j of value k.  The
// results are on thread 0.
template <int K>
__device__ __forceinline__ void tile_pairwise(const double (&t)[K][kSlots], double (&out)[K]) {
  __shared__ double lc[K][kChunks];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const double c = wave_pairwise(t[k][j]);
      if (lane == 0) lc[k][j * kWaves + w] = c;
    }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = pairwise_fixed<kChunks>(lc[k]);
  }
  __syncthreads();
}

// Pairwise sum of 16 values in registers (zeros pad shorter runs: a
// subtree sum plus zeros keeps its value).
__device__ __forceinline__ double pairwise16(double (&r)[16]) {
#pragma unroll
  for (int w = 16; w > 1; w >>= 1)
#pragma unroll
    for (int i = 0; i < w / 2; ++i) r[i] = r[2 * i] + r[2 * i + 1];
  return r[0];
}

// Pairwise sum over tiles 0..nt-1 of part[tile*stride + k] (nt <= kMaxTiles).
// With ntp = nt rounded up to a power of two, thread t sums the aligned run
// [t P, (t+1) P), P = max(1, ntp / kBlock), in registers (runs of 16, then
// the run sums), and the workgroup adds the kBlock run sums pairwise.
// Result on thread 0.  The partials come from an earlier kernel, so plain
// loads see them.
__device__ __forceinline__ double tiles_pairwise(const double* part, uint32_t nt, int stride, int k) {
  uint32_t ntp = 1;
  while (ntp < nt) ntp <<= 1;
  const uint32_t P = ntp > (uint32_t)kBlock ? ntp / kBlock : 1u;  // 1 .. 256
  const uint32_t base = threadIdx.x * P;
  const uint32_t runs = P > 16u ? P / 16u : 1u;                     // 1 .. 16
  const uint32_t len = P > 16u ? 16u : P;
  double g[16];
#pragma unroll
  for (int gi = 0; gi < 16; ++gi) {
    g[gi] = 0.0;
    if ((uint32_t)gi < runs) {
      double r[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const uint32_t q = base + (uint32_t)gi * 16u + (uint32_t)i;
        r[i] = ((uint32_t)i < len && q < nt) ? part[(size_t)q * stride + k] : 0.0;
      }
      g[gi] = pairwise16(r);
    }
  }
  return block_pairwise(pairwise16(g));
}

// Tile of this workgroup.  The dispatcher deals workgroup b to XCD b % 8;
// XCD x is given the contiguous tiles [start_x, start_x + count_x), in
// order, so the rows a stencil tile shares with its neighbours are already
// in that XCD's L2 instead of being fetched once per XCD.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t tile_id() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t per = nb / kXcds, rem = nb % kXcds;
  const uint32_t x = b % kXcds, j = b / kXcds;
  return x * per + (x < rem ? x : rem) + j;
}

// Index of slot j of segment s in this workgroup's tile, or -1 past the end
// of the segment (segments of length seg, laid out one after the other).
__device__ __forceinline__ int64_t tile_elem(int64_t seg, int s, int j) {
  const int64_t o = (int64_t)tile_id() * kTile + j * kBlock + threadIdx.x;
  return o < seg ? (int64_t)s * seg + o : -1;
}

// The same for an explicit tile index (workgroups that sweep several tiles).
template <int NS, int NV>
__device__ __forceinline__ void tile_publish_at(const double (&t)[NS * NV][kSlots], double* part, uint32_t tile) {
  constexpr int K = NS * NV;
  double p[K];
  tile_pairwise<K>(t, p);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) part[(size_t)tile * K + k] = p[k];
  }
}

// Publishes this workgroup's partial sums: value v of segment s at
// part[tile*(NS*NV) + s*NV + v].  A one-workgroup kernel launched after the
// producer (k_finalize) adds them up; the kernel boundary makes them
// visible across the XCDs' L2 caches without a per-workgroup fence.
template <int NS, int NV>
__device__ __forceinline__ void tile_publish(const double (&t)[NS * NV][kSlots], double* part) {
  constexpr int K = NS * NV;
  double p[K];
  tile_pairwise<K>(t, p);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) part[(size_t)tile_id() * K + k] = p[k];
  }
}

}  // namespace drd
