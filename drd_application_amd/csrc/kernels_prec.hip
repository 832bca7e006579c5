// kernels_prec.hip -- preconditioners of the fast path (no reference
// counterpart; the reference's own ones are in kernels_stencil.hip).
//
// 1. CH 2x2 block Jacobi: per cell, the exact 2x2 diagonal block of the
//    ph/mu operator of MATRIX_VECTOR_PH_MU_CG (TP/ph_mu_solve.f90:203-299),
//    inverted in closed form.  The reference's scalar Jacobi ignores the
//    phi-mu coupling, which makes its BiCGSTAB stall.
//
// 2. Velocity fast-Poisson preconditioner M^-1 = S^-1 A0^-1 S^-1, with A0 the
//    constant-coefficient operator a0 I - dt (2 dxx + dyy) on u and
//    a0 I - dt (dxx + 2 dyy) on v under the same ghost rules as the
//    reference's operator (u: mirror at i=0,nx -> DCT-I in x, odd at the
//    walls -> DST-II in y; v: even at the side walls -> DCT-II in x,
//    Dirichlet at k=0,ny -> DST-I in y), and S = sqrt(diag(A)/diag(A0)).
//    The transforms and the diagonal scalings run in T = float (FP32) or
//    double; input and output vectors are FP64.
//
// Lines are transformed in LDS by a complex radix-2 Stockham FFT whose length
// is a power of two (2nx, nx, ny, 2ny); other grids are rejected at set-up.
#include <cmath>
#include <vector>

#include "fft_lines.hpp"
#include "kernels.hpp"

namespace drd {

// ---- CH block Jacobi --------------------------------------------------------
// diagonal coefficient of op_double_gradient_Neum at (i,j) (TP/matrix_tool.f90:178-219)
__device__ __forceinline__ double dgn_diag(const Grid& g, double dx2, double dy2, const double* __restrict__ mo,
                                           int i, int j) {
  const double m0 = mo[g.ci(i, j)];
  double dx = 0.0, dy = 0.0;
  if (i > 0) dx += (mo[g.ci(i - 1, j)] + m0) / 2.0;
  if (i < g.nx - 1) dx += (m0 + mo[g.ci(i + 1, j)]) / 2.0;
  if (j > 0) dy += (mo[g.ci(i, j - 1)] + m0) / 2.0;
  if (j < g.ny - 1) dy += (m0 + mo[g.ci(i, j + 1)]) / 2.0;
  return -(dx / dx2 + dy / dy2);
}

__global__ __launch_bounds__(kBlock) void k_phmu_block(Grid g, Scal s, const double* __restrict__ bu,
                                                       const double* __restrict__ minner,
                                                       const double* __restrict__ m0,
                                                       const double* __restrict__ b, double* __restrict__ x,
                                                       const KState* st) {
  if (st && (st->done || st->early)) return;
  const int64_t C = g.ncell();
  const int64_t q = (int64_t)tile_id() * kBlock + threadIdx.x;  // XCD-contiguous rows
  if (q >= C) return;
  const int j = g.by_nx.div(q), i = (int)(q - (int64_t)j * g.nx);
  const double bv = bu[g.ci(i, j)];
  const double d11 = s.s / s.I - s.I * dgn_diag(g, s.dr2, s.dz2, minner, i, j) +
                     ((j == 0) ? s.I / s.dz2 * (1.0 - s.coef_bot) : 0.0) +
                     ((j == g.ny - 1) ? s.I / s.dz2 * (1.0 - s.coef_top) : 0.0);
  const double d12 = -bv;
  const double d21 = bv + s.adj;
  // mobility-free floor keeps the block invertible where M_0 vanishes
  const double floor22 = 1e-3 * s.dt * s.xld * (2.0 / s.dr2 + 2.0 / s.dz2);
  double d22 = -s.dt * s.xld * dgn_diag(g, s.dr2, s.dz2, m0, i, j);
  d22 = d22 > floor22 ? d22 : floor22;
  const double det = d11 * d22 - d12 * d21;
  const double r1 = b[q], r2 = b[C + q];
  x[q] = (d22 * r1 - d12 * r2) / det;
  x[C + q] = (-d21 * r1 + d11 * r2) / det;
}

void launch_phmu_block(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const double* b, double* x,
                       const KState* ks) {
  hipLaunchKernelGGL(k_phmu_block, dim3((unsigned)((g.ncell() + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, g,
                     s, c.bu, c.minner, c.m0, b, x, ks);
}

// ---- exact diagonal of the BDF2 velocity operator and the scaling S ---------
// (MatrixVector_uv_2_CG TP/VeloSolve_2_combined.f90:607-670 restricted to the
// coefficient of the row's own unknown)
__global__ __launch_bounds__(kBlock) void k_velo_scale(Grid g, Scal s, const double* __restrict__ rho,
                                                       const double* __restrict__ eta,
                                                       const double* __restrict__ chi, double a0,
                                                       double* __restrict__ inv_s64, float* __restrict__ inv_s32) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (q >= g.nuv()) return;
  double d, d0;
  if (q < g.mn()) {
    const int k = g.by_nx1.div(q), i = (int)(q - (int64_t)k * (nx + 1));
    d = 1.5 * s.re * (rho[g.ci(i, k)] + rho[g.ci(i - 1, k)]) / 2.0 +
        s.penalty_k * (chi[g.ci(i, k)] + chi[g.ci(i - 1, k)]) / 2.0 * s.dt;
    d += 2.0 * s.dt * (eta[g.ci(i, k)] + eta[g.ci(i - 1, k)]) / s.dr2;
    const double cr = (eta[g.ci(i, k + 1)] + eta[g.ci(i - 1, k + 1)] + eta[g.ci(i, k)] + eta[g.ci(i - 1, k)]) / 4.0;
    const double cl = (eta[g.ci(i, k)] + eta[g.ci(i - 1, k)] + eta[g.ci(i, k - 1)] + eta[g.ci(i - 1, k - 1)]) / 4.0;
    d += s.dt * ((k == 0 ? 2.0 * cl : cl) + (k == ny - 1 ? 2.0 * cr : cr)) / s.dz2;
    d0 = a0 + s.dt * (4.0 / s.dr2 + 2.0 / s.dz2);
  } else {
    const int64_t t = q - g.mn();
    const int k = g.by_nx.div(t) + 1, i = (int)(t - (int64_t)(k - 1) * nx);
    d = 1.5 * s.re * (rho[g.ci(i, k)] + rho[g.ci(i, k - 1)]) / 2.0 +
        s.penalty_k * (chi[g.ci(i, k)] + chi[g.ci(i, k - 1)]) / 2.0 * s.dt;
    d += 2.0 * s.dt * (eta[g.ci(i, k)] + eta[g.ci(i, k - 1)]) / s.dz2;
    const double cr = (eta[g.ci(i + 1, k)] + eta[g.ci(i, k)] + eta[g.ci(i + 1, k - 1)] + eta[g.ci(i, k - 1)]) / 4.0;
    const double cl = (eta[g.ci(i, k)] + eta[g.ci(i - 1, k)] + eta[g.ci(i, k - 1)] + eta[g.ci(i - 1, k - 1)]) / 4.0;
    d += s.dt * ((i == 0 ? 0.0 : cl) + (i == nx - 1 ? 0.0 : cr)) / s.dr2;
    d0 = a0 + s.dt * (2.0 / s.dr2 + 4.0 / s.dz2);
  }
  const double is = sqrt(d0 / d);
  inv_s64[q] = is;
  inv_s32[q] = (float)is;
}

template <class T>
struct PrecTables {
  const Cx<T>* tw2nx;  // L = 2nx (u rows, DCT-I)
  const Cx<T>* twnx;   // L = nx  (v rows, DCT-II/III)
  const Cx<T>* rotnx;
  const Cx<T>* twny;   // L = ny  (u columns, DST-II/III)
  const Cx<T>* rotny;
  const Cx<T>* tw2ny;  // L = 2ny (v columns, DST-I)
  const T* lam_u;      // (nx+1) x ny, column-major by kx: lam_u[kx*ny + ky]
  const T* lam_v;      // nx x (ny-1): lam_v[kx*(ny-1) + ky]
};

// pass 1 (fwd) / 3 (inv) over rows: u rows (blocks [0,ny)), v rows ([ny, 2ny-1)).
// fwd: w = DCT_x(inv_s * r) in T;  inv: z = inv_s * DCT_x^-1(w) in FP64
template <class T, bool FWD>
__global__ __launch_bounds__(kBlock) void k_prec_rows(Grid g, PrecTables<T> tb, const double* __restrict__ r,
                                                      const T* __restrict__ inv_s, T* __restrict__ wu,
                                                      T* __restrict__ wv, double* __restrict__ z,
                                                      const KState* st) {
  if (st && (st->done || st->early)) return;
  extern __shared__ unsigned char smem_raw[];
  const int nx = g.nx, ny = g.ny;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(smem_raw);
  Cx<T>* B = A + 2 * nx;
  T* line = reinterpret_cast<T*>(B + 2 * nx);
  const bool is_u = blockIdx.x < (unsigned)ny;
  const int n = is_u ? nx + 1 : nx;
  const int64_t base = is_u ? (int64_t)blockIdx.x * (nx + 1) : g.mn() + (int64_t)(blockIdx.x - ny) * nx;
  T* w = is_u ? wu + (int64_t)blockIdx.x * (nx + 1) : wv + (int64_t)(blockIdx.x - ny) * nx;
  for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = FWD ? T(r[base + i]) * inv_s[base + i] : w[i];
  __syncthreads();
  if (is_u)
    line_transform<T>(TR_DCT1, line, n, A, B, tb.tw2nx, nullptr);
  else
    line_transform<T>(FWD ? TR_DCT2 : TR_DCT3, line, n, A, B, tb.twnx, tb.rotnx);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (FWD)
      w[i] = line[i];
    else
      z[base + i] = (double)(line[i] * inv_s[base + i]);
  }
}

// pass 2 over columns of the transposed blocks: u columns (blocks [0,nx+1)),
// v columns ([nx+1, 2nx+1)); forward transform, divide by the eigenvalue, inverse.
template <class T>
__global__ __launch_bounds__(kBlock) void k_prec_cols(Grid g, PrecTables<T> tb, T* __restrict__ tu,
                                                      T* __restrict__ tv, const KState* st) {
  if (st && (st->done || st->early)) return;
  extern __shared__ unsigned char smem_raw[];
  const int nx = g.nx, ny = g.ny;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(smem_raw);
  Cx<T>* B = A + 2 * ny;
  T* line = reinterpret_cast<T*>(B + 2 * ny);
  const bool is_u = blockIdx.x < (unsigned)(nx + 1);
  const int kx = is_u ? blockIdx.x : blockIdx.x - (nx + 1);
  const int n = is_u ? ny : ny - 1;
  T* col = is_u ? tu + (int64_t)kx * ny : tv + (int64_t)kx * (ny - 1);
  const T* lam = is_u ? tb.lam_u + (int64_t)kx * ny : tb.lam_v + (int64_t)kx * (ny - 1);
  for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = col[i];
  __syncthreads();
  if (is_u) {
    line_transform<T>(TR_DST2, line, n, A, B, tb.twny, tb.rotny);
    for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = line[i] / lam[i];
    __syncthreads();
    line_transform<T>(TR_DST3, line, n, A, B, tb.twny, tb.rotny);
  } else {
    line_transform<T>(TR_DST1, line, n, A, B, tb.tw2ny, nullptr);
    for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = line[i] / lam[i];
    __syncthreads();
    line_transform<T>(TR_DST1, line, n, A, B, tb.tw2ny, nullptr);
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] = line[i];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_transposeT(int rows, int cols, const T* __restrict__ in,
                                                       T* __restrict__ out, const KState* st) {
  if (st && (st->done || st->early)) return;
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int rr = by + j, c = bx + tx;
    if (rr < rows && c < cols) tile[j][tx] = in[(size_t)rr * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int rr = bx + j, c = by + tx;
    if (rr < cols && c < rows) out[(size_t)rr * rows + c] = tile[tx][j];
  }
}

// ---- host side -------------------------------------------------------------------
struct VeloPrec {
  int nx, ny;
  bool fp32;
  void* tabs_mem = nullptr;  // all tables in one allocation
  PrecTables<float> tf{};
  PrecTables<double> td{};
  double* inv_s64 = nullptr;
  float* inv_s32 = nullptr;
  void* work = nullptr;  // wu, wv, tu, tv in T
  double a0 = 0;
  bool scaled = false;
};

template <class T>
static void host_tw(std::vector<Cx<T>>& v, int L) {
  for (int t = 0; t < L / 2; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)L;
    v.push_back(Cx<T>{(T)std::cos(a), (T)std::sin(a)});
  }
}
template <class T>
static void host_rot(std::vector<Cx<T>>& v, int L) {
  for (int k = 0; k < L; ++k) {
    const double a = -M_PI * (double)k / (2.0 * L);
    v.push_back(Cx<T>{(T)std::cos(a), (T)std::sin(a)});
  }
}

// eigenvalues of A0 times the unnormalised transform pair scale (2nx)(2ny)
template <class T>
static int build_prec_tables(VeloPrec* pc, const Scal& s, PrecTables<T>& tb) {
  const int nx = pc->nx, ny = pc->ny;
  std::vector<Cx<T>> cx;
  std::vector<size_t> off;
  off.push_back(cx.size());
  host_tw<T>(cx, 2 * nx);
  off.push_back(cx.size());
  host_tw<T>(cx, nx);
  off.push_back(cx.size());
  host_rot<T>(cx, nx);
  off.push_back(cx.size());
  host_tw<T>(cx, ny);
  off.push_back(cx.size());
  host_rot<T>(cx, ny);
  off.push_back(cx.size());
  host_tw<T>(cx, 2 * ny);
  off.push_back(cx.size());
  std::vector<T> lam;
  const double sc = (2.0 * nx) * (2.0 * ny);
  for (int kx = 0; kx <= nx; ++kx)
    for (int ky = 0; ky < ny; ++ky) {
      const double sx = std::sin(M_PI * kx / (2.0 * nx)), sy = std::sin(M_PI * (ky + 1) / (2.0 * ny));
      lam.push_back((T)(sc * (pc->a0 + s.dt * (2.0 * 4.0 * sx * sx / s.dr2 + 4.0 * sy * sy / s.dz2))));
    }
  const size_t lam_v_off = lam.size();
  for (int kx = 0; kx < nx; ++kx)
    for (int ky = 1; ky < ny; ++ky) {
      const double sx = std::sin(M_PI * kx / (2.0 * nx)), sy = std::sin(M_PI * ky / (2.0 * ny));
      lam.push_back((T)(sc * (pc->a0 + s.dt * (4.0 * sx * sx / s.dr2 + 2.0 * 4.0 * sy * sy / s.dz2))));
    }
  const size_t bytes_c = cx.size() * sizeof(Cx<T>), bytes_l = lam.size() * sizeof(T);
  if (hipMalloc(&pc->tabs_mem, bytes_c + bytes_l) != hipSuccess) return DRD_E_HIP;
  if (hipMemcpy(pc->tabs_mem, cx.data(), bytes_c, hipMemcpyHostToDevice) != hipSuccess) return DRD_E_HIP;
  if (hipMemcpy((char*)pc->tabs_mem + bytes_c, lam.data(), bytes_l, hipMemcpyHostToDevice) != hipSuccess)
    return DRD_E_HIP;
  const Cx<T>* base = reinterpret_cast<const Cx<T>*>(pc->tabs_mem);
  tb.tw2nx = base + off[0];
  tb.twnx = base + off[1];
  tb.rotnx = base + off[2];
  tb.twny = base + off[3];
  tb.rotny = base + off[4];
  tb.tw2ny = base + off[5];
  const T* lb = reinterpret_cast<const T*>((char*)pc->tabs_mem + bytes_c);
  tb.lam_u = lb;
  tb.lam_v = lb + lam_v_off;
  return 0;
}

static bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

VeloPrec* velo_prec_create(const Grid& g, const Scal& s, bool fp32, int* err) {
  *err = 0;
  if (!pow2(g.nx) || !pow2(g.ny) || g.nx > 4096 || g.ny > 4096) {
    *err = DRD_E_UNSUPPORTED;
    return nullptr;
  }
  VeloPrec* pc = new VeloPrec();
  pc->nx = g.nx;
  pc->ny = g.ny;
  pc->fp32 = fp32;
  pc->a0 = 1.5 * s.re;  // unit density, unit viscosity, no penalty
  int rc = fp32 ? build_prec_tables<float>(pc, s, pc->tf) : build_prec_tables<double>(pc, s, pc->td);
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  if (!rc && hipMalloc(&pc->inv_s64, g.nuv() * sizeof(double)) != hipSuccess) rc = DRD_E_HIP;
  if (!rc && hipMalloc(&pc->inv_s32, g.nuv() * sizeof(float)) != hipSuccess) rc = DRD_E_HIP;
  if (!rc && hipMalloc(&pc->work, 2 * g.nuv() * esz) != hipSuccess) rc = DRD_E_HIP;
  if (rc) {
    velo_prec_destroy(pc);
    *err = rc;
    return nullptr;
  }
  static bool raised = false;
  if (!raised) {  // rows of 2nx complex (+ the line) may exceed 64 KiB
    const int lim = 160 * 1024;
    (void)hipFuncSetAttribute((const void*)k_prec_rows<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_prec_rows<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_prec_rows<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_prec_rows<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_prec_cols<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_prec_cols<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    raised = true;
  }
  return pc;
}

void velo_prec_destroy(VeloPrec* pc) {
  if (!pc) return;
  (void)hipFree(pc->tabs_mem);
  (void)hipFree(pc->inv_s64);
  (void)hipFree(pc->inv_s32);
  (void)hipFree(pc->work);
  delete pc;
}

void velo_prec_setup(hipStream_t st, VeloPrec* pc, const Grid& g, const Scal& s, const Coef& c) {
  hipLaunchKernelGGL(k_velo_scale, dim3((unsigned)((g.nuv() + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, g, s,
                     c.rho, c.eta, c.chi, pc->a0, pc->inv_s64, pc->inv_s32);
}

template <class T>
static void apply_T(hipStream_t st, VeloPrec* pc, const Grid& g, const PrecTables<T>& tb, const T* inv_s,
                    const double* r, double* z, const KState* ks) {
  const int nx = g.nx, ny = g.ny;
  T* wu = reinterpret_cast<T*>(pc->work);
  T* wv = wu + (size_t)(nx + 1) * ny;
  T* tu = wv + (size_t)nx * (ny - 1);
  T* tv = tu + (size_t)(nx + 1) * ny;
  const size_t lds_rows = 4 * (size_t)nx * sizeof(Cx<T>) + (size_t)(nx + 1) * sizeof(T);
  const size_t lds_cols = 4 * (size_t)ny * sizeof(Cx<T>) + (size_t)(ny + 1) * sizeof(T);
  hipLaunchKernelGGL((k_prec_rows<T, true>), dim3(2 * ny - 1), dim3(kBlock), lds_rows, st, g, tb, r, inv_s, wu, wv,
                     (double*)nullptr, ks);
  hipLaunchKernelGGL(k_transposeT<T>, dim3((nx + 1 + 31) / 32, (ny + 31) / 32), dim3(kBlock), 0, st, ny, nx + 1,
                     (const T*)wu, tu, ks);
  hipLaunchKernelGGL(k_transposeT<T>, dim3((nx + 31) / 32, (ny - 1 + 31) / 32), dim3(kBlock), 0, st, ny - 1, nx,
                     (const T*)wv, tv, ks);
  hipLaunchKernelGGL(k_prec_cols<T>, dim3(2 * nx + 1), dim3(kBlock), lds_cols, st, g, tb, tu, tv, ks);
  hipLaunchKernelGGL(k_transposeT<T>, dim3((ny + 31) / 32, (nx + 1 + 31) / 32), dim3(kBlock), 0, st, nx + 1, ny,
                     (const T*)tu, wu, ks);
  hipLaunchKernelGGL(k_transposeT<T>, dim3((ny - 1 + 31) / 32, (nx + 31) / 32), dim3(kBlock), 0, st, nx, ny - 1,
                     (const T*)tv, wv, ks);
  hipLaunchKernelGGL((k_prec_rows<T, false>), dim3(2 * ny - 1), dim3(kBlock), lds_rows, st, g, tb,
                     (const double*)nullptr, inv_s, wu, wv, z, ks);
}

void velo_prec_apply(hipStream_t st, VeloPrec* pc, const Grid& g, const double* r, double* z, const KState* ks) {
  if (pc->fp32)
    apply_T<float>(st, pc, g, pc->tf, pc->inv_s32, r, z, ks);
  else
    apply_T<double>(st, pc, g, pc->td, pc->inv_s64, r, z, ks);
}

}  // namespace drd
