// kernels_stencil.hip -- MAC-grid and Cahn-Hilliard stencil kernels (FP64).
//
// Every kernel keeps the reference's floating-point evaluation order and is
// compiled with -ffp-contract=off, so its results equal the CPU oracle (and
// the reference Fortran) bit for bit.  One thread owns one unknown of the
// packed Krylov vector; a workgroup covers kTile consecutive unknowns, so
// neighbouring rows of a tile are served from L2 and each HBM line of x is
// fetched once per sweep.
#include <hip/hip_ext.h>

#include <algorithm>

#include "kernels.hpp"
#include "velo_rows.hpp"

namespace drd {

__device__ __forceinline__ bool kactive(const KState* st) {
  return st == nullptr || (!st->done && !st->early);
}

// Generic matvec with the reductions BiCGSTAB fuses into it.
//   MV_PLAIN : y = alpha*A*x + beta*y
//   MV_DOT1  : y = A x, partial (a0, y)                 -> PH_MV1
//   MV_DOT2  : y = A x, partials (y, a0), (y, y)        -> PH_MV2
//   MV_INIT  : y = b - A x (if ||x||!=0, else b); a0 := y; partial (y,y) -> PH_INIT_R
template <int BDF, int MODE, bool PX>
__global__ __launch_bounds__(kBlock) void k_velo_matvec(Grid g, Scal s, const double* __restrict__ rho,
                                                        const double* __restrict__ eta,
                                                        const double* __restrict__ chi,
                                                        const double* __restrict__ x, double* __restrict__ y,
                                                        double alpha, double beta, double* __restrict__ a0,
                                                        const double* __restrict__ b, double* part,
                                                        KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (MODE != MV_PLAIN && !kactive(st)) return;
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  chi = bofs(chi, bs);
  x = bofs(x, bs);
  y = bofs(y, bs);
  a0 = bofs(a0, bs);
  b = bofs(b, bs);
  part = bofs(part, bs);
  VeloInT<double, PX> X{{x}, g.nx, g.ny, s.lr_d, (uint32_t)g.mn()};
  const uint32_t n = (uint32_t)g.nuv();
  const bool xnz = (MODE == MV_INIT) ? (st->xnz != 0) : true;
  // A workgroup sweeps tile `pair` (u half of the vector) and tile
  // pair + half (v half): both cover about the same grid rows, so the u-v
  // cross terms and the shared viscosity rows are read once.
  const uint32_t nt = (uint32_t)((n + kTile - 1) / kTile), half = (nt + 1) / 2;
  const uint32_t pair = tile_id();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t tile = pair + (uint32_t)h * half;
    if (tile >= nt) break;
    double tm[2][kSlots] = {};
    // A x for the tile's four slots first, then the stores: a tile lies in
    // the u half or the v half of the vector except the one tile across the
    // boundary, so in the others the four rows are straight-line code whose
    // loads the compiler can issue together.
    const uint32_t t0 = tile * (uint32_t)kTile;
    double z[kSlots];
    const bool need = MODE != MV_INIT || xnz;
    if (need && t0 + (uint32_t)kTile <= X.mn) {
#pragma unroll
      for (int j = 0; j < kSlots; ++j)
        z[j] = velo_u_row<BDF>(g, s, rho, eta, chi, X, t0 + (uint32_t)(j * kBlock) + threadIdx.x);
    } else if (need && t0 >= X.mn) {
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const uint32_t q = t0 + (uint32_t)(j * kBlock) + threadIdx.x;
        z[j] = velo_v_row<BDF>(g, s, rho, eta, chi, X, q < n ? q : n - 1);
      }
    } else if (need) {
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const uint32_t q = t0 + (uint32_t)(j * kBlock) + threadIdx.x;
        z[j] = velo_row<BDF>(g, s, rho, eta, chi, X, q < n ? q : n - 1);
      }
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const uint32_t q = t0 + (uint32_t)(j * kBlock) + threadIdx.x;
      if (q >= n) break;
      if (MODE == MV_PLAIN) {
        // y is not read when beta == 0 (as in BLAS), so it may be uninitialised
        st32(y, q, beta == 0.0 ? alpha * z[j] : alpha * z[j] + beta * ld32((const double*)y, q));
      } else if (MODE == MV_INIT) {
        double r = ld32(b, q);
        if (xnz) r = -1.0 * z[j] + 1.0 * r;
        st32(y, q, r);
        st32(a0, q, r);
        tm[0][j] = r * r;
      } else {
        st32(y, q, z[j]);
        if (MODE == MV_DOT1) {
          tm[0][j] = ld32((const double*)a0, q) * z[j];
        } else {
          tm[0][j] = z[j] * ld32((const double*)a0, q);
          tm[1][j] = z[j] * z[j];
        }
      }
    }
    if (MODE == MV_DOT2) {
      tile_publish_at<1, 2>(tm, part, tile);
    } else if (MODE != MV_PLAIN) {
      double t1[1][kSlots];
#pragma unroll
      for (int j = 0; j < kSlots; ++j) t1[0][j] = tm[0][j];
      tile_publish_at<1, 1>(t1, part, tile);
    }
  }
}

// PSolve_uv_CG (Jacobi) TP/VeloSolve_combined.f90:692-741: the diagonal,
// with the reference's rho(i-1,j) average on the v rows (:720).
__global__ __launch_bounds__(kBlock) void k_velo_jacobi_diag(Grid g, Scal s, const double* __restrict__ rho,
                                                             const double* __restrict__ eta,
                                                             double* __restrict__ diag, double mcoef, int64_t bs) {
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  diag = bofs(diag, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (q >= g.nuv()) return;
  double tmp;
  if (q < g.mn()) {
    const int j = g.by_nxu.div(q);
    const int i = (int)(q - (int64_t)j * g.nxu());
    tmp = mcoef * (rho[g.ci(i, j)] + rho[g.ci(i - 1, j)]) / 2.0;
    double cr = eta[g.ci(i, j)], cl = eta[g.ci(i - 1, j)];
    tmp = tmp + 2.0 * s.dt * (cl + cr) / s.dr2;
    cr = (eta[g.ci(i, j + 1)] + eta[g.ci(i - 1, j + 1)] + eta[g.ci(i, j)] + eta[g.ci(i - 1, j)]) / 4.0;
    cl = (eta[g.ci(i, j)] + eta[g.ci(i - 1, j)] + eta[g.ci(i, j - 1)] + eta[g.ci(i - 1, j - 1)]) / 4.0;
    if (j == 0)
      tmp = tmp + s.dt * (cl + (cl + cr)) / s.dz2;
    else if (j == ny - 1)
      tmp = tmp + s.dt * (cr + (cl + cr)) / s.dz2;
    else
      tmp = tmp + s.dt * (cl + cr) / s.dz2;
  } else {
    const int64_t t = q - g.mn();
    const int j = g.by_nx.div(t) + 1;
    const int i = (int)(t - (int64_t)(j - 1) * nx);
    tmp = mcoef * (rho[g.ci(i, j)] + rho[g.ci(i - 1, j)]) / 2.0;
    double cr = eta[g.ci(i, j)], cl = eta[g.ci(i, j - 1)];
    tmp = tmp + 2.0 * s.dt * (cl + cr) / s.dz2;
    cr = (eta[g.ci(i + 1, j)] + eta[g.ci(i, j)] + eta[g.ci(i + 1, j - 1)] + eta[g.ci(i, j - 1)]) / 4.0;
    cl = (eta[g.ci(i, j)] + eta[g.ci(i - 1, j)] + eta[g.ci(i, j - 1)] + eta[g.ci(i - 1, j - 1)]) / 4.0;
    if (!g.px && i == 0)
      tmp = tmp + s.dt * cr / s.dr2;
    else if (!g.px && i == nx - 1)
      tmp = tmp + s.dt * cl / s.dr2;
    else
      tmp = tmp + s.dt * (cl + cr) / s.dr2;
  }
  diag[q] = tmp;
}

// x_out = b_in / diag (any diagonal preconditioner)
__global__ __launch_bounds__(kBlock) void k_div_diag(int64_t n, const double* __restrict__ b,
                                                     const double* __restrict__ d, double* __restrict__ x,
                                                     const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (!kactive(st)) return;
  b = bofs(b, bs);
  d = bofs(d, bs);
  x = bofs(x, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q < n) x[q] = b[q] / d[q];
}

// Atob_uv(u, v -> x) TP/matrix_tool.f90:5-26
__global__ __launch_bounds__(kBlock) void k_velo_pack(Grid g, const double* __restrict__ u,
                                                      const double* __restrict__ v, double* __restrict__ x,
                                                      int64_t bs) {
  u = bofs(u, bs);
  v = bofs(v, bs);
  x = bofs(x, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nuv()) return;
  const int nx = g.nx;
  if (q < g.mn()) {
    const int k = g.by_nxu.div(q);
    const int i = (int)(q - (int64_t)k * g.nxu());
    x[q] = u[g.ui(i, k)];
  } else {
    const int64_t t = q - g.mn();
    const int k = g.by_nx.div(t) + 1;
    const int i = (int)(t - (int64_t)(k - 1) * nx);
    x[q] = v[g.vi(i, k)];
  }
}

// btoA_uv + the ghost rows of VeloSolve_2_combined TP/VeloSolve_2_combined.f90:426-457.
// One thread per element of the u field and one per element of the v field.
// PX: periodic side ghosts (the wall ghosts then follow from them).
template <bool PX>
__global__ __launch_bounds__(kBlock) void k_velo_unpack(Grid g, Scal s, const double* __restrict__ x,
                                                        const double* __restrict__ uL,
                                                        double* __restrict__ u, double* __restrict__ v,
                                                        int64_t bs) {
  x = bofs(x, bs);  // uL (inflow profile) is shared by the group
  u = bofs(u, bs);
  v = bofs(v, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  const int64_t nu = g.nuf();
  VeloInT<double, PX> X{{x}, nx, ny, 0, (uint32_t)g.mn()};
  if (q < nu) {
    const int k = (int)(q / (nx + 3)) - 1;
    const int i = (int)(q - (int64_t)(k + 1) * (nx + 3)) - 1;
    // interior / left-right ghosts for rows 0..ny-1, then wall rows
    auto side = [&](int ii, int kk) -> double {
      if (PX) return X.uraw(ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii), kk);
      if (ii >= 0 && ii <= nx) return X.uraw(ii, kk);
      if (s.lr_d) return ii < 0 ? 2.0 * uL[kk] - X.uraw(1, kk) : 2.0 * uL[kk] - X.uraw(nx - 1, kk);
      return ii < 0 ? X.uraw(1, kk) : X.uraw(nx - 1, kk);
    };
    double val;
    if (k < 0)
      val = s.uwB * 2.0 - side(i, 0);
    else if (k >= ny)
      val = s.uwT * 2.0 - side(i, ny - 1);
    else
      val = side(i, k);
    u[q] = val;
  } else if (q < nu + g.nvf()) {
    const int64_t t = q - nu;
    const int k = (int)(t / (nx + 2));
    const int i = (int)(t - (int64_t)k * (nx + 2)) - 1;
    double val;
    if (k == 0 || k == ny)
      val = 0.0;
    else
      val = X.V(i, k);
    v[t] = val;
  }
}

// Momentum right-hand side, one thread per unknown.
// VeloSolve_combined TP/VeloSolve_combined.f90:48-396 (BDF = 1)
// VeloSolve_2_combined TP/VeloSolve_2_combined.f90:62-400 (BDF = 2)
template <int BDF>
__global__ __launch_bounds__(kBlock) void k_velo_rhs(Grid g, Scal s, VeloRhsArgs a, double* __restrict__ rhs,
                                                     int64_t bs) {
  a.u = bofs(a.u, bs);  // every field but the shared inflow profile uL
  a.v = bofs(a.v, bs);
  a.u_pre = bofs(a.u_pre, bs);
  a.v_pre = bofs(a.v_pre, bs);
  a.p = bofs(a.p, bs);
  a.psi1 = bofs(a.psi1, bs);
  a.psi2 = bofs(a.psi2, bs);
  a.rho = bofs(a.rho, bs);
  a.eta = bofs(a.eta, bs);
  a.bu = bofs(a.bu, bs);
  a.ph_after = bofs(a.ph_after, bs);
  a.rmu_after = bofs(a.rmu_after, bs);
  a.pf1 = bofs(a.pf1, bs);
  a.pf2 = bofs(a.pf2, bs);
  a.bdr = bofs(a.bdr, bs);
  rhs = bofs(rhs, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nuv()) return;
  const int nx = g.nx, ny = g.ny;
  const double *rho = a.rho, *eta = a.eta, *bu = a.bu, *ph = a.ph_after, *mu = a.rmu_after, *bd = a.bdr;
  const double *u = a.u, *v = a.v, *up = a.u_pre, *vp = a.v_pre;
  // extrapolated velocities of the convection term (:109-110)
  auto UT = [&](int i, int k) -> double {
    return BDF == 2 ? 2.0 * u[g.ui(i, k)] - up[g.ui(i, k)] : u[g.ui(i, k)];
  };
  auto VT = [&](int i, int k) -> double {
    return BDF == 2 ? 2.0 * v[g.vi(i, k)] - vp[g.vi(i, k)] : v[g.vi(i, k)];
  };
  auto TP = [&](int i, int k) -> double {
    const int64_t c = g.ci(i, k);
    return BDF == 2 ? a.p[c] + 4 * a.psi2[c] / 3.0 - a.psi1[c] / 3.0 : a.p[c];
  };
  double f;
  if (q < g.mn()) {
    const int k = g.by_nxu.div(q);
    const int i = (int)(q - (int64_t)k * g.nxu());
    const int64_t c0 = g.ci(i, k), c1 = g.ci(i - 1, k);
    if (s.mf)  // MF/VeloSolve_2_combined.f90:64-65
      f = s.ca * (mu[c0] + mu[c1]) / 2.0 * (ph[c0] - ph[c1]) / s.dr * (1.0 - (bd[c0] + bd[c1]) / 2.0);
    else
      f = s.ca * (mu[c0] + mu[c1]) / 2.0 * (ph[c0] * bu[c0] - ph[c1] * bu[c1]) / s.dr;
    f = f - s.re * (rho[c0] + rho[c1]) / 2.0 *
                (UT(i, k) * (UT(i + 1, k) - UT(i - 1, k)) / 2.0 / s.dr +
                 (VT(i, k) + VT(i - 1, k) + VT(i, k + 1) + VT(i - 1, k + 1)) / 4.0 * (UT(i, k + 1) - UT(i, k - 1)) / 2.0 / s.dz);
    const double rb = (rho[c0] + rho[c1]) / 2.0;
    double in;
    if (BDF == 2 && s.mf)  // MF/VeloSolve_2_combined.f90:201
      in = 2.0 * u[g.ui(i, k)] * s.re / s.dt - up[g.ui(i, k)] * s.re / 2.0 / s.dt + s.ggx;
    else if (BDF == 2)
      in = 2.0 * u[g.ui(i, k)] * s.re / s.dt - up[g.ui(i, k)] * s.re / 2.0 / s.dt + (ph[c0] + ph[c1]) / 2.0 * s.ggx;
    else
      in = u[g.ui(i, k)] * s.re / s.dt + s.ggx;
    f = f - (TP(i, k) - TP(i - 1, k)) / s.dr + rb * in;
    f = f + (a.pf1[c0] + a.pf1[c1]) / 2.0;
    f = f * s.dt;
    if (s.lr_d) {
      if (i == 0) f = f + 2.0 * a.uL[k] * eta[c0] * 2.0 * s.dt / s.dr2;
      if (i == nx) f = f + 2.0 * a.uL[k] * eta[c0] * 2.0 * s.dt / s.dr2;
    }
    if (k == 0) {
      const double etay = (eta[g.ci(i, 0)] + eta[g.ci(i - 1, 0)] + eta[g.ci(i, -1)] + eta[g.ci(i - 1, -1)]) / 4.0;
      f = f + s.dt * (2.0 * s.uwB / s.dz2 * etay);
    }
    if (k == ny - 1) {
      const double etay = (eta[g.ci(i, ny)] + eta[g.ci(i - 1, ny)] + eta[g.ci(i, ny - 1)] + eta[g.ci(i - 1, ny - 1)]) / 4.0;
      f = f + s.dt * (2.0 * s.uwT / s.dz2 * etay);
    }
  } else {
    const int64_t t = q - g.mn();
    const int k = g.by_nx.div(t) + 1;
    const int i = (int)(t - (int64_t)(k - 1) * nx);
    const int64_t c0 = g.ci(i, k), c1 = g.ci(i, k - 1);
    if (s.mf)  // MF/VeloSolve_2_combined.f90:78-79
      f = s.ca * (mu[c0] + mu[c1]) / 2.0 * (ph[c0] - ph[c1]) / s.dz * (1.0 - (bd[c0] + bd[c1]) / 2.0);
    else
      f = s.ca * (mu[c0] + mu[c1]) / 2.0 * (ph[c0] * bu[c0] - ph[c1] * bu[c1]) / s.dz;
    f = f - s.re * (rho[c0] + rho[c1]) / 2.0 *
                ((UT(i, k) + UT(i + 1, k) + UT(i, k - 1) + UT(i + 1, k - 1)) / 4.0 * (VT(i + 1, k) - VT(i - 1, k)) / 2.0 / s.dr +
                 VT(i, k) * (VT(i, k + 1) - VT(i, k - 1)) / 2.0 / s.dz);
    const double rb = (rho[c0] + rho[c1]) / 2.0;
    double in;
    if (BDF == 2 && s.mf)  // MF/VeloSolve_2_combined.f90:208
      in = 2.0 * v[g.vi(i, k)] * s.re / s.dt - vp[g.vi(i, k)] * s.re / 2.0 / s.dt + s.ggz;
    else if (BDF == 2)
      in = 2.0 * v[g.vi(i, k)] * s.re / s.dt - vp[g.vi(i, k)] * s.re / 2.0 / s.dt +
           (ph[c0] + ph[g.ci(i - 1, k)]) / 2.0 * s.ggz;
    else
      in = v[g.vi(i, k)] * s.re / s.dt + s.ggz;
    f = f - (TP(i, k) - TP(i, k - 1)) / s.dz + rb * in;
    f = f + (a.pf2[c0] + a.pf2[c1]) / 2.0;
    f = f * s.dt;
  }
  rhs[q] = f;
}

// ---- Cahn-Hilliard ------------------------------------------------------

// op_double_gradient_Neum TP/matrix_tool.f90:178-219 at cell (i,j), reading
// the unknown through `get` (interior values only) and mobility `mo`.
// The reference's boundary branches replace the missing neighbour's value
// and mobility by the cell's own (i = 0: moL = (m0 + m0)/2 and r0 in place of
// the left value; likewise on the other walls).  Clamping the neighbour
// index to the cell does exactly that with the interior expression, term
// for term, so every load is unconditional and a thread issues them all at
// once (nx, ny >= 2, which drd_init requires).
// idx2, idy2 = 1.0 / dx2, 1.0 / dy2, the factors the reference evaluates
template <class R = double>
__device__ __forceinline__ R op_dgn_v(R idx2, R idy2, const Nb5T<R>& mo, const Nb5T<R>& x) {
  R moL, moR, rhs = R(0.0);
  moL = (mo.xl + mo.c) / R(2.0);
  moR = (mo.c + mo.xr) / R(2.0);
  rhs = rhs + idx2 * (moL * x.xl + moR * x.xr - (moL + moR) * x.c);
  moL = (mo.yl + mo.c) / R(2.0);
  moR = (mo.c + mo.yr) / R(2.0);
  rhs = rhs + idy2 * (moL * x.yl + moR * x.yr - (moL + moR) * x.c);
  return rhs;
}

// MATRIX_VECTOR_PH_MU_CG TP/ph_mu_solve.f90:212-284 for cell (i, j) (both
// rows) from the values around it: phi and mu (f1, f2) and the two mobilities
// (R = float: FP32 arithmetic and s.s / s.I * f1 for s.s * f1 / s.I, not the
// reference's rounding; for the FP32 preconditioner only)
template <class R = double, bool WALLS = true>
__device__ __forceinline__ void phmu_rows_v(const Grid g, const Scal& s, int j, R b, const Nb5T<R>& mi,
                                            const Nb5T<R>& m0, const Nb5T<R>& f1, const Nb5T<R>& f2, R& o1,
                                            R& o2) {
  constexpr bool kF32 = sizeof(R) == sizeof(float);
  R ff1 = op_dgn_v<R>(R(s.idr2), R(s.idz2), mi, f1);
  ff1 = (kF32 ? R(s.s_I) * f1.c : R(s.s) * f1.c / R(s.I)) - R(s.I) * ff1 - f2.c * b;
  // wall rows: ghost f1(i,-1) = f1(i,0) from btoA_ph_mu (:229-232)
  if (WALLS && j == 0) ff1 = ff1 + R(s.I_dz2) * (f1.c - R(s.coef_bot) * f1.c);
  if (WALLS && j == g.ny - 1) ff1 = ff1 + R(s.I_dz2) * (f1.c - R(s.coef_top) * f1.c);
  R ff2 = op_dgn_v<R>(R(s.idr2), R(s.idz2), m0, f2);
  ff2 = f1.c * (b + R(s.adj)) - R(s.dt) * R(s.xld) * ff2;
  o1 = ff1;
  o2 = ff2;
}

// MATRIX_VECTOR_PH_MU_CG TP/ph_mu_solve.f90:212-284 for cell q (both rows).
__device__ __forceinline__ void phmu_rows(const Grid g, const Scal& s, const double* __restrict__ bu,
                                          const double* __restrict__ minner, const double* __restrict__ m0,
                                          const double* __restrict__ x, int64_t q, double& o1, double& o2) {
  const int64_t C = g.ncell();
  const int j = g.by_nx.div(q), i = (int)(q - (int64_t)j * g.nx);
  phmu_rows_v(g, s, j, bu[g.ci(i, j)], nb5_cell(g, minner, i, j), nb5_cell(g, m0, i, j), nb5_packed(g, x, i, j),
              nb5_packed(g, x + C, i, j), o1, o2);
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_phmu_matvec(Grid g, Scal s, const double* __restrict__ bu,
                                                        const double* __restrict__ minner,
                                                        const double* __restrict__ m0,
                                                        const double* __restrict__ x, double* __restrict__ y,
                                                        double alpha, double beta, double* __restrict__ a0,
                                                        const double* __restrict__ b, double* part,
                                                        KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (MODE != MV_PLAIN && !kactive(st)) return;
  bu = bofs(bu, bs);
  minner = bofs(minner, bs);
  m0 = bofs(m0, bs);
  x = bofs(x, bs);
  y = bofs(y, bs);
  a0 = bofs(a0, bs);
  b = bofs(b, bs);
  part = bofs(part, bs);
  const int64_t C = g.ncell();
  // terms tm[seg*2 + v]: segment 0 = phi rows (element q), 1 = mu rows (C + q)
  double tm[4][kSlots] = {};
  const bool xnz = (MODE == MV_INIT) ? (st->xnz != 0) : true;
#pragma unroll
  for (int jj = 0; jj < kSlots; ++jj) {
    const int64_t q = tile_elem(C, 0, jj);
    if (q < 0) break;
    double z1, z2;
    if (MODE != MV_INIT || xnz) phmu_rows(g, s, bu, minner, m0, x, q, z1, z2);
    if (MODE == MV_PLAIN) {
      y[q] = beta == 0.0 ? alpha * z1 : alpha * z1 + beta * y[q];
      y[C + q] = beta == 0.0 ? alpha * z2 : alpha * z2 + beta * y[C + q];
    } else if (MODE == MV_INIT) {
      double r1 = b[q], r2 = b[C + q];
      if (xnz) {
        r1 = -1.0 * z1 + 1.0 * r1;
        r2 = -1.0 * z2 + 1.0 * r2;
      }
      y[q] = r1;
      y[C + q] = r2;
      a0[q] = r1;
      a0[C + q] = r2;
      tm[0][jj] = r1 * r1;
      tm[2][jj] = r2 * r2;
    } else {
      y[q] = z1;
      y[C + q] = z2;
      if (MODE == MV_DOT1) {
        tm[0][jj] = a0[q] * z1;
        tm[2][jj] = a0[C + q] * z2;
      } else {
        tm[0][jj] = z1 * a0[q];
        tm[1][jj] = z1 * z1;
        tm[2][jj] = z2 * a0[C + q];
        tm[3][jj] = z2 * z2;
      }
    }
  }
  if (MODE == MV_PLAIN) return;
  if (MODE == MV_DOT2) {
    tile_publish<2, 2>(tm, part);
  } else {
    double t1[2][kSlots];
#pragma unroll
    for (int jj = 0; jj < kSlots; ++jj) {
      t1[0][jj] = tm[0][jj];
      t1[1][jj] = tm[2][jj];
    }
    tile_publish<2, 1>(t1, part);
  }
}

// PSolve_PH_MU_CG TP/ph_mu_solve.f90:319-353 (note the reference's
// `j == nr-1` / `i == 0` tests, kept as written).
__global__ __launch_bounds__(kBlock) void k_phmu_jacobi(Grid g, Scal s, const double* __restrict__ b,
                                                        double* __restrict__ x, const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (!kactive(st)) return;
  b = bofs(b, bs);
  x = bofs(x, bs);
  const int64_t C = g.ncell();
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= C) return;
  const int nx = g.nx, ny = g.ny;
  const int j = g.by_nx.div(q), i = (int)(q - (int64_t)j * nx);
  // periodic x (no reference counterpart): the unit-mobility diagonal
  const double tmpx = (!g.px && (i == 0 || j == nx - 1)) ? 1.0 / s.dr2 : 2.0 / s.dr2;
  const double tmpz = ((g.px ? j == 0 : i == 0) || j == ny - 1) ? 1.0 / s.dz2 : 2.0 / s.dz2;
  x[q] = b[q] / ((tmpx + tmpz) * s.I + s.s / s.I);
  x[C + q] = b[C + q] / (s.dt * s.xld * (tmpx + tmpz));
}

// ch_bulk_energy TP/ch_bulk_energy.f90:4-39 and ch_bdr_condition
// TP/ch_bdr_condition.f90:43-67
// phi**3.0 is evaluated as phi*phi*phi, as the reference build does.
__device__ __forceinline__ double ch_bulk(double phi, double psi, int mode) {
  if (mode == 1) return 0.5 * phi * (1.0 - phi) * (1.0 - 2.0 * phi) * (1 - psi);
  return (phi * phi * phi - phi) * (1 - psi);
}
__device__ __forceinline__ double ch_bdr(double cf, double phv, int mode) {
  if (mode == 1) return cf * (1.0 - phv) * phv;
  return cf * (1.0 - phv * phv);
}

// ph_M_0 = (1-bdr_ph) f(ph) over the whole field TP/ph_mu_solve.f90:31-35
__global__ __launch_bounds__(kBlock) void k_phmu_m0(int64_t nc, int mode, const double* __restrict__ ph,
                                                    const double* __restrict__ bdr, double* __restrict__ m0,
                                                    int64_t bs) {
  ph = bofs(ph, bs);
  bdr = bofs(bdr, bs);
  m0 = bofs(m0, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= nc) return;
  const double p = ph[q];
  if (mode == 1)
    m0[q] = (1.0 - bdr[q]) * (p * p) * ((p - 1.0) * (p - 1.0));
  else
    m0[q] = (1.0 - bdr[q]) * ((p * p - 1.0) * (p * p - 1.0));
}

// Right-hand side of ph_mu_solve TP/ph_mu_solve.f90:40-153 and the
// initial guess (Atob_ph_mu of ph, rmu :152-153).  pn[0..nx) = pnb,
// pn[nx..2nx) = pnt for the ghost rows written after the solve.
__global__ __launch_bounds__(kBlock) void k_phmu_rhs(Grid g, Scal s, PhmuRhsArgs a, double* __restrict__ rhs,
                                                     double* __restrict__ x0, double* __restrict__ pn, int64_t bs) {
  a.ph = bofs(a.ph, bs);  // every field and the per-instance wetting coefficients
  a.rmu = bofs(a.rmu, bs);
  a.cfB = bofs(a.cfB, bs);
  a.cfT = bofs(a.cfT, bs);
  a.u = bofs(a.u, bs);
  a.v = bofs(a.v, bs);
  a.bdr = bofs(a.bdr, bs);
  a.bu = bofs(a.bu, bs);
  a.bg = bofs(a.bg, bs);
  rhs = bofs(rhs, bs);
  x0 = bofs(x0, bs);
  pn = bofs(pn, bs);
  const int64_t C = g.ncell();
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= C) return;
  const int nx = g.nx, ny = g.ny;
  const int k = g.by_nx.div(q), i = (int)(q - (int64_t)k * nx);
  const double *ph = a.ph, *bu = a.bu, *u = a.u, *v = a.v;
  const int64_t c = g.ci(i, k);
  double f2 = ph[c] * (bu[c] + s.adj);
  f2 = f2 - s.dt * (u[g.ui(i + 1, k)] + u[g.ui(i, k)]) / 2.0 *
                (ph[g.ci(i + 1, k)] * bu[g.ci(i + 1, k)] - ph[g.ci(i - 1, k)] * bu[g.ci(i - 1, k)]) / 2.0 / s.dr -
       s.dt * (v[g.vi(i, k + 1)] + v[g.vi(i, k)]) / 2.0 *
           (ph[g.ci(i, k + 1)] * bu[g.ci(i, k + 1)] - ph[g.ci(i, k - 1)] * bu[g.ci(i, k - 1)]) / 2.0 / s.dz;
  const double phv = ph[c];
  double f1;
  if (s.phi_mode == 1)
    f1 = s.s * phv / s.I - ch_bulk(phv, a.bdr[c], 1) / s.I - phv * (phv - 1.0) * a.bg[c] * a.cosa / a.sqrt2;
  else
    f1 = s.s * phv / s.I - ch_bulk(phv, a.bdr[c], 2) / s.I - (phv * phv - 1.0) * a.bg[c] * a.cosa / a.sqrt2;
  const double tmp = s.I / s.dz;
  if (k == 0) {
    const double pv = (ph[g.ci(i, -1)] + ph[g.ci(i, 0)]) / 2.0;
    double pnb = ch_bdr(a.cfB[i], pv, s.phi_mode);
    pnb = pnb / tmp;
    f1 = f1 + pnb * s.I / s.dz2;
    pn[i] = pnb;
  }
  if (k == ny - 1) {
    const double pv = (ph[g.ci(i, ny - 1)] + ph[g.ci(i, ny)]) / 2.0;
    double pnt = ch_bdr(a.cfT[i], pv, s.phi_mode);
    pnt = pnt / tmp;
    f1 = f1 + pnt * s.I / s.dz2;
    pn[nx + i] = pnt;
  }
  rhs[q] = f1;
  rhs[C + q] = f2;
  x0[q] = phv;
  x0[C + q] = a.rmu[c];
}

// btoA_ph_mu TP/matrix_tool.f90:104-136 and the ph ghost rows of
// ph_mu_solve TP/ph_mu_solve.f90:173-177.  The corner ghosts keep the value
// btoA_ph_mu gave them (copied before the ghost-row override).
__global__ __launch_bounds__(kBlock) void k_phmu_unpack(Grid g, Scal s, const double* __restrict__ x,
                                                        const double* __restrict__ pn, double* __restrict__ ph,
                                                        double* __restrict__ mu, int64_t bs) {
  x = bofs(x, bs);
  pn = bofs(pn, bs);
  ph = bofs(ph, bs);
  mu = bofs(mu, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (q >= g.nc()) return;
  const int k = (int)(q / (nx + 2)) - 1;
  const int i = (int)(q - (int64_t)(k + 1) * (nx + 2)) - 1;
  const int64_t C = g.ncell();
  const int ic = g.xcol(i);
  const int kc = k < 0 ? 0 : (k >= ny ? ny - 1 : k);
  const double p = x[ic + (int64_t)nx * kc], m = x[C + ic + (int64_t)nx * kc];
  double pv = p;
  if (g.px || (i >= 0 && i < nx)) {  // periodic x: the corners copy the wall ghosts too
    if (k < 0) pv = p * s.coef_bot + pn[ic];
    if (k >= ny) pv = p * s.coef_top + pn[nx + ic];
  }
  ph[q] = pv;
  mu[q] = m;
}

// Atob_ph_mu of two cell fields (the packed phi|mu vector) TP/matrix_tool.f90:82-98
__global__ __launch_bounds__(kBlock) void k_phmu_pack(Grid g, const double* __restrict__ ph,
                                                      const double* __restrict__ mu, double* __restrict__ x,
                                                      int64_t bs) {
  ph = bofs(ph, bs);
  mu = bofs(mu, bs);
  x = bofs(x, bs);
  const int64_t C = g.ncell();
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= C) return;
  const int k = g.by_nx.div(q), i = (int)(q - (int64_t)k * g.nx);
  x[q] = ph[g.ci(i, k)];
  x[C + q] = mu[g.ci(i, k)];
}

// ---- red-black block Gauss-Seidel for the phi/mu system -------------------
// NS full sweeps per launch: each the red cells ((i + j) even), then the
// black ones, updated as x_c += w B_c^-1 (b - A x)_c (over-relaxation,
// w = kRbOmega) with the rows of the matvec (phmu_rows_v) and the exact
// 2x2 block solve (phmu_block_v).
// A workgroup stages its tile of kRbTX x kRbTY cells and a halo of 2 NS
// cells (phi, mu, b, both mobilities, bu) in LDS.  Sweep k of NS updates the
// red cells of the tile and a ring of 2 (NS - k) - 1 cells around it, then
// the black cells of the tile and a ring one narrower; the last sweep's
// black phase covers the tile alone, which is written out.  A red cell's
// rows read only black neighbours (and a black cell's only red ones), so
// each colour updates in place in LDS, and a ring cell's value after each
// phase equals the one the tile that owns it computes, because every
// workgroup recomputes it from the same values of xi.  A launch therefore
// gives bit for bit 2 NS colour half-sweeps.  xo must not alias xi: a
// workgroup that starts after its neighbour has written would read that
// neighbour's new values in its halo.  FIRST: from x = 0 (xi unread).
// The coefficients are stored as TC, b as TB, xi as TI and xo as TO; the
// tile is held in LDS and computed in TC: FP64 (the reference's arithmetic),
// or FP32 (the FP32 preconditioner: FP32 arithmetic with reciprocals), where
// bc, if given, receives the FP32 b of the tile for the later launches.
// FP32 tiles also stage each cell's block factors (d11, d22, 1/det of
// phmu_block_f, formed once per solve by k_phmu_coef32) instead of forming
// them at every update.  A tile whose cells and neighbours all lie at least
// one cell from the walls and side columns (interior) takes the update
// without the clamped neighbour indices and wall rows, which are identities
// there.
// 32 x 16 tiles of 512 threads (a thread per cell of the widest phase): the
// FP32 middle launch at 2048 x 1024 took 29.6 us, against 31.3-52.9 us for
// 64 x 16, 32 x 32, 48 x 24, 32 x 24, 48 x 16, 24 x 16, 16 x 16, 64 x 8,
// 32 x 8 and 128 x 8 tiles (256 to 1024 threads)
constexpr int kRbTX = 32, kRbTY = 16, kRbThreads = 512;
template <bool FIRST, int NS, class TC, class TB, class TI, class TO>
__global__ __launch_bounds__(kRbThreads) void k_phmu_rb_sweep(Grid g, Scal s, const TC* __restrict__ bu,
                                                          const TC* __restrict__ minner,
                                                          const TC* __restrict__ m0, const TB* __restrict__ b,
                                                          float* __restrict__ bc, const TI* __restrict__ xi,
                                                          TO* __restrict__ xo, const KState* st, int64_t bs,
                                                          const float* __restrict__ fd11,
                                                          const float* __restrict__ fd22,
                                                          const float* __restrict__ fq) {
  constexpr int kH = 2 * NS, kW = kRbTX + 2 * kH, kHt = kRbTY + 2 * kH;
  constexpr bool BF = sizeof(TC) == sizeof(float);  // staged block factors
  constexpr int kF = BF ? kHt : 1, kFW = BF ? kW : 1;
  st = bofs(st, bs);
  if (st && (st->done || st->early)) return;
  bu = bofs(bu, bs);
  minner = bofs(minner, bs);
  m0 = bofs(m0, bs);
  b = bofs(b, bs);
  bc = bofs(bc, bs);
  xi = bofs(xi, bs);
  xo = bofs(xo, bs);
  fd11 = bofs(fd11, bs);
  fd22 = bofs(fd22, bs);
  fq = bofs(fq, bs);
  __shared__ TC sx1[kHt][kW], sx2[kHt][kW];
  __shared__ TC smi[kHt][kW], smz[kHt][kW], sbu[kHt][kW], sb1[kHt][kW], sb2[kHt][kW];
  __shared__ float sd11[kF][kFW], sd22[kF][kFW], sq[kF][kFW];
  const int nx = g.nx, ny = g.ny;
  const int64_t C = g.ncell();
  // tiles in row-major order, a contiguous band of them per XCD (tile_id),
  // so neighbouring tiles share their halo lines in one L2
  const int ntx = (nx + kRbTX - 1) / kRbTX;
  const int t = (int)tile_id();
  const int i0 = (t % ntx) * kRbTX, j0 = (t / ntx) * kRbTY;
  // staging: every load of the thread is issued before the first LDS store
  // (indices clamped into the grid; the stores keep the cells inside it),
  // held in registers as TC (the stored values are (TC) of the loads)
  constexpr int kStage = kW * kHt, kStageIt = (kStage + kRbThreads - 1) / kRbThreads;
  constexpr int kV = BF ? 10 : 7;
  TC v[kStageIt][kV];
#pragma unroll
  for (int it = 0; it < kStageIt; ++it) {
    const int e = min((int)threadIdx.x + it * kRbThreads, kStage - 1);
    const int i = min(max(i0 - kH + e % kW, 0), nx - 1), j = min(max(j0 - kH + e / kW, 0), ny - 1);
    const int64_t q = i + (int64_t)nx * j, c = g.ci(i, j);
    v[it][0] = FIRST ? TC(0) : (TC)xi[q];
    v[it][1] = FIRST ? TC(0) : (TC)xi[C + q];
    v[it][2] = (TC)minner[c];
    v[it][3] = (TC)m0[c];
    v[it][4] = (TC)bu[c];
    v[it][5] = (TC)b[q];
    v[it][6] = (TC)b[C + q];
    if (BF) {
      v[it][kV - 3] = (TC)fd11[c];
      v[it][kV - 2] = (TC)fd22[c];
      v[it][kV - 1] = (TC)fq[c];
    }
  }
#pragma unroll
  for (int it = 0; it < kStageIt; ++it) {
    const int e = (int)threadIdx.x + it * kRbThreads;
    const int li = e % kW, lj = e / kW;
    const int i = i0 - kH + li, j = j0 - kH + lj;
    if (e >= kStage || i < 0 || i >= nx || j < 0 || j >= ny) continue;
    sx1[lj][li] = v[it][0];
    sx2[lj][li] = v[it][1];
    smi[lj][li] = v[it][2];
    smz[lj][li] = v[it][3];
    sbu[lj][li] = v[it][4];
    sb1[lj][li] = v[it][5];
    sb2[lj][li] = v[it][6];
    if (BF) {
      sd11[BF ? lj : 0][BF ? li : 0] = (float)v[it][kV - 3];
      sd22[BF ? lj : 0][BF ? li : 0] = (float)v[it][kV - 2];
      sq[BF ? lj : 0][BF ? li : 0] = (float)v[it][kV - 1];
    }
  }
  __syncthreads();
  const bool interior = i0 - kH >= 0 && i0 + kRbTX + kH <= nx && j0 - kH >= 0 && j0 + kRbTY + kH <= ny;
  auto update = [&](int i, int j, auto in) {
    constexpr bool IN = decltype(in)::value;
    const int il = (IN || i > 0) ? i - 1 : i, ir = (IN || i < nx - 1) ? i + 1 : i;
    const int jl = (IN || j > 0) ? j - 1 : j, jr = (IN || j < ny - 1) ? j + 1 : j;
    const int a = i - i0 + kH, al = il - i0 + kH, ar = ir - i0 + kH;
    const int r = j - j0 + kH, rl = jl - j0 + kH, rr = jr - j0 + kH;
    const Nb5T<TC> mi{smi[r][a], smi[r][al], smi[r][ar], smi[rl][a], smi[rr][a]};
    const Nb5T<TC> mz{smz[r][a], smz[r][al], smz[r][ar], smz[rl][a], smz[rr][a]};
    const Nb5T<TC> f1{sx1[r][a], sx1[r][al], sx1[r][ar], sx1[rl][a], sx1[rr][a]};
    const Nb5T<TC> f2{sx2[r][a], sx2[r][al], sx2[r][ar], sx2[rl][a], sx2[rr][a]};
    const TC bv = sbu[r][a];
    TC z1, z2, y1, y2;
    phmu_rows_v<TC, !IN>(g, s, j, bv, mi, mz, f1, f2, z1, z2);
    if constexpr (BF)
      phmu_block_x<TC>(s, bv, (TC)sd11[r][a], (TC)sd22[r][a], (TC)sq[r][a], sb1[r][a] - z1, sb2[r][a] - z2, y1, y2);
    else
      phmu_block_v<TC>(g, s, bv, mi, mz, i, j, sb1[r][a] - z1, sb2[r][a] - z2, y1, y2);
    sx1[r][a] = f1.c + (TC)kRbOmega * y1;  // two roundings: no contraction in this build
    sx2[r][a] = f2.c + (TC)kRbOmega * y2;
  };
  // threads take the cells of one colour in row order: a row of a region
  // kRbTX + 2 w wide holds (kRbTX + 2 w) / 2 cells of each colour
  auto sweeps = [&](auto in) {
    constexpr bool IN = decltype(in)::value;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
#pragma unroll
      for (int colour = 0; colour < 2; ++colour) {
        const int w = 2 * (NS - k) - 1 - colour;  // ring width of this phase
        const int rowc = (kRbTX + 2 * w) / 2, rows = kRbTY + 2 * w;
        for (int e = threadIdx.x; e < rowc * rows; e += kRbThreads) {
          const int j = j0 - w + e / rowc;
          const int i = i0 - w + 2 * (e % rowc) + ((i0 - w + j + colour) & 1);
          if (!IN && (i < 0 || i >= nx || j < 0 || j >= ny)) continue;
          update(i, j, in);
        }
        __syncthreads();
      }
    }
  };
  if (interior)
    sweeps(std::true_type{});
  else
    sweeps(std::false_type{});
  for (int e = threadIdx.x; e < kRbTX * kRbTY; e += kRbThreads) {
    const int i = i0 + e % kRbTX, j = j0 + e / kRbTX;
    if (i >= nx || j >= ny) continue;
    const int64_t q = i + (int64_t)nx * j;
    xo[q] = (TO)sx1[j - j0 + kH][i - i0 + kH];
    xo[C + q] = (TO)sx2[j - j0 + kH][i - i0 + kH];
    if (bc) {
      bc[q] = (float)sb1[j - j0 + kH][i - i0 + kH];
      bc[C + q] = (float)sb2[j - j0 + kH][i - i0 + kH];
    }
  }
}

// interpolate_coef TP/interpolate_coef.f90:4-115 for one cell of the
// ghosted field (one thread per cell, ghosts included): a ghost takes the
// value of the nearest interior cell, as the reference's copies leave it.
// d = fpd_dist TP/particles_move.f90:361-378 at the cell centre (finit
// TP/finit.f90:44-45): 1e10, then the minimum over the particles and the
// swimmers, in that order, of sqrt(dx^2 + dy^2) - radius.  pts holds
// (x, y) pairs, np particles then nswm swimmers.
__global__ __launch_bounds__(kBlock) void k_interpolate_coef(Grid g, double dr, double dz, double I,
                                                             double lrho, double leta,
                                                             const double* __restrict__ pts, int np,
                                                             double radius, int nswm, double swm_radius,
                                                             const double* __restrict__ bdr,
                                                             double* __restrict__ fpd, double* __restrict__ rho,
                                                             double* __restrict__ eta, int64_t bs) {
  bdr = bofs(bdr, bs);
  fpd = bofs(fpd, bs);
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nc()) return;
  const int ld = g.nx + 2;
  const int k = (int)(q / ld) - 1, i = (int)(q - (int64_t)(k + 1) * ld) - 1;
  const int ic = g.xcol(i);  // periodic x: the ghost column's periodic copy
  const int kc = k < 0 ? 0 : (k > g.ny - 1 ? g.ny - 1 : k);
  const double x = (ic + 0.5) * dr, y = (kc + 0.5) * dz;
  double d = 1.0E10;
  for (int p = 0; p < np + nswm; ++p) {
    const double ddx = x - pts[2 * p], ddy = y - pts[2 * p + 1];
    const double e = sqrt(ddx * ddx + ddy * ddy) - (p < np ? radius : swm_radius);
    d = fmin(d, e);
  }
  double f = tanh(-1.0 * d / sqrt(2.0) / I);
  f = (f + 1.0) / 2.0;
  f = f + bdr[g.ci(ic, kc)];
  fpd[q] = f;
  rho[q] = 1.0 + (lrho - 1.0) * f;
  eta[q] = 1.0 + (leta - 1.0) * f;
}

// ---- particle coupling (TP/particles_move.f90:326-357, 758-1008) ----------
// A body is 5 doubles: centre x, y, radius, force x, y.  Its weight at a cell
// centre is (tanh(-d / sqrt(2) / I) + 1) / 2 with d the periodic distance of
// fpd_dist_periodic: the nearer image across each side within one radius.
// Beyond kBodyCut * sqrt(2) * I from a body its tanh argument is below -20,
// where tanh is -1 to double precision (1 - |tanh| < 1e-17), so the weight is
// 0 and the sums skip it.
constexpr double kBodyCut = 20.0;

__device__ __forceinline__ double body_weight(double x, double y, double lx, double ly, const double* b,
                                              double I) {
  const double bx = b[0], by = b[1], r = b[2];
  double xs = 0.0, ys = 0.0;
  if (bx < r) xs = -lx;
  else if (lx - bx < r) xs = lx;
  if (by < r) ys = -ly;
  else if (ly - by < r) ys = ly;
  const double dx0 = x - bx, dx1 = x + xs - bx, dy0 = y - by, dy1 = y + ys - by;
  double d = sqrt(dx0 * dx0 + dy0 * dy0) - r;
  d = fmin(d, sqrt(dx1 * dx1 + dy0 * dy0) - r);
  d = fmin(d, sqrt(dx0 * dx0 + dy1 * dy1) - r);
  d = fmin(d, sqrt(dx1 * dx1 + dy1 * dy1) - r);
  if (d >= kBodyCut * sqrt(2.0) * I) return 0.0;
  return (tanh(-1.0 * d / sqrt(2.0) / I) + 1.0) / 2.0;
}

// True when every cell centre of rows [k0, k1) lies beyond the cut of body b
// (its centre and the y image fpd_dist_periodic may use).
__device__ __forceinline__ bool band_beyond(int k0, int k1, double dz, double ly, const double* b, double I) {
  const double by = b[1], r = b[2], y0 = (k0 + 0.5) * dz, y1 = (k1 - 0.5) * dz;
  double ys = 0.0;
  if (by < r) ys = -ly;
  else if (ly - by < r) ys = ly;
  const double lim = r + kBodyCut * sqrt(2.0) * I;
  const double c0 = by, c1 = by - ys;
  const double g0 = fmax(0.0, fmax(y0 - c0, c0 - y1)), g1 = fmax(0.0, fmax(y0 - c1, c1 - y1));
  return fmin(g0, g1) > lim * (1.0 + 1e-12);
}

// Per body (blockIdx.y) and row chunk (blockIdx.x): the interior sums of the
// weight and, with VEL, of the weighted row-averaged velocities
// (u(i,k) + u(i,k+1)) / 2 and (v(i,k) + v(i,k+1)) / 2.  part[(b * chunks + c) * 3 + j].
template <bool VEL>
__global__ __launch_bounds__(kBlock) void k_body_sums(Grid g, double dr, double dz, double I, double lx, double ly,
                                                      const double* __restrict__ bodies,
                                                      const double* __restrict__ u, const double* __restrict__ v,
                                                      double* __restrict__ part) {
  const int chunks = gridDim.x, c = blockIdx.x, b = blockIdx.y;
  const int k0 = (int)((int64_t)g.ny * c / chunks), k1 = (int)((int64_t)g.ny * (c + 1) / chunks);
  const double* body = bodies + 5 * (int64_t)b;
  double s = 0.0, su = 0.0, sv = 0.0;
  // A body with no x image (fpd_dist_periodic takes one within a radius of
  // x = 0 or lx) weighs 0 at every column whose centre is beyond the cut.
  int ia = 0, ib = g.nx - 1;
  if (!(body[0] < body[2]) && !(lx - body[0] < body[2])) {
    const double lim = (body[2] + kBodyCut * sqrt(2.0) * I) * (1.0 + 1e-12);
    ia = (int)fmax(0.0, floor((body[0] - lim) / dr - 0.5));
    ib = (int)fmin((double)(g.nx - 1), ceil((body[0] + lim) / dr - 0.5));
  }
  const int wdt = ib - ia + 1;
  const int64_t n = (band_beyond(k0, k1, dz, ly, body, I) || wdt <= 0) ? 0 : (int64_t)(k1 - k0) * wdt;
  for (int64_t q = threadIdx.x; q < n; q += kBlock) {
    const int k = k0 + (int)(q / wdt), i = ia + (int)(q % wdt);
    const double w = body_weight((i + 0.5) * dr, (k + 0.5) * dz, lx, ly, body, I);
    s += w;
    if (VEL) {
      su += (u[g.ui(i, k)] + u[g.ui(i, k + 1)]) / 2.0 * w;
      sv += (v[g.vi(i, k)] + v[g.vi(i, k + 1)]) / 2.0 * w;
    }
  }
  __shared__ double red[3][kBlock];
  red[0][threadIdx.x] = s;
  red[1][threadIdx.x] = su;
  red[2][threadIdx.x] = sv;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[((int64_t)b * chunks + c) * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// Per body: the chunk sums in chunk order; fin[3b] = sum * dr * dz (the force
// normaliser), fin[3b+1], fin[3b+2] = the weighted velocity averages.
__global__ __launch_bounds__(kBlock) void k_body_finish(int nbody, int chunks, double dr, double dz,
                                                        const double* __restrict__ part, double* __restrict__ fin) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= nbody) return;
  double s = 0.0, su = 0.0, sv = 0.0;
  for (int c = 0; c < chunks; ++c) {
    const double* pc = part + ((int64_t)b * chunks + c) * 3;
    s += pc[0];
    su += pc[1];
    sv += pc[2];
  }
  fin[3 * b] = s * dr * dz;
  fin[3 * b + 1] = su / s;
  fin[3 * b + 2] = sv / s;
}

// True when every cell centre of the box [i0, i1] x [k0, k1] lies beyond the
// cut of body b, for each of the centres fpd_dist_periodic may use (a lower
// bound of the distance: the nearest x image and the nearest y image).
__device__ __forceinline__ bool box_beyond(int i0, int i1, int k0, int k1, double dr, double dz, double lx,
                                           double ly, const double* b, double I) {
  const double bx = b[0], by = b[1], r = b[2];
  double xs = 0.0, ys = 0.0;
  if (bx < r) xs = -lx;
  else if (lx - bx < r) xs = lx;
  if (by < r) ys = -ly;
  else if (ly - by < r) ys = ly;
  const double x0 = (i0 + 0.5) * dr, x1 = (i1 + 0.5) * dr, y0 = (k0 + 0.5) * dz, y1 = (k1 + 0.5) * dz;
  const double cx1 = bx - xs, cy1 = by - ys;
  const double gx = fmin(fmax(0.0, fmax(x0 - bx, bx - x1)), fmax(0.0, fmax(x0 - cx1, cx1 - x1)));
  const double gy = fmin(fmax(0.0, fmax(y0 - by, by - y1)), fmax(0.0, fmax(y0 - cy1, cy1 - y1)));
  const double lim = r + kBodyCut * sqrt(2.0) * I;
  return sqrt(gx * gx + gy * gy) > lim * (1.0 + 1e-12);
}

// particle_force1/2 per ghosted cell on kSpTX x kSpTY tiles: the bodies' (F *
// weight) / normaliser added in body order; ghosts 0.  Per batch of kBlock
// bodies the tile keeps, in order, those that can reach one of its cells
// (wave ballots, then per-wave offsets); the others add exactly 0 at every
// cell of the tile.
constexpr int kSpTX = 64, kSpTY = kBlock / kSpTX;

__global__ __launch_bounds__(kBlock) void k_body_spread(Grid g, double dr, double dz, double I, double lx, double ly,
                                                        const double* __restrict__ bodies, int nbody,
                                                        const double* __restrict__ fin, double* __restrict__ pf1,
                                                        double* __restrict__ pf2) {
  constexpr int kWaves = kBlock / 64;
  __shared__ int list[kBlock];
  __shared__ int wcount[kWaves];
  const int tx = (int)threadIdx.x % kSpTX, ty = (int)threadIdx.x / kSpTX;
  const int ox = (int)blockIdx.x * kSpTX - 1, oy = (int)blockIdx.y * kSpTY - 1;  // ghosted tile origin
  const int i = ox + tx, k = oy + ty;
  const int i0 = max(ox, 0), i1 = min(ox + kSpTX - 1, g.nx - 1);
  const int k0 = max(oy, 0), k1 = min(oy + kSpTY - 1, g.ny - 1);
  const bool empty = i0 > i1 || k0 > k1;
  const bool inside = i >= 0 && i < g.nx && k >= 0 && k < g.ny;
  const double x = (i + 0.5) * dr, y = (k + 0.5) * dz;
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  double a1 = 0.0, a2 = 0.0;
  for (int base = 0; base < nbody; base += kBlock) {
    const int b = base + (int)threadIdx.x;
    const bool keep = b < nbody && !empty && !box_beyond(i0, i1, k0, k1, dr, dz, lx, ly, bodies + 5 * (int64_t)b, I);
    const unsigned long long m = __ballot(keep);
    if (lane == 0) wcount[wave] = __popcll(m);
    __syncthreads();
    int off = 0, total = 0;
    for (int w = 0; w < kWaves; ++w) {
      off += w < wave ? wcount[w] : 0;
      total += wcount[w];
    }
    if (keep) list[off + __popcll(m & ((1ull << lane) - 1ull))] = b;
    __syncthreads();
    if (inside)
      for (int j = 0; j < total; ++j) {
        const int bj = list[j];
        const double* body = bodies + 5 * (int64_t)bj;
        const double w = body_weight(x, y, lx, ly, body, I), S = fin[3 * bj];
        if (w == 0.0) continue;  // adds +0
        a1 = a1 + body[3] * w / S;
        a2 = a2 + body[4] * w / S;
      }
    __syncthreads();  // list and wcount are rewritten by the next batch
  }
  if (i <= g.nx && k <= g.ny) {
    const int64_t q = g.ci(i, k);
    pf1[q] = a1;
    pf2[q] = a2;
  }
}

// FP32 copies of the CH block coefficients bu, M_inner, M_0 (ghosts included)
// and, at the interior cells, the FP32 block factors of phmu_block_f<float>
// from those copies (the red-black sweeps read them instead of forming them
// at every update)
__global__ __launch_bounds__(kBlock) void k_phmu_coef32(Grid g, Scal s, const double* __restrict__ bu,
                                                        const double* __restrict__ mi,
                                                        const double* __restrict__ m0, PhmuCoef32 o, int64_t bs) {
  bu = bofs(bu, bs);
  mi = bofs(mi, bs);
  m0 = bofs(m0, bs);
  float* bu32 = bofs(const_cast<float*>(o.bu), bs);
  float* mi32 = bofs(const_cast<float*>(o.minner), bs);
  float* m032 = bofs(const_cast<float*>(o.m0), bs);
  float* d11 = bofs(const_cast<float*>(o.d11), bs);
  float* d22 = bofs(const_cast<float*>(o.d22), bs);
  float* qq = bofs(const_cast<float*>(o.q), bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nc()) return;
  bu32[q] = (float)bu[q];
  mi32[q] = (float)mi[q];
  m032[q] = (float)m0[q];
  const int ld = g.nx + 2;
  const int j = (int)(q / ld) - 1, i = (int)(q - (int64_t)(j + 1) * ld) - 1;
  float f11 = 0.f, f22 = 0.f, fq = 0.f;
  if (i >= 0 && i < g.nx && j >= 0 && j < g.ny) {
    const Nb5 a = nb5_cell(g, mi, i, j), z = nb5_cell(g, m0, i, j);
    const Nb5T<float> af{(float)a.c, (float)a.xl, (float)a.xr, (float)a.yl, (float)a.yr};
    const Nb5T<float> zf{(float)z.c, (float)z.xl, (float)z.xr, (float)z.yl, (float)z.yr};
    phmu_block_f<float>(g, s, (float)bu[q], af, zf, i, j, f11, f22, fq);
  }
  d11[q] = f11;
  d22[q] = f22;
  qq[q] = fq;
}

// ---- multigrid pieces for the BDF2 velocity operator ----------------------
// Level l+1 halves level l in both directions.  The transfers between the
// packed (u, v) vectors of two levels are staggered bilinear (prolongation
// P) and R = P^T / 4 (restriction); both evaluate one set of 1-D weights, so
// R is exactly the transpose of P up to the factor.  P[f, C] per direction:
//   u along x (faces 0..2n <- 0..n): even f: C = f/2 (1); odd f: f/2, f/2+1 (1/2 each)
//   u along y (cells, odd walls):    C = f/2 (3/4), its other neighbour (1/4);
//                                    past a wall the neighbour is -C (3/4 - 1/4)
//   v along x (cells, even walls):   C = f/2 (3/4), other neighbour (1/4), clamped
//   v along y (inner faces 1..2n-1 <- 1..n-1): as u along x, wall faces dropped
__device__ __forceinline__ double mg_w_face(int f, int C) {
  return (f & 1) ? ((C == (f >> 1) || C == (f >> 1) + 1) ? 0.5 : 0.0) : (C == (f >> 1) ? 1.0 : 0.0);
}
__device__ __forceinline__ double mg_w_cell_odd(int f, int C, int nc) {
  const int c = f >> 1, o = (f & 1) ? c + 1 : c - 1;
  const bool oin = o >= 0 && o < nc;
  double w = 0.0;
  if (C == c) w = oin ? 0.75 : 0.5;
  if (oin && C == o) w = 0.25;
  return w;
}
__device__ __forceinline__ double mg_w_cell_even(int f, int C, int nc) {
  const int c = f >> 1;
  int o = (f & 1) ? c + 1 : c - 1;
  o = o < 0 ? 0 : (o > nc - 1 ? nc - 1 : o);
  return (C == c ? 0.75 : 0.0) + (C == o ? 0.25 : 0.0);
}

// Level 0's rho, eta, chi (ghosts included) stored as FP32.
__global__ __launch_bounds__(kBlock) void k_mg_convert(int64_t n, const double* __restrict__ r,
                                                       const double* __restrict__ e,
                                                       const double* __restrict__ c, float* __restrict__ rf,
                                                       float* __restrict__ ef, float* __restrict__ cf, int64_t bs) {
  r = bofs(r, bs);
  e = bofs(e, bs);
  c = bofs(c, bs);
  rf = bofs(rf, bs);
  ef = bofs(ef, bs);
  cf = bofs(cf, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  rf[q] = (float)r[q];
  ef[q] = (float)e[q];
  cf[q] = (float)c[q];
}

// Coarse cell field (ghosts included, one thread per coarse cell): the 2x2
// average of the fine interior cells under it; a ghost takes the value of
// the nearest interior cell.
template <class T, class TO>
__global__ __launch_bounds__(kBlock) void k_mg_coarsen(Grid gf, Grid gc, const T* __restrict__ rf,
                                                       const T* __restrict__ ef, const T* __restrict__ cf,
                                                       TO* __restrict__ rc, TO* __restrict__ ec,
                                                       TO* __restrict__ cc, int64_t bs) {
  rf = bofs(rf, bs);
  ef = bofs(ef, bs);
  cf = bofs(cf, bs);
  rc = bofs(rc, bs);
  ec = bofs(ec, bs);
  cc = bofs(cc, bs);
  const int64_t q = (int64_t)tile_id() * kBlock + threadIdx.x;  // XCD-contiguous blocks
  if (q >= gc.nc()) return;
  const int ld = gc.nx + 2;
  const int k = (int)(q / ld) - 1, i = (int)(q - (int64_t)(k + 1) * ld) - 1;
  const int ic = i < 0 ? 0 : (i > gc.nx - 1 ? gc.nx - 1 : i);
  const int kc = k < 0 ? 0 : (k > gc.ny - 1 ? gc.ny - 1 : k);
  const int64_t a = gf.ci(2 * ic, 2 * kc), b = gf.ci(2 * ic, 2 * kc + 1);
  const int64_t c = gf.ci(2 * ic + 1, 2 * kc), d = gf.ci(2 * ic + 1, 2 * kc + 1);
  auto avg = [&](const T* f) {
    double v = (double)f[a] + (double)f[b];
    v = v + (double)f[c];
    v = v + (double)f[d];
    return (TO)(0.25 * v);
  };
  rc[q] = avg(rf);
  ec[q] = avg(ef);
  cc[q] = avg(cf);
}

// dw = omega / (the exact diagonal): the smoothers multiply by it
template <class T>
__global__ __launch_bounds__(kBlock) void k_mg_diag(Grid g, Scal s, const T* __restrict__ rho,
                                                    const T* __restrict__ eta, const T* __restrict__ chi,
                                                    T* __restrict__ dw, double omega, int64_t bs) {
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  chi = bofs(chi, bs);
  dw = bofs(dw, bs);
  const int64_t q = (int64_t)tile_id() * kBlock + threadIdx.x;  // XCD-contiguous blocks
  if (q < g.nuv()) dw[q] = (T)(omega / velo_diag2(g, s, rho, eta, chi, q));
}

// x of the first sweep (x = omega b / d from 0) at packed element i, in the
// level's type: the second sweep forms its neighbours' values on the fly
template <class T, class TB>
struct LdFirst {
  using R = T;
  const TB* __restrict__ b;
  const T* __restrict__ dw;
  __device__ __forceinline__ T operator()(uint32_t i) const { return (T)ld32(b, i) * ld32(dw, i); }
};
// levels below this many unknowns are latency-bound: their kernels issue the
// data loads before testing the solve's state
constexpr uint32_t kMgLate = 1u << 18;

// Damped Jacobi sweep xo = x + (b - A x) dw, dw = omega / d; FIRST: from
// x = 0, xo = b dw, and bc (if given) receives b as FP32.  SWEEP0: FIRST and
// one sweep in one pass (the neighbours' first-sweep values formed on the
// fly, rounded as FIRST stores them); bc as for FIRST.  RES: r = b - A x.
// The level's coefficients, x and dw are stored as T; b as TB, the output as
// TO.  A sweep's b is the level's own (FP32 levels: rounded to FP32, as the
// copy the later sweeps read).
template <int KIND, class T, class TB, class TO>
__global__ __launch_bounds__(kBlock) void k_mg_smooth(Grid g, Scal s, const T* __restrict__ rho,
                                                      const T* __restrict__ eta, const T* __restrict__ chi,
                                                      const T* __restrict__ x, const TB* __restrict__ b,
                                                      const T* __restrict__ dw, TO* __restrict__ out,
                                                      float* __restrict__ bc, const KState* st, int64_t ibs) {
  rho = bofs(rho, ibs);
  eta = bofs(eta, ibs);
  chi = bofs(chi, ibs);
  x = bofs(x, ibs);
  b = bofs(b, ibs);
  dw = bofs(dw, ibs);
  out = bofs(out, ibs);
  bc = bofs(bc, ibs);
  st = bofs(st, ibs);
  const uint32_t n = (uint32_t)g.nuv();
  const bool late = n < kMgLate;
  if (!late && st && (st->done || st->early)) return;
  const uint32_t q = tile_id() * kBlock + threadIdx.x;  // XCD-contiguous blocks
  if (q >= n) return;
  const TB bq = b[q];
  const double bs = (double)(T)bq, w = (double)dw[q];
  if (KIND == MG_FIRST) {
    if (late && st && (st->done || st->early)) return;
    out[q] = (TO)((double)bq * w);
    if (bc) bc[q] = (float)bq;
    return;
  }
  double z, xq;
  // FP32 levels: the regrouped form of velo_row_pc
  constexpr bool PC = sizeof(T) == sizeof(float);
  if (KIND == MG_SWEEP0) {
    const VeloAcc<LdFirst<T, TB>> X{{b, dw}, g.nx, g.ny, s.lr_d, (uint32_t)g.mn()};
    if constexpr (PC)
      z = velo_row_pc(g, s, rho, eta, chi, X, q);
    else
      z = velo_row<2>(g, s, rho, eta, chi, X, q);
    xq = X.ld(q);
  } else {
    const VeloInT<T> X{{x}, g.nx, g.ny, s.lr_d, (uint32_t)g.mn()};
    if constexpr (PC)
      z = velo_row_pc(g, s, rho, eta, chi, X, q);
    else
      z = velo_row<2>(g, s, rho, eta, chi, X, q);
    xq = (KIND == MG_RES) ? 0.0 : (double)x[q];
  }
  if (late && st && (st->done || st->early)) return;
  if (KIND == MG_RES) {
    out[q] = (TO)(bs - z);
  } else {
    out[q] = (TO)(xq + (bs - z) * w);
    if (KIND == MG_SWEEP0 && bc) bc[q] = (float)bq;
  }
}

// bc = R r = P^T r / 4 (one thread per coarse unknown)
template <class TR, class TB>
__global__ __launch_bounds__(kBlock) void k_mg_restrict(Grid gf, Grid gc, const TR* __restrict__ r,
                                                        TB* __restrict__ bc, const KState* st, int64_t bs) {
  r = bofs(r, bs);
  bc = bofs(bc, bs);
  st = bofs(st, bs);
  const bool late = gf.nuv() < kMgLate;
  if (!late && st && (st->done || st->early)) return;
  const int64_t q = (int64_t)tile_id() * kBlock + threadIdx.x;  // XCD-contiguous blocks
  if (q >= gc.nuv()) return;
  const PackedUV<TR> R{r, gf.nx, gf.mn()};
  double acc;
  if (q < gc.mn()) {
    const int K = (int)(q / (gc.nx + 1)), I = (int)(q - (int64_t)K * (gc.nx + 1));
    acc = mg_restrict_u(gf, gc, R, I, K);
  } else {
    const int64_t t = q - gc.mn();
    const int K = (int)(t / gc.nx) + 1, I = (int)(t - (int64_t)(K - 1) * gc.nx);
    acc = mg_restrict_v(gf, gc, R, I, K);
  }
  if (late && st && (st->done || st->early)) return;
  bc[q] = (TB)(0.25 * acc);
}

// x += P xc (one thread per fine unknown)
template <class TC, class TF>
__global__ __launch_bounds__(kBlock) void k_mg_prolong_add(Grid gc, Grid gf, const TC* __restrict__ xc,
                                                           TF* __restrict__ x, const KState* st, int64_t bs) {
  xc = bofs(xc, bs);
  x = bofs(x, bs);
  st = bofs(st, bs);
  const bool late = gf.nuv() < kMgLate;
  if (!late && st && (st->done || st->early)) return;
  const int64_t q = (int64_t)tile_id() * kBlock + threadIdx.x;  // XCD-contiguous blocks
  if (q >= gf.nuv()) return;
  double acc;
  int I[4], K[4];
  double xv[4];
  if (q < gf.mn()) {
    const int k = (int)(q / (gf.nx + 1)), i = (int)(q - (int64_t)k * (gf.nx + 1));
    mg_prolong_u_at(gc, i, k, I, K);
#pragma unroll
    for (int m = 0; m < 4; ++m) xv[m] = (double)xc[mg_uidx(gc, I[m], K[m])];
    acc = mg_prolong_u_sum(gc, i, k, xv);
  } else {
    const int64_t t = q - gf.mn();
    const int k = (int)(t / gf.nx) + 1, i = (int)(t - (int64_t)(k - 1) * gf.nx);
    mg_prolong_v_at(gc, i, k, I, K);
#pragma unroll
    for (int m = 0; m < 4; ++m) xv[m] = (double)xc[mg_vidx(gc, I[m], K[m])];
    acc = mg_prolong_v_sum(gc, i, k, xv);
  }
  const double xq = (double)x[q];
  if (late && st && (st->done || st->early)) return;
  x[q] = (TF)(xq + acc);
}

// ---- host launchers -----------------------------------------------------
static inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

void launch_velo_matvec(hipStream_t st, int bdf, int mode, const Grid& g, const Scal& s, const Coef& c,
                        const double* x, double* y, double alpha, double beta, double* a0,
                        const double* b, double* part, KState* ks, bool finish, hipEvent_t ev0,
                        hipEvent_t ev1) {
  const unsigned ntiles = nblk(g.nuv(), kTile);
  dim3 grid((ntiles + 1) / 2, 1, bat_n());  // two tiles per workgroup
#define VM(B, M)                                                                                               \
  do {                                                                                                         \
    if (g.px)                                                                                                  \
      hipExtLaunchKernelGGL((k_velo_matvec<B, M, true>), grid, dim3(kBlock), 0, st, ev0, ev1, 0u, g, s, c.rho, \
                            c.eta, c.chi, x, y, alpha, beta, a0, b, part, ks, bat_bs());                       \
    else                                                                                                       \
      hipExtLaunchKernelGGL((k_velo_matvec<B, M, false>), grid, dim3(kBlock), 0, st, ev0, ev1, 0u, g, s,       \
                            c.rho, c.eta, c.chi, x, y, alpha, beta, a0, b, part, ks, bat_bs());                \
  } while (0)
  if (bdf == 1) {
    switch (mode) {
      case MV_PLAIN: VM(1, MV_PLAIN); break;
      case MV_DOT1: VM(1, MV_DOT1); break;
      case MV_DOT2: VM(1, MV_DOT2); break;
      case MV_INIT: VM(1, MV_INIT); break;
    }
  } else {
    switch (mode) {
      case MV_PLAIN: VM(2, MV_PLAIN); break;
      case MV_DOT1: VM(2, MV_DOT1); break;
      case MV_DOT2: VM(2, MV_DOT2); break;
      case MV_INIT: VM(2, MV_INIT); break;
    }
  }
#undef VM
  if (finish) launch_mv_finalize(st, mode, part, ntiles, 1, ks);
}

void launch_phmu_matvec(hipStream_t st, int mode, const Grid& g, const Scal& s, const Coef& c,
                        const double* x, double* y, double alpha, double beta, double* a0,
                        const double* b, double* part, KState* ks, bool finish) {
  dim3 grid(nblk(g.ncell(), kTile), 1, bat_n());
#define PM(M)                                                                                          \
  hipLaunchKernelGGL((k_phmu_matvec<M>), grid, dim3(kBlock), 0, st, g, s, c.bu, c.minner, c.m0, x, y, alpha, \
                     beta, a0, b, part, ks, bat_bs())
  switch (mode) {
    case MV_PLAIN: PM(MV_PLAIN); break;
    case MV_DOT1: PM(MV_DOT1); break;
    case MV_DOT2: PM(MV_DOT2); break;
    case MV_INIT: PM(MV_INIT); break;
  }
#undef PM
  if (finish) launch_mv_finalize(st, mode, part, grid.x, 2, ks);
}

void launch_velo_jacobi_diag(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, double* diag,
                             double mcoef) {
  hipLaunchKernelGGL(k_velo_jacobi_diag, dim3(nblk(g.nuv(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, s,
                     c.rho, c.eta, diag, mcoef, bat_bs());
}

void launch_div_diag(hipStream_t st, int64_t n, const double* b, const double* d, double* x,
                     const KState* ks) {
  hipLaunchKernelGGL(k_div_diag, dim3(nblk(n, kBlock), 1, bat_n()), dim3(kBlock), 0, st, n, b, d, x, ks, bat_bs());
}

void launch_phmu_jacobi(hipStream_t st, const Grid& g, const Scal& s, const double* b, double* x,
                        const KState* ks) {
  hipLaunchKernelGGL(k_phmu_jacobi, dim3(nblk(g.ncell(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, s, b, x, ks,
                     bat_bs());
}

void launch_velo_pack(hipStream_t st, const Grid& g, const double* u, const double* v, double* x) {
  hipLaunchKernelGGL(k_velo_pack, dim3(nblk(g.nuv(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, u, v, x,
                     bat_bs());
}

void launch_velo_unpack(hipStream_t st, const Grid& g, const Scal& s, const double* x, const double* uL,
                        double* u, double* v) {
  if (g.px)
    hipLaunchKernelGGL(k_velo_unpack<true>, dim3(nblk(g.nuf() + g.nvf(), kBlock), 1, bat_n()), dim3(kBlock), 0, st,
                       g, s, x, uL, u, v, bat_bs());
  else
    hipLaunchKernelGGL(k_velo_unpack<false>, dim3(nblk(g.nuf() + g.nvf(), kBlock), 1, bat_n()), dim3(kBlock), 0, st,
                       g, s, x, uL, u, v, bat_bs());
}

// Periodic x: the ghost columns of a field from its interior, over every row
// (wall-ghost rows included): cell and v fields f(-1) = f(nx-1), f(nx) =
// f(0); u fields also f(nx+1) = f(1).  One thread per row.
__global__ __launch_bounds__(kBlock) void k_xghost_fill(Grid g, double* __restrict__ f, int kind, int64_t bs) {
  f = bofs(f, bs);
  const int r = (int)(blockIdx.x * kBlock + threadIdx.x);
  const int rows = kind == 2 ? g.ny + 1 : g.ny + 2, ld = kind == 1 ? g.nx + 3 : g.nx + 2;
  if (r >= rows) return;
  double* row = f + (int64_t)r * ld;  // element i of the row at row[i + 1]
  row[0] = row[g.nx];
  row[g.nx + 1] = row[1];
  if (kind == 1) row[g.nx + 2] = row[2];
}

void launch_xghost_fill(hipStream_t st, const Grid& g, double* f, int kind) {
  hipLaunchKernelGGL(k_xghost_fill, dim3(nblk(g.ny + 2, kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, f, kind,
                     bat_bs());
}

void launch_velo_rhs(hipStream_t st, int bdf, const Grid& g, const Scal& s, const VeloRhsArgs& a,
                     double* rhs) {
  if (bdf == 1)
    hipLaunchKernelGGL(k_velo_rhs<1>, dim3(nblk(g.nuv(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, s, a, rhs,
                       bat_bs());
  else
    hipLaunchKernelGGL(k_velo_rhs<2>, dim3(nblk(g.nuv(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, s, a, rhs,
                       bat_bs());
}

void launch_phmu_m0(hipStream_t st, const Grid& g, int mode, const double* ph, const double* bdr,
                    double* m0) {
  hipLaunchKernelGGL(k_phmu_m0, dim3(nblk(g.nc(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g.nc(), mode, ph, bdr,
                     m0, bat_bs());
}

void launch_phmu_rhs(hipStream_t st, const Grid& g, const Scal& s, const PhmuRhsArgs& a, double* rhs,
                     double* x0, double* pn) {
  hipLaunchKernelGGL(k_phmu_rhs, dim3(nblk(g.ncell(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, s, a, rhs, x0,
                     pn, bat_bs());
}

void launch_phmu_pack(hipStream_t st, const Grid& g, const double* ph, const double* mu, double* x) {
  hipLaunchKernelGGL(k_phmu_pack, dim3(nblk(g.ncell(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, ph, mu, x,
                     bat_bs());
}

void launch_phmu_unpack(hipStream_t st, const Grid& g, const Scal& s, const double* x, const double* pn,
                        double* ph, double* mu) {
  hipLaunchKernelGGL(k_phmu_unpack, dim3(nblk(g.nc(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, s, x, pn, ph,
                     mu, bat_bs());
}


// ---- multigrid launchers ------------------------------------------------------
void launch_mg_convert(hipStream_t st, const Grid& g, const Coef& c, float* rho, float* eta, float* chi) {
  hipLaunchKernelGGL(k_mg_convert, dim3(nblk(g.nc(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g.nc(), c.rho, c.eta,
                     c.chi, rho, eta, chi, bat_bs());
}
void launch_mg_coarsen(hipStream_t st, bool f32, bool o32, const Grid& gf, const Grid& gc, const MgCoef& f,
                       void* rc, void* ec, void* cc) {
  const dim3 grid(nblk(gc.nc(), kBlock), 1, bat_n());
#define MGC(T, TO)                                                                                     \
  hipLaunchKernelGGL((k_mg_coarsen<T, TO>), grid, dim3(kBlock), 0, st, gf, gc, (const T*)f.rho,           \
                     (const T*)f.eta, (const T*)f.chi, (TO*)rc, (TO*)ec, (TO*)cc, bat_bs())
  if (f32 && o32)
    MGC(float, float);
  else if (f32)  // into an FP64 coarsest level
    MGC(float, double);
  else
    MGC(double, double);
#undef MGC
}
void launch_mg_diag(hipStream_t st, bool f32, const Grid& g, const Scal& s, const MgCoef& c, void* dw,
                    double omega) {
  const dim3 grid(nblk(g.nuv(), kBlock), 1, bat_n());
  if (f32)
    hipLaunchKernelGGL(k_mg_diag<float>, grid, dim3(kBlock), 0, st, g, s, (const float*)c.rho, (const float*)c.eta,
                       (const float*)c.chi, (float*)dw, omega, bat_bs());
  else
    hipLaunchKernelGGL(k_mg_diag<double>, grid, dim3(kBlock), 0, st, g, s, (const double*)c.rho,
                       (const double*)c.eta, (const double*)c.chi, (double*)dw, omega, bat_bs());
}
template <int K, class T, class TB, class TO>
static void mg_smooth_t(hipStream_t st, const Grid& g, const Scal& s, const MgCoef& c, const void* x,
                        const void* b, const void* dw, void* out, float* bc, const KState* ks) {
  hipLaunchKernelGGL((k_mg_smooth<K, T, TB, TO>), dim3(nblk(g.nuv(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g,
                     s, (const T*)c.rho, (const T*)c.eta, (const T*)c.chi, (const T*)x, (const TB*)b, (const T*)dw,
                     (TO*)out, bc, ks, bat_bs());
}
template <int K>
static void mg_smooth_k(hipStream_t st, bool f32, bool b32, bool o32, const Grid& g, const Scal& s,
                        const MgCoef& c, const void* x, const void* b, const void* dw, void* out, float* bc,
                        const KState* ks) {
  if (!f32)  // an FP64 hierarchy is FP64 throughout
    mg_smooth_t<K, double, double, double>(st, g, s, c, x, b, dw, out, bc, ks);
  else if (b32 && o32)
    mg_smooth_t<K, float, float, float>(st, g, s, c, x, b, dw, out, bc, ks);
  else if (b32)  // the last sweep of level 0: FP64 output
    mg_smooth_t<K, float, float, double>(st, g, s, c, x, b, dw, out, bc, ks);
  else  // the first sweep of level 0: FP64 b
    mg_smooth_t<K, float, double, float>(st, g, s, c, x, b, dw, out, bc, ks);
}
void launch_mg_smooth(hipStream_t st, int kind, bool f32, bool b32, bool o32, const Grid& g, const Scal& s,
                      const MgCoef& c, const void* x, const void* b, const void* dw, void* out, float* bc,
                      const KState* ks) {
  switch (kind) {
    case MG_FIRST: mg_smooth_k<MG_FIRST>(st, f32, b32, o32, g, s, c, x, b, dw, out, bc, ks); break;
    case MG_SWEEP0: mg_smooth_k<MG_SWEEP0>(st, f32, b32, o32, g, s, c, x, b, dw, out, bc, ks); break;
    case MG_RES: mg_smooth_k<MG_RES>(st, f32, b32, o32, g, s, c, x, b, dw, out, bc, ks); break;
    default: mg_smooth_k<MG_SWEEP>(st, f32, b32, o32, g, s, c, x, b, dw, out, bc, ks); break;
  }
}
void launch_mg_restrict(hipStream_t st, bool r32, bool b32, const Grid& gf, const Grid& gc, const void* r,
                        void* bc, const KState* ks) {
  const dim3 grid(nblk(gc.nuv(), kBlock), 1, bat_n());
  if (r32 && b32)
    hipLaunchKernelGGL((k_mg_restrict<float, float>), grid, dim3(kBlock), 0, st, gf, gc, (const float*)r,
                       (float*)bc, ks, bat_bs());
  else if (r32)  // into an FP64 coarsest level
    hipLaunchKernelGGL((k_mg_restrict<float, double>), grid, dim3(kBlock), 0, st, gf, gc, (const float*)r,
                       (double*)bc, ks, bat_bs());
  else
    hipLaunchKernelGGL((k_mg_restrict<double, double>), grid, dim3(kBlock), 0, st, gf, gc, (const double*)r,
                       (double*)bc, ks, bat_bs());
}
void launch_mg_prolong_add(hipStream_t st, bool c32, bool f32, const Grid& gc, const Grid& gf, const void* xc,
                           void* x, const KState* ks) {
  const dim3 grid(nblk(gf.nuv(), kBlock), 1, bat_n());
  if (c32 && f32)
    hipLaunchKernelGGL((k_mg_prolong_add<float, float>), grid, dim3(kBlock), 0, st, gc, gf, (const float*)xc,
                       (float*)x, ks, bat_bs());
  else if (f32)  // from an FP64 coarsest level
    hipLaunchKernelGGL((k_mg_prolong_add<double, float>), grid, dim3(kBlock), 0, st, gc, gf, (const double*)xc,
                       (float*)x, ks, bat_bs());
  else
    hipLaunchKernelGGL((k_mg_prolong_add<double, double>), grid, dim3(kBlock), 0, st, gc, gf, (const double*)xc,
                       (double*)x, ks, bat_bs());
}

static dim3 rb_grid(const Grid& g) {  // one workgroup per tile (tile_id)
  return dim3((unsigned)(((g.nx + kRbTX - 1) / kRbTX) * ((g.ny + kRbTY - 1) / kRbTY)), 1, bat_n());
}
void launch_phmu_rb_sweep32(hipStream_t st, const Grid& g, const Scal& s, const PhmuCoef32& c32, const float* b,
                            const float* xi, float* xo, const KState* ks) {
  hipLaunchKernelGGL((k_phmu_rb_sweep<false, kRbNS, float, float, float, float>), rb_grid(g), dim3(kRbThreads), 0, st,
                     g, s, c32.bu, c32.minner, c32.m0, b, nullptr, xi, xo, ks, bat_bs(), c32.d11, c32.d22, c32.q);
}
// x = the red-black block Gauss-Seidel preconditioner applied to b:
// sixteen full sweeps (red cells, then black), from x = 0, kRbNS per launch,
// alternating between work and x so that the last launch writes x.
// FP64: work holds 2 ncell doubles.  FP32 storage (c32 = FP32 bu, M_inner,
// M_0): work holds 6 ncell floats, two FP32 iterates and the FP32 b the
// first launch writes.
constexpr int kRbgsLaunches = kRbgsMiddle + 2;
static_assert(kRbgsLaunches % 2 == 0, "the last launch must write x");
static_assert(kRbgsLaunches * kRbNS == 16, "sixteen sweeps per application");
void launch_phmu_rbgs(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const PhmuCoef32* c32,
                      const double* b, double* x, void* work, const KState* ks, const KTimer* tm) {
  const dim3 grid = rb_grid(g);
#define RB(F, TC, TB, TI, TO, CO, B, BC, XI, XO)                                                                 \
  hipLaunchKernelGGL((k_phmu_rb_sweep<F, kRbNS, TC, TB, TI, TO>), grid, dim3(kRbThreads), 0, st, g, s, CO->bu,         \
                     CO->minner, CO->m0, (const TB*)(B), (float*)(BC), (const TI*)(XI), (TO*)(XO), ks, bat_bs(), \
                     fd11, fd22, fq)
  const float* fd11 = c32 ? c32->d11 : nullptr;  // the FP32 sweeps' block factors
  const float* fd22 = c32 ? c32->d22 : nullptr;
  const float* fq = c32 ? c32->q : nullptr;
  // FP32 storage: two iterates and the FP32 b
  const int64_t n2 = 2 * g.ncell();
  float* wa = (float*)work;
  float* wb = wa + n2;
  float* bf = wb + n2;
  auto mid = [&](const float* xi, float* xo) { launch_phmu_rb_sweep32(st, g, s, *c32, bf, xi, xo, ks); };
  if (!c32) {
    double* w = (double*)work;
    const Coef* co = &c;
    RB(true, double, double, double, double, co, b, nullptr, nullptr, w);
    for (int k = 1; k < kRbgsLaunches; ++k) {
      if (k & 1)
        RB(false, double, double, double, double, co, b, nullptr, w, x);
      else
        RB(false, double, double, double, double, co, b, nullptr, x, w);
    }
  } else {
    RB(true, float, double, float, float, c32, b, bf, nullptr, wa);
    // the middle launches, timed as one group while tm has event pairs left:
    // markers on the stream, so each launch runs as it does untimed
    const int tk = (tm && *tm->n < tm->cap) ? (*tm->n)++ : -1;
    if (tk >= 0) (void)hipEventRecord(tm->ev[2 * tk], st);
    for (int k = 1; k + 1 < kRbgsLaunches; ++k) {
      if (k & 1)
        mid(wa, wb);
      else
        mid(wb, wa);
    }
    if (tk >= 0) (void)hipEventRecord(tm->ev[2 * tk + 1], st);
    RB(false, float, float, float, double, c32, bf, nullptr, wa, x);  // launch kRbgsLaunches - 2 (even) wrote wa
  }
#undef RB
}

void launch_interpolate_coef(hipStream_t st, const Grid& g, double dr, double dz, double I, double lrho,
                             double leta, const double* pts, int np, double radius, int nswm,
                             double swm_radius, const double* bdr, double* fpd, double* rho, double* eta) {
  hipLaunchKernelGGL(k_interpolate_coef, dim3(nblk(g.nc(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, dr, dz,
                     I, lrho, leta, pts, np, radius, nswm, swm_radius, bdr, fpd, rho, eta, bat_bs());
}

int body_chunks(const Grid& g, int nbody, int64_t cap) {
  int64_t c = std::min<int64_t>({(int64_t)g.ny, 256, cap / (3 * (int64_t)std::max(nbody, 1))});
  return (int)std::max<int64_t>(c, 1);
}

void launch_body_sums(hipStream_t st, const Grid& g, double dr, double dz, double I, double lx, double ly,
                      const double* bodies, int nbody, int chunks, const double* u, const double* v, double* part,
                      double* fin) {
  if (nbody <= 0) return;
  if (u)
    hipLaunchKernelGGL(k_body_sums<true>, dim3(chunks, nbody), dim3(kBlock), 0, st, g, dr, dz, I, lx, ly, bodies, u,
                       v, part);
  else
    hipLaunchKernelGGL(k_body_sums<false>, dim3(chunks, nbody), dim3(kBlock), 0, st, g, dr, dz, I, lx, ly, bodies,
                       u, v, part);
  hipLaunchKernelGGL(k_body_finish, dim3(nblk(nbody, kBlock)), dim3(kBlock), 0, st, nbody, chunks, dr, dz, part,
                     fin);
}

void launch_body_spread(hipStream_t st, const Grid& g, double dr, double dz, double I, double lx, double ly,
                        const double* bodies, int nbody, const double* fin, double* pf1, double* pf2) {
  const dim3 grid((unsigned)((g.nx + 2 + kSpTX - 1) / kSpTX), (unsigned)((g.ny + 2 + kSpTY - 1) / kSpTY));
  hipLaunchKernelGGL(k_body_spread, grid, dim3(kBlock), 0, st, g, dr, dz, I, lx, ly, bodies, nbody, fin, pf1, pf2);
}

void launch_phmu_coef32(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const PhmuCoef32& c32) {
  hipLaunchKernelGGL(k_phmu_coef32, dim3(nblk(g.nc(), kBlock), 1, bat_n()), dim3(kBlock), 0, st, g, s, c.bu,
                     c.minner, c.m0, c32, bat_bs());
}

}  // namespace drd
