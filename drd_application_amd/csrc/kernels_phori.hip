// kernels_phori.hip -- PhSolve_ori (TP/PhSolve_ori.f90:1-416, the same in
// MF/) and the CH re-initialisation of solid domains around it
// (MF/finit.f90:132-188 = TP/reinint_by_PhSolve.f90:1-91).
//
// PhSolve_ori is an implicit CH step for phi with unit mobility:
//   (I - dt M lap (s/eps - eps lap + walls)) phi = phi - dt u.grad(phi)
//                                                  - dt M lap(f(phi) + walls)
// with laplace_nn (the 5-point Laplacian whose ghosts copy their
// neighbour), wall terms from alpha_s and the contact angles, and the DCT
// preconditioner of PhSolve.  Every expression keeps the reference's order,
// so the operator equals the oracle (and the reference) bit for bit; the
// right-hand side and the finish evaluate cos on the device, so they agree
// to rounding.  Neumann sides: the re-initialisation runs with lr_bd_type 'n'.
#include "kernels.hpp"

namespace drd {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;  // pi of TP/global.f90:39 (rounded to double)

// a(i, k) of a ghost-copied field: neighbours beyond the interior clamp
__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// laplace_nn TP/PhSolve_ori.f90:463-487 at (i, k) of a value function F(i, k)
// over the interior (ghosts = copies of their neighbour)
template <class F>
__device__ __forceinline__ double lap_nn(const Grid& g, const Scal& s, int i, int k, F f) {
  const double a = f(i, k);
  return (f(clampi(i + 1, g.nx), k) + f(clampi(i - 1, g.nx), k) - 2 * a) / s.dr2 +
         (f(i, clampi(k + 1, g.ny)) + f(i, clampi(k - 1, g.ny)) - 2 * a) / s.dz2;
}

// pnb(i) / pnt(i) of :48-118 (wall type 'n'): (cf cos(tmp pi/2) + alpha_s tmp) / tmp1
__device__ __forceinline__ double wall_pn(const Scal& s, double angle, double tmp, double alpha_s) {
  const double cf = sqrt(2.0) * cos(angle * kPi / 180.0) / 3.0 * kPi / 2.;
  const double pn = cf * cos(tmp * kPi / 2) + alpha_s * tmp;
  const double tmp1 = s.I / s.dz + 0.5 * alpha_s;
  return pn / tmp1;
}

}  // namespace

// rhs = f2 - dt M laplace_nn(f1), x0 = phi (:16-139); pn[0, nx) = pnb,
// pn[nx, 2nx) = pnt for the finish.  One thread per interior cell.
__global__ __launch_bounds__(kBlock) void k_phori_rhs(Grid g, Scal s, const double* __restrict__ ph,
                                                      const double* __restrict__ u, const double* __restrict__ v,
                                                      const double* __restrict__ angB, const double* __restrict__ angT,
                                                      double alpha_s, double* __restrict__ rhs, double* __restrict__ x0,
                                                      double* __restrict__ pn) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.ncell()) return;
  const int nx = g.nx, ny = g.ny;
  const int k = g.by_nx.div(q), i = (int)(q - (int64_t)k * nx);
  auto P = [&](int a, int b) { return ph[g.ci(a, b)]; };
  auto f1 = [&](int a, int b) {
    const double p = P(a, b);
    double f = ((1. + s.s) * p - p * p * p) / s.I;
    if (b == 0) f = f + wall_pn(s, angB[a], (P(a, -1) + P(a, 0)) / 2.0, alpha_s) * s.I / s.dz2;
    if (b == ny - 1) f = f + wall_pn(s, angT[a], (P(a, ny - 1) + P(a, ny)) / 2.0, alpha_s) * s.I / s.dz2;
    return f;
  };
  double f2 = P(i, k) - s.dt * (u[g.ui(i + 1, k)] + u[g.ui(i, k)]) / 2.0 * (P(i + 1, k) - P(i - 1, k)) / 2.0 / s.dr -
              s.dt * (v[g.vi(i, k + 1)] + v[g.vi(i, k)]) / 2.0 * (P(i, k + 1) - P(i, k - 1)) / 2.0 / s.dz;
  f2 = f2 - s.dt * s.xld * lap_nn(g, s, i, k, f1);
  rhs[q] = f2;
  x0[q] = P(i, k);
  if (k == 0) pn[i] = wall_pn(s, angB[i], (P(i, -1) + P(i, 0)) / 2.0, alpha_s);
  if (k == ny - 1) pn[nx + i] = wall_pn(s, angT[i], (P(i, ny - 1) + P(i, ny)) / 2.0, alpha_s);
}

// MATRIX_VECTOR_CG_ori :348-397, stage 1: t = -eps laplace_nn(x) + walls + s x / eps
__global__ __launch_bounds__(kBlock) void k_phori_mv1(Grid g, Scal s, double cbot, double ctop,
                                                      const double* __restrict__ x, double* __restrict__ t,
                                                      const KState* st) {
  if (st && (st->done || st->early)) return;
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.ncell()) return;
  const int nx = g.nx, ny = g.ny;
  const int k = g.by_nx.div(q), i = (int)(q - (int64_t)k * nx);
  auto X = [&](int a, int b) { return x[a + (int64_t)nx * b]; };
  const double f = X(i, k);
  double ff1 = -s.I * lap_nn(g, s, i, k, X);
  if (k == 0) ff1 = ff1 + s.I / s.dz2 * (f - cbot * f);  // the ghost f1(i,-1) = f1(i,0)
  if (k == ny - 1) ff1 = ff1 + s.I / s.dz2 * (f - ctop * f);
  t[q] = ff1 + f * s.s / s.I;
}

// stage 2: y = alpha (x - dt M laplace_nn(t)) + beta y, with the BiCGSTAB
// fusions of the other operators
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_phori_mv2(Grid g, Scal s, const double* __restrict__ x,
                                                      const double* __restrict__ t, double* __restrict__ y,
                                                      double alpha, double beta, double* __restrict__ a0,
                                                      const double* __restrict__ b, double* part, KState* st) {
  if (MODE != MV_PLAIN && (st->done || st->early)) return;
  const int64_t n = g.ncell();
  const int nx = g.nx;
  double tm[2][kSlots] = {};
  const bool xnz = (MODE == MV_INIT) ? (st->xnz != 0) : true;
#pragma unroll
  for (int jj = 0; jj < kSlots; ++jj) {
    const int64_t q = tile_elem(n, 0, jj);
    if (q < 0) break;
    const int k = g.by_nx.div(q), i = (int)(q - (int64_t)k * nx);
    double z = 0.0;
    if (MODE != MV_INIT || xnz)
      z = x[q] + -s.dt * s.xld * lap_nn(g, s, i, k, [&](int a, int bb) { return t[a + (int64_t)nx * bb]; });
    if (MODE == MV_PLAIN) {
      y[q] = beta == 0.0 ? alpha * z : alpha * z + beta * y[q];
    } else if (MODE == MV_INIT) {
      double r = b[q];
      if (xnz) r = -1.0 * z + 1.0 * r;
      y[q] = r;
      a0[q] = r;
      tm[0][jj] = r * r;
    } else {
      y[q] = z;
      if (MODE == MV_DOT1) {
        tm[0][jj] = a0[q] * z;
      } else {
        tm[0][jj] = z * a0[q];
        tm[1][jj] = z * z;
      }
    }
  }
  if (MODE == MV_PLAIN) return;
  if (MODE == MV_DOT2) {
    tile_publish<1, 2>(tm, part);
  } else {
    double t1[1][kSlots];
#pragma unroll
    for (int jj = 0; jj < kSlots; ++jj) t1[0][jj] = tm[0][jj];
    tile_publish<1, 1>(t1, part);
  }
}

// :160-233: phi_after = x with the wall ghosts x cbot + pnb (top: ctop, pnt)
// and the side ghosts copied (corners the copies of the wall ghosts);
// mu_after = -eps lap(phi_after) - phi/eps + phi^3/eps on the interior, its
// ghosts copied from the nearest interior cell.  One thread per cell.
__global__ __launch_bounds__(kBlock) void k_phori_finish(Grid g, Scal s, double cbot, double ctop,
                                                         const double* __restrict__ x, const double* __restrict__ pn,
                                                         double* __restrict__ ph, double* __restrict__ mu) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nc()) return;
  const int nx = g.nx, ny = g.ny;
  const int k = (int)(q / (nx + 2)) - 1, i = (int)(q - (int64_t)(k + 1) * (nx + 2)) - 1;
  auto PA = [&](int a, int b) {  // phi_after at any cell of the ghosted field
    const int ac = clampi(a, nx);
    if (b < 0) return x[ac] * cbot + pn[ac];
    if (b >= ny) return x[ac + (int64_t)nx * (ny - 1)] * ctop + pn[nx + ac];
    return x[ac + (int64_t)nx * b];
  };
  ph[q] = PA(i, k);
  const int ic = clampi(i, nx), kc = clampi(k, ny);
  const double a = PA(ic, kc);
  mu[q] = -s.I * ((PA(ic + 1, kc) + PA(ic - 1, kc) - 2 * a) / s.dr2 + (PA(ic, kc + 1) + PA(ic, kc - 1) - 2 * a) / s.dz2) -
          a / s.I + a * a * a / s.I;
}

// re-initialisation, start: phi = bdr_ph with its ghosts copied from the
// nearest interior cell, M_0 = M_inner = 1 (MF/finit.f90:132-157)
__global__ __launch_bounds__(kBlock) void k_reinit_start(Grid g, const double* __restrict__ bdr,
                                                         double* __restrict__ ph, double* __restrict__ m0,
                                                         double* __restrict__ minner) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nc()) return;
  const int nx = g.nx;
  const int k = (int)(q / (nx + 2)) - 1, i = (int)(q - (int64_t)(k + 1) * (nx + 2)) - 1;
  ph[q] = bdr[g.ci(clampi(i, nx), clampi(k, g.ny))];
  m0[q] = 1.0;
  minner[q] = 1.0;
}

// end: bdr_ph = (phi + 1) / 2, the interior clamped to [0, 1], and the
// penalty mask bdr_ph > 0.5 (MF/finit.f90:176-200, TP/reinint_by_PhSolve.f90:63-89)
__global__ __launch_bounds__(kBlock) void k_reinit_end(Grid g, const double* __restrict__ ph, double* __restrict__ bdr,
                                                       double* __restrict__ pen) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nc()) return;
  const int nx = g.nx, ny = g.ny;
  const int k = (int)(q / (nx + 2)) - 1, i = (int)(q - (int64_t)(k + 1) * (nx + 2)) - 1;
  double b = (ph[q] + 1.0) / 2.0;
  if (i >= 0 && i < nx && k >= 0 && k < ny) {
    if (b >= 1.0)
      b = 1.0;
    else if (b <= 0.0)
      b = 0.0;
  }
  bdr[q] = b;
  pen[q] = b > 0.5 ? 1.0 : 0.0;
}

static inline unsigned nbo(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void launch_phori_rhs(hipStream_t st, const Grid& g, const Scal& s, const double* ph, const double* u,
                      const double* v, const double* angB, const double* angT, double alpha_s, double* rhs,
                      double* x0, double* pn) {
  hipLaunchKernelGGL(k_phori_rhs, dim3(nbo(g.ncell())), dim3(kBlock), 0, st, g, s, ph, u, v, angB, angT, alpha_s,
                     rhs, x0, pn);
}

void launch_phori_matvec(hipStream_t st, int mode, const Grid& g, const Scal& s, double cbot, double ctop,
                         const double* x, double* t, double* y, double alpha, double beta, double* a0,
                         const double* b, double* part, KState* ks) {
  const unsigned n1 = nbo(g.ncell());
  hipLaunchKernelGGL(k_phori_mv1, dim3(n1), dim3(kBlock), 0, st, g, s, cbot, ctop, x, t,
                     mode == MV_PLAIN ? nullptr : ks);
  const dim3 grid(tiles_of(g.ncell()));
#define PO(M) hipLaunchKernelGGL((k_phori_mv2<M>), grid, dim3(kBlock), 0, st, g, s, x, t, y, alpha, beta, a0, b, part, ks)
  switch (mode) {
    case MV_PLAIN: PO(MV_PLAIN); break;
    case MV_DOT1: PO(MV_DOT1); break;
    case MV_DOT2: PO(MV_DOT2); break;
    case MV_INIT: PO(MV_INIT); break;
  }
#undef PO
  launch_mv_finalize(st, mode, part, grid.x, 1, ks);
}

void launch_phori_finish(hipStream_t st, const Grid& g, const Scal& s, double cbot, double ctop, const double* x,
                         const double* pn, double* ph, double* mu) {
  hipLaunchKernelGGL(k_phori_finish, dim3(nbo(g.nc())), dim3(kBlock), 0, st, g, s, cbot, ctop, x, pn, ph, mu);
}

void launch_reinit_start(hipStream_t st, const Grid& g, const double* bdr, double* ph, double* m0, double* minner) {
  hipLaunchKernelGGL(k_reinit_start, dim3(nbo(g.nc())), dim3(kBlock), 0, st, g, bdr, ph, m0, minner);
}

void launch_reinit_end(hipStream_t st, const Grid& g, const double* ph, double* bdr, double* pen) {
  hipLaunchKernelGGL(k_reinit_end, dim3(nbo(g.nc())), dim3(kBlock), 0, st, g, ph, bdr, pen);
}

}  // namespace drd
