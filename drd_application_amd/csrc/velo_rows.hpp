// velo_rows.hpp -- the MAC-grid velocity operator evaluated at one unknown,
// shared by the matvec, the multigrid smoothers (stored vectors) and the
// LDS-tiled multigrid kernels (vectors and coefficients staged in LDS).
//
// The arithmetic is written once (velo_u_at / velo_v_at) over two views:
// the vector view X returns u(i,k) and v(i,k) with the ghost rules applied,
// the coefficient view C returns rho, eta, chi at a cell index and maps
// (i,k) to that index.  Every view computes the same values in the same
// order, so a kernel's result does not depend on where its operands sit.
#pragma once

#include "kernels.hpp"

namespace drd {

// Ghost rules applied on the fly to a packed velocity vector:
// btoA_uv TP/matrix_tool.f90:32-75 plus the overrides of
// MatrixVector_uv_2_CG TP/VeloSolve_2_combined.f90:554-583.
// The vector's values come from a loader L (value of packed element i, as
// double): a stored FP64 or FP32 vector (LdVec), or values formed on the fly.
template <class T>
struct LdVec {
  using R = T;  // the arithmetic type of values read through it
  const T* __restrict__ x;
  __device__ __forceinline__ T operator()(uint32_t i) const { return ld32(x, i); }
};
// PX: periodic in x, a u row of nx unknowns and every side ghost the
// periodic copy (u(-1) = u(nx-1), u(nx) = u(0), u(nx+1) = u(1); v likewise).
template <class L, bool PX = false>
struct VeloAcc {
  using R = typename L::R;
  L ld;
  int nx, ny, lr_d;
  uint32_t mn;
  __device__ __forceinline__ R uraw(int i, int k) const { return ld((uint32_t)(i + (PX ? nx : nx + 1) * k)); }
  // f1(i,k) for -1<=i<=nx+1, -1<=k<=ny (corners unused by the stencils):
  // wall ghosts are odd mirrors of the first/last row, side ghosts even or
  // odd mirrors of column 1 / nx-1.  Negation is exact, so the mirrorings
  // commute; both are applied branch-free.
  __device__ __forceinline__ R U(int i, int k) const {
    const bool kw = (k < 0) | (k >= ny);
    const int kk = k < 0 ? 0 : (k >= ny ? ny - 1 : k);
    if (PX) {
      const int ii = i < 0 ? i + nx : (i >= nx ? i - nx : i);
      const R v = ld((uint32_t)(ii + nx * kk));
      return kw ? -v : v;
    }
    const bool iw = ((i < 0) | (i > nx)) & (lr_d != 0);
    const int ii = i < 0 ? 1 : (i > nx ? nx - 1 : i);
    const R v = ld((uint32_t)(ii + (nx + 1) * kk));
    const R w = iw ? -v : v;
    return kw ? -w : w;
  }
  // f2(i,k) for -1<=i<=nx, 0<=k<=ny (0 on the walls k = 0, ny)
  __device__ __forceinline__ R V(int i, int k) const {
    const int ic = PX ? (i < 0 ? i + nx : (i >= nx ? i - nx : i)) : (i < 0 ? 0 : (i >= nx ? nx - 1 : i));
    const int kc = k < 1 ? 1 : (k > ny - 1 ? ny - 1 : k);
    const R v = ld(mn + (uint32_t)(ic + nx * (kc - 1)));
    return (k <= 0 || k >= ny) ? R(0.0) : v;
  }
};
template <class T, bool PX = false>
using VeloInT = VeloAcc<LdVec<T>, PX>;
using VeloIn = VeloInT<double>;

// The same ghost rules (side walls, not periodic) over u and v values held
// in LDS as two row-major arrays with row stride ld whose element 0 is the
// unknown (i0, k0).  Every position the rules map to must be held.
template <class T>
struct VeloLds {
  using R = T;
  const T* u;
  const T* v;
  int nx, ny, lr_d, i0, k0, ld;
  __device__ __forceinline__ T U(int i, int k) const {
    const bool kw = (k < 0) | (k >= ny);
    const int kk = k < 0 ? 0 : (k >= ny ? ny - 1 : k);
    const bool iw = ((i < 0) | (i > nx)) & (lr_d != 0);
    const int ii = i < 0 ? 1 : (i > nx ? nx - 1 : i);
    const T x = u[(ii - i0) + ld * (kk - k0)];
    const T w = iw ? -x : x;
    return kw ? -w : w;
  }
  __device__ __forceinline__ T V(int i, int k) const {
    const int ic = i < 0 ? 0 : (i >= nx ? nx - 1 : i);
    const int kc = k < 1 ? 1 : (k > ny - 1 ? ny - 1 : k);
    const T x = v[(ic - i0) + ld * (kc - k0)];
    return (k <= 0 || k >= ny) ? T(0.0) : x;
  }
};

// The same for blocks whose every read position is an unknown of the grid
// away from the walls and side columns, where the ghost rules map each
// position to itself: plain LDS reads at fixed offsets from the row's base.
template <class T>
struct VeloLdsIn {
  using R = T;
  const T* u;
  const T* v;
  int i0, k0, ld;
  __device__ __forceinline__ T U(int i, int k) const { return u[(i - i0) + ld * (k - k0)]; }
  __device__ __forceinline__ T V(int i, int k) const { return v[(i - i0) + ld * (k - k0)]; }
};

// Cell coefficients rho, eta, chi of the stored fields (layout (-1:nx,-1:ny),
// row stride nx + 2).
template <class T>
struct CoefG {
  const T* __restrict__ rho;
  const T* __restrict__ eta;
  const T* __restrict__ chi;
  uint32_t ld;
  __device__ __forceinline__ uint32_t at(int i, int k) const { return (uint32_t)((i + 1) + (int)ld * (k + 1)); }
  __device__ __forceinline__ T r(uint32_t c) const { return ld32(rho, c); }
  __device__ __forceinline__ T e(uint32_t c) const { return ld32(eta, c); }
  __device__ __forceinline__ T x(uint32_t c) const { return ld32(chi, c); }
};
// The same coefficients staged in LDS: row stride ld, element 0 the cell
// (i0, k0).
template <class T>
struct CoefL {
  const T* rho;
  const T* eta;
  const T* chi;
  int i0, k0;
  uint32_t ld;
  __device__ __forceinline__ uint32_t at(int i, int k) const { return (uint32_t)((i - i0) + (int)ld * (k - k0)); }
  __device__ __forceinline__ T r(uint32_t c) const { return rho[c]; }
  __device__ __forceinline__ T e(uint32_t c) const { return eta[c]; }
  __device__ __forceinline__ T x(uint32_t c) const { return chi[c]; }
};

// y = A x for one unknown of the velocity operator.
// MatrixVector_uv_CG  TP/VeloSolve_combined.f90:606-670 (bdf 1)
// MatrixVector_uv_2_CG TP/VeloSolve_2_combined.f90:607-670 (bdf 2)
// Computed in the coefficients' type T: FP64 is the reference's arithmetic;
// FP32 (the FP32 multigrid levels) multiplies by the reciprocals of the
// spacings where the reference divides.  The vector's loader reads T.
template <class T, class D>
__device__ __forceinline__ T vdiv(T a, D d, D inv) {
  return sizeof(T) == sizeof(float) ? a * T(inv) : a / T(d);
}
// The scalars a stencil reads, converted once to the stencil's type T (the
// same values T(s.re), ... that the stencil would form from Scal; FP32
// kernels convert them at their start instead of at every unknown).
template <class T>
struct ScalT {
  T re, penalty_k, dt, dr, dz, dr2, dz2, idr2, idz2, idrdz;
  __device__ __forceinline__ static ScalT make(const Scal& s) {
    return ScalT{T(s.re), T(s.penalty_k), T(s.dt), T(s.dr), T(s.dz), T(s.dr2), T(s.dz2), T(s.idr2), T(s.idz2), T(s.idrdz)};
  }
};
// u row at face (i, k)
template <int BDF, class T, class XA, class CV, class S>
__device__ __forceinline__ T velo_u_at(const S& s, const CV& C, const XA& X, int i, int k) {
  T r;
  const uint32_t c = C.at(i, k), cw = c - 1, ld = C.ld;
  const T f = X.U(i, k);
  if (BDF == 2)
    r = f * T(1.5) * T(s.re) * (C.r(c) + C.r(cw)) / T(2.0) + f * T(s.penalty_k) * (C.x(c) + C.x(cw)) / T(2.0) * T(s.dt);
  else
    r = f * T(s.re) * (C.r(c) + C.r(cw)) / T(2.0);
  const T e00 = C.e(c), e10 = C.e(cw);
  T cr = e00, cl = e10;
  r = r - vdiv(T(2.0) * T(s.dt) * (cr * X.U(i + 1, k) + cl * X.U(i - 1, k) - (cl + cr) * f), s.dr2, s.idr2);
  cr = (C.e(c + ld) + C.e(cw + ld) + e00 + e10) / T(4.0);
  cl = (e00 + e10 + C.e(c - ld) + C.e(cw - ld)) / T(4.0);
  r = r - vdiv(T(s.dt) * (cr * X.U(i, k + 1) + cl * X.U(i, k - 1) - (cl + cr) * f), s.dz2, s.idz2);
  const T cross = T(s.dt) * (cr * (X.V(i, k + 1) - X.V(i - 1, k + 1)) - cl * (X.V(i, k) - X.V(i - 1, k)));
  r = r - (sizeof(T) == sizeof(float) ? cross * T(s.idrdz) : cross / T(s.dr) / T(s.dz));
  return r;
}
// v row at face (i, k)
template <int BDF, class T, class XA, class CV, class S>
__device__ __forceinline__ T velo_v_at(const S& s, const CV& C, const XA& X, int i, int k) {
  T r;
  const uint32_t c = C.at(i, k), ld = C.ld, cs = c - ld;
  const T f = X.V(i, k);
  if (BDF == 2)
    r = f * T(1.5) * T(s.re) * (C.r(c) + C.r(cs)) / T(2.0) + f * T(s.penalty_k) * (C.x(c) + C.x(cs)) / T(2.0) * T(s.dt);
  else
    r = f * T(s.re) * (C.r(c) + C.r(cs)) / T(2.0);
  const T e00 = C.e(c), e01 = C.e(cs);
  T cr = e00, cl = e01;
  r = r - vdiv(T(2.0) * T(s.dt) * (cr * X.V(i, k + 1) + cl * X.V(i, k - 1) - (cl + cr) * f), s.dz2, s.idz2);
  cr = (C.e(c + 1) + e00 + C.e(cs + 1) + e01) / T(4.0);
  cl = (e00 + C.e(c - 1) + e01 + C.e(cs - 1)) / T(4.0);
  r = r - vdiv(T(s.dt) * (cr * X.V(i + 1, k) + cl * X.V(i - 1, k) - (cl + cr) * f), s.dr2, s.idr2);
  const T cross = T(s.dt) * (cr * (X.U(i + 1, k) - X.U(i + 1, k - 1)) - cl * (X.U(i, k) - X.U(i, k - 1)));
  r = r - (sizeof(T) == sizeof(float) ? cross * T(s.idrdz) : cross / T(s.dr) / T(s.dz));
  return r;
}

// The same at packed unknown q of a stored vector and the stored fields.
template <int BDF, class T, class XA>
__device__ __forceinline__ T velo_u_row(const Grid g, const Scal& s, const T* __restrict__ rho,
                                        const T* __restrict__ eta, const T* __restrict__ chi, const XA& X,
                                        uint32_t q) {
  const int k = g.by_nxu.div(q);
  const int i = (int)(q - (uint32_t)k * (uint32_t)g.nxu());
  const CoefG<T> C{rho, eta, chi, (uint32_t)(g.nx + 2)};
  return velo_u_at<BDF, T>(s, C, X, i, k);
}

template <int BDF, class T, class XA>
__device__ __forceinline__ T velo_v_row(const Grid g, const Scal& s, const T* __restrict__ rho,
                                        const T* __restrict__ eta, const T* __restrict__ chi, const XA& X,
                                        uint32_t q) {
  const uint32_t t = q - X.mn;
  const int k = g.by_nx.div(t) + 1;
  const int i = (int)(t - (uint32_t)(k - 1) * (uint32_t)g.nx);
  const CoefG<T> C{rho, eta, chi, (uint32_t)(g.nx + 2)};
  return velo_v_at<BDF, T>(s, C, X, i, k);
}

// ---- the FP32 multigrid levels' form of the BDF2 rows ---------------------
// The same operator as velo_u_at / velo_v_at<2>, with its terms regrouped:
//   (A x)_u = D f - [kx (E_e U_e + E_w U_w) + ky (Q_n U_n + Q_s U_s)
//                    + kxy (Q_n (V_ne - V_nw) - Q_s (V_se - V_sw))]
//   D_u = 0.75 re (rho_e + rho_w) + 0.5 pk dt (chi_e + chi_w)
//         + kx (E_e + E_w) + ky (Q_n + Q_s)
// (v alike with x and y exchanged), E the cell viscosity, Q the corner
// viscosity (the mean of the four cells around the corner, as velo_u_at
// forms cr and cl), kx = 2 dt / dr^2, ky = dt / dz^2, kxy = dt / (dr dz).
// D and Q are formed once per position and cell (the tiled legs keep them
// for the whole leg), so an evaluation is 13 fused multiply-adds and
// subtractions instead of about 45 operations.  FP32 only: the ghost rules
// are the vector view's, as in velo_u_at; the rounding is this form's own.
template <class T>
struct PcK {
  T cr, cx, kx, ky, kxy, kxv, kyv;  // 0.75 re, 0.5 pk dt, 2 dt / dr^2, dt / dz^2, dt / (dr dz), dt / dr^2, 2 dt / dz^2
};
template <class T, class S>
__device__ __forceinline__ PcK<T> pc_k(const S& s) {
  const T dt = T(s.dt);
  return PcK<T>{T(1.5) * T(s.re) * T(0.5), T(s.penalty_k) * dt * T(0.5), T(2.0) * dt * T(s.idr2), dt * T(s.idz2),
                dt * T(s.idrdz),           dt * T(s.idr2),               T(2.0) * dt * T(s.idz2)};
}
// the corner viscosity at the north-east corner of cell c of view C
template <class T, class CV>
__device__ __forceinline__ T pc_corner(const CV& C, uint32_t c) {
  return ((C.e(c) + C.e(c + 1)) + (C.e(c + C.ld) + C.e(c + C.ld + 1))) * T(0.25);
}
template <class T, class CV>
struct PcQFly {  // corners formed where they are read
  const CV& C;
  __device__ __forceinline__ T operator()(uint32_t c) const { return pc_corner<T>(C, c); }
};
template <class T>
struct PcQArr {  // corners formed beforehand, in C's index space
  const T* q;
  __device__ __forceinline__ T operator()(uint32_t c) const { return q[c]; }
};
template <class T, class CV, class QV>
__device__ __forceinline__ T pc_du(const PcK<T>& K, const CV& C, const QV& Q, int i, int k) {
  const uint32_t c = C.at(i, k), cw = c - 1;
  T d = (Q(cw) + Q(cw - C.ld)) * K.ky;
  d = __builtin_fmaf(C.e(c) + C.e(cw), K.kx, d);
  d = __builtin_fmaf(C.x(c) + C.x(cw), K.cx, d);
  return __builtin_fmaf(C.r(c) + C.r(cw), K.cr, d);
}
template <class T, class CV, class QV>
__device__ __forceinline__ T pc_dv(const PcK<T>& K, const CV& C, const QV& Q, int i, int k) {
  const uint32_t c = C.at(i, k), cs = c - C.ld;
  T d = (Q(cs) + Q(cs - 1)) * K.kxv;
  d = __builtin_fmaf(C.e(c) + C.e(cs), K.kyv, d);
  d = __builtin_fmaf(C.x(c) + C.x(cs), K.cx, d);
  return __builtin_fmaf(C.r(c) + C.r(cs), K.cr, d);
}
template <class T, class XA, class CV, class QV>
__device__ __forceinline__ T velo_u_pc(const PcK<T>& K, const CV& C, const QV& Q, const XA& X, int i, int k, T D) {
  const uint32_t c = C.at(i, k), cw = c - 1;
  const T qn = Q(cw), qs = Q(cw - C.ld);
  T a = C.e(c) * X.U(i + 1, k);
  a = __builtin_fmaf(C.e(cw), X.U(i - 1, k), a);
  T b = qn * X.U(i, k + 1);
  b = __builtin_fmaf(qs, X.U(i, k - 1), b);
  T x = qn * (X.V(i, k + 1) - X.V(i - 1, k + 1));
  x = __builtin_fmaf(-qs, X.V(i, k) - X.V(i - 1, k), x);
  T t = K.kx * a;
  t = __builtin_fmaf(K.ky, b, t);
  t = __builtin_fmaf(K.kxy, x, t);
  return __builtin_fmaf(D, X.U(i, k), -t);
}
template <class T, class XA, class CV, class QV>
__device__ __forceinline__ T velo_v_pc(const PcK<T>& K, const CV& C, const QV& Q, const XA& X, int i, int k, T D) {
  const uint32_t c = C.at(i, k), cs = c - C.ld;
  const T qe = Q(cs), qw = Q(cs - 1);
  T a = C.e(c) * X.V(i, k + 1);
  a = __builtin_fmaf(C.e(cs), X.V(i, k - 1), a);
  T b = qe * X.V(i + 1, k);
  b = __builtin_fmaf(qw, X.V(i - 1, k), b);
  T x = qe * (X.U(i + 1, k) - X.U(i + 1, k - 1));
  x = __builtin_fmaf(-qw, X.U(i, k) - X.U(i, k - 1), x);
  T t = K.kyv * a;
  t = __builtin_fmaf(K.kxv, b, t);
  t = __builtin_fmaf(K.kxy, x, t);
  return __builtin_fmaf(D, X.V(i, k), -t);
}
// The same at packed unknown q of a stored vector and the stored fields,
// D and the corners formed on the spot.
template <class T, class XA>
__device__ __forceinline__ T velo_row_pc(const Grid g, const Scal& s, const T* __restrict__ rho,
                                         const T* __restrict__ eta, const T* __restrict__ chi, const XA& X,
                                         uint32_t q) {
  const CoefG<T> C{rho, eta, chi, (uint32_t)(g.nx + 2)};
  const PcQFly<T, CoefG<T>> Q{C};
  const PcK<T> K = pc_k<T>(s);
  if (q < X.mn) {
    const int k = g.by_nxu.div(q);
    const int i = (int)(q - (uint32_t)k * (uint32_t)g.nxu());
    return velo_u_pc(K, C, Q, X, i, k, pc_du(K, C, Q, i, k));
  }
  const uint32_t t = q - X.mn;
  const int k = g.by_nx.div(t) + 1;
  const int i = (int)(t - (uint32_t)(k - 1) * (uint32_t)g.nx);
  return velo_v_pc(K, C, Q, X, i, k, pc_dv(K, C, Q, i, k));
}

template <int BDF, class T, class XA>
__device__ __forceinline__ T velo_row(const Grid g, const Scal& s, const T* __restrict__ rho,
                                      const T* __restrict__ eta, const T* __restrict__ chi, const XA& X,
                                      uint32_t q) {
  return q < X.mn ? velo_u_row<BDF, T, XA>(g, s, rho, eta, chi, X, q)
                   : velo_v_row<BDF, T, XA>(g, s, rho, eta, chi, X, q);
}

// ---- multigrid transfers at one unknown -------------------------------------
// Level l+1 halves level l in both directions.  Prolongation P is staggered
// bilinear, restriction R = P^T / 4; both evaluate one set of 1-D weights in
// closed form (the derivation sits with k_mg_restrict in
// kernels_stencil.hip).  Loads are clamped into the grid and their weights
// zeroed instead.  The vector views return a level's value at u face (i, k)
// or v face (i, k) as double.
template <class T>
struct PackedUV {  // a stored packed vector of the level
  const T* __restrict__ x;
  int nx;
  int64_t mn;
  __device__ __forceinline__ double operator()(bool is_u, int i, int k) const {
    return is_u ? (double)x[(int64_t)i + (int64_t)(nx + 1) * k] : (double)x[mn + i + (int64_t)nx * (k - 1)];
  }
};
template <class T>
struct LdsUV {  // u and v arrays in LDS (row stride ld, element 0 at (i0, k0)), no ghost rules
  const T* u;
  const T* v;
  int i0, k0, ld;
  __device__ __forceinline__ double operator()(bool is_u, int i, int k) const {
    return (double)(is_u ? u : v)[(i - i0) + ld * (k - k0)];
  }
};

// (R r)(I, K) * 4 at coarse u face (I, K): fine faces 2I-1..2I+1
// (1/2, 1, 1/2), rows 2K-1..2K+2.
template <class RU>
__device__ __forceinline__ double mg_restrict_u(const Grid& gf, const Grid& gc, const RU& r, int I, int K) {
  const bool lo = K >= 1, hi = K + 1 <= gc.ny - 1;  // the neighbouring coarse rows exist
  const double wy[4] = {lo ? 0.25 : 0.0, lo ? 0.75 : 0.5, hi ? 0.75 : 0.5, hi ? 0.25 : 0.0};
  const int il = I > 0 ? 2 * I - 1 : 0, ir = 2 * I + 1 <= gf.nx ? 2 * I + 1 : gf.nx;
  const double wl = I > 0 ? 0.5 : 0.0, wr = 2 * I + 1 <= gf.nx ? 0.5 : 0.0;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    int k = 2 * K - 1 + m;
    k = k < 0 ? 0 : (k > gf.ny - 1 ? gf.ny - 1 : k);
    const double sx = wl * r(true, il, k) + r(true, 2 * I, k) + wr * r(true, ir, k);
    acc += wy[m] * sx;
  }
  return acc;
}
// (R r)(I, K) * 4 at coarse v face (I, K), K = 1..ny-1: fine inner faces
// 2K-1..2K+1 (1/2, 1, 1/2), cells 2I-1..2I+2.
template <class RV>
__device__ __forceinline__ double mg_restrict_v(const Grid& gf, const Grid& gc, const RV& r, int I, int K) {
  const bool left = I >= 1, right = I + 1 <= gc.nx - 1;  // the neighbouring coarse cells exist
  const double wx[4] = {left ? 0.25 : 0.0, left ? 0.75 : 1.0, right ? 0.75 : 1.0, right ? 0.25 : 0.0};
  const int i0 = left ? 2 * I - 1 : 0, i3 = right ? 2 * I + 2 : gf.nx - 1;
  const double wy[3] = {0.5, 1.0, 0.5};
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int k = 2 * K - 1 + m;
    const double sx = wx[0] * r(false, i0, k) + wx[1] * r(false, 2 * I, k) + wx[2] * r(false, 2 * I + 1, k) +
                      wx[3] * r(false, i3, k);
    acc += wy[m] * sx;
  }
  return acc;
}
// (P xc) at fine u face (i, k): coarse faces i/2 (and i/2 + 1), rows k/2
// and its neighbour.  mg_prolong_u_at gives the four coarse faces read, in
// the order mg_prolong_u_sum takes their values (callers issue the four
// loads together).
__device__ __forceinline__ void mg_prolong_u_at(const Grid& gc, int i, int k, int (&I)[4], int (&K)[4]) {
  const int I0 = i >> 1, I1 = (i & 1) ? I0 + 1 : I0;  // odd i: 1/2 each; even i: I0 alone
  const int K0 = k >> 1, K1 = (k & 1) ? K0 + 1 : K0 - 1;
  const bool in1 = K1 >= 0 && K1 <= gc.ny - 1;
  const int K1c = in1 ? K1 : K0;
  I[0] = I0, K[0] = K0;
  I[1] = I1, K[1] = K0;
  I[2] = I0, K[2] = K1c;
  I[3] = I1, K[3] = K1c;
}
__device__ __forceinline__ double mg_prolong_u_sum(const Grid& gc, int i, int k, const double (&x)[4]) {
  const double wx = (i & 1) ? 0.5 : 1.0;
  const int K0 = k >> 1, K1 = (k & 1) ? K0 + 1 : K0 - 1;
  const bool in1 = K1 >= 0 && K1 <= gc.ny - 1;
  const double w0 = in1 ? 0.75 : 0.5, w1 = in1 ? 0.25 : 0.0;
  const double v0 = x[0], u0 = x[1], v1 = x[2], u1 = x[3];
  const double s0 = (i & 1) ? wx * v0 + wx * u0 : v0, s1 = (i & 1) ? wx * v1 + wx * u1 : v1;
  return w0 * s0 + w1 * s1;
}
// (P xc) at fine v face (i, k), k = 1..ny-1: coarse inner faces k/2 (and
// k/2 + 1), cells i/2 and its neighbour; the four coarse faces read, then
// the sum of their values.
__device__ __forceinline__ void mg_prolong_v_at(const Grid& gc, int i, int k, int (&I)[4], int (&K)[4]) {
  // faces: even k -> K = k/2 (1); odd k -> k>>1 and k>>1 + 1 (1/2 each), wall faces (0, ny) dropped
  const int Ka = k >> 1, Kb = (k & 1) ? Ka + 1 : Ka;
  const bool ina = Ka >= 1 && Ka <= gc.ny - 1, inb = Kb >= 1 && Kb <= gc.ny - 1;
  const int Kac = ina ? Ka : 1, Kbc = inb ? Kb : 1;
  // cells: c = i/2 (3/4), its neighbour (1/4), clamped (the weights add where they meet)
  const int c = i >> 1;
  int o = (i & 1) ? c + 1 : c - 1;
  o = o < 0 ? 0 : (o > gc.nx - 1 ? gc.nx - 1 : o);
  I[0] = c, K[0] = Kac;
  I[1] = o, K[1] = Kac;
  I[2] = c, K[2] = Kbc;
  I[3] = o, K[3] = Kbc;
}
__device__ __forceinline__ double mg_prolong_v_sum(const Grid& gc, int i, int k, const double (&x)[4]) {
  const int Ka = k >> 1, Kb = (k & 1) ? Ka + 1 : Ka;
  const bool ina = Ka >= 1 && Ka <= gc.ny - 1, inb = Kb >= 1 && Kb <= gc.ny - 1;
  const double wa = ina ? ((k & 1) ? 0.5 : 1.0) : 0.0, wb = ((k & 1) && inb) ? 0.5 : 0.0;
  const double sa = 0.75 * x[0] + 0.25 * x[1];
  const double sb = 0.75 * x[2] + 0.25 * x[3];
  return wa * sa + wb * sb;
}
// packed index of coarse u face (I, K) / v face (I, K)
__device__ __forceinline__ int64_t mg_uidx(const Grid& g, int I, int K) { return (int64_t)I + (int64_t)(g.nx + 1) * K; }
__device__ __forceinline__ int64_t mg_vidx(const Grid& g, int I, int K) {
  return g.mn() + I + (int64_t)g.nx * (K - 1);
}

}  // namespace drd
