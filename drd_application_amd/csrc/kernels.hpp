// kernels.hpp -- kernel launchers and the device-side BiCGSTAB state machine.
#pragma once

#include <string>

#include "drd_common.hpp"

namespace drd {

// the message drd_last_error() returns (entry points outside drd_api.hip)
void api_set_error(const std::string& msg);

// matvec modes (see k_velo_matvec)
enum { MV_PLAIN = 0, MV_DOT1 = 1, MV_DOT2 = 2, MV_INIT = 3 };

// phases whose finished reductions drive the state machine
enum { PH_INIT_NORMS = 0, PH_INIT_R, PH_MV1, PH_S, PH_MV2, PH_XR, PH_NONE };

struct Coef {
  const double *rho, *eta, *chi, *bdr, *bu, *bg, *minner, *m0;
};

struct VeloRhsArgs {
  const double *u, *v, *u_pre, *v_pre, *p, *psi1, *psi2;
  const double *rho, *eta, *bu, *ph_after, *rmu_after, *pf1, *pf2, *uL;
  const double* bdr;  // bdr_ph (the MF variant's surface-tension term)
};

struct PhmuRhsArgs {
  const double *ph, *rmu, *u, *v, *bdr, *bu, *bg, *cfB, *cfT;
  double cosa, sqrt2;
};

// Decisions of LIB/BiCGSTABREVCOM.f taken where each reduction finishes.
// Runs on thread 0 of the last workgroup of the producing kernel.
__device__ inline void kfinalize(KState* st, int phase, const double* t) {
  switch (phase) {
    case PH_INIT_NORMS: {  // t = {x.x, b.b}            REVCOM:229-253
      st->xnz = (sqrt(t[0]) != 0.0);
      double bn = sqrt(t[1]);
      st->bnrm2 = (bn == 0.0) ? 1.0 : bn;
      break;
    }
    case PH_INIT_R: {  // t = {r.r} with rtld = r      REVCOM:248-263
      if (st->xnz && sqrt(t[0]) <= st->tol) {  // ITER and RESID stay as passed
        st->info = 0;
        st->done = 1;
        break;
      }
      st->iter = 1;
      st->rho = t[0];
      if (fabs(st->rho) < st->breaktol) {
        st->info = -10;
        st->done = 1;
      }
      break;
    }
    case PH_MV1:  // t = {rtld.v}                     REVCOM:307
      st->alpha = st->rho / t[0];
      break;
    case PH_S: {  // t = {s.s}                        REVCOM:311-316
      const double sn = sqrt(t[0]);
      st->snrm = sn;
      if (sn <= st->tol) {
        st->resid = sn / st->bnrm2;
        st->info = 0;
        st->early = 1;
      }
      break;
    }
    case PH_MV2:  // t = {t.s, t.t}                   REVCOM:349-350
      st->omega = t[0] / t[1];
      break;
    case PH_XR: {  // t = {r.r, rtld.r}  STOPTEST2.f:13-16, REVCOM:374-388, :261-264
      st->resid = sqrt(t[0]) / st->bnrm2;
      if (st->resid <= st->tol) {
        st->info = 0;
        st->done = 1;
        break;
      }
      if (st->iter == st->maxit) {
        st->info = 1;
        st->done = 1;
        break;
      }
      if (fabs(st->omega) < st->breaktol) {
        st->info = -11;
        st->done = 1;
        break;
      }
      st->rho1 = st->rho;
      st->iter = st->iter + 1;
      st->rho = t[1];
      if (fabs(st->rho) < st->breaktol) {
        st->info = -10;
        st->done = 1;
      }
      break;
    }
    default:
      break;
  }
  st->red[0] = t[0];
}

// Exact diagonal of the BDF2 velocity operator at packed unknown q
// (MatrixVector_uv_2_CG TP/VeloSolve_2_combined.f90:607-670 restricted to the
// coefficient of the row's own unknown, wall and side rules included).
template <class T>
__device__ inline double velo_diag2(const Grid& g, const Scal& s, const T* __restrict__ rho,
                                    const T* __restrict__ eta, const T* __restrict__ chi, int64_t q) {
  const int nx = g.nx, ny = g.ny;
  auto R = [&](int i, int k) { return (double)rho[g.ci(i, k)]; };
  auto E = [&](int i, int k) { return (double)eta[g.ci(i, k)]; };
  auto X = [&](int i, int k) { return (double)chi[g.ci(i, k)]; };
  double d;
  if (q < g.mn()) {
    const int k = g.by_nxu.div(q), i = (int)(q - (int64_t)k * g.nxu());
    d = 1.5 * s.re * (R(i, k) + R(i - 1, k)) / 2.0 + s.penalty_k * (X(i, k) + X(i - 1, k)) / 2.0 * s.dt;
    d += 2.0 * s.dt * (E(i, k) + E(i - 1, k)) / s.dr2;
    const double cr = (E(i, k + 1) + E(i - 1, k + 1) + E(i, k) + E(i - 1, k)) / 4.0;
    const double cl = (E(i, k) + E(i - 1, k) + E(i, k - 1) + E(i - 1, k - 1)) / 4.0;
    d += s.dt * ((k == 0 ? 2.0 * cl : cl) + (k == ny - 1 ? 2.0 * cr : cr)) / s.dz2;
  } else {
    const int64_t t = q - g.mn();
    const int k = g.by_nx.div(t) + 1, i = (int)(t - (int64_t)(k - 1) * nx);
    d = 1.5 * s.re * (R(i, k) + R(i, k - 1)) / 2.0 + s.penalty_k * (X(i, k) + X(i, k - 1)) / 2.0 * s.dt;
    d += 2.0 * s.dt * (E(i, k) + E(i, k - 1)) / s.dz2;
    const double cr = (E(i + 1, k) + E(i, k) + E(i + 1, k - 1) + E(i, k - 1)) / 4.0;
    const double cl = (E(i, k) + E(i - 1, k) + E(i, k - 1) + E(i - 1, k - 1)) / 4.0;
    d += s.dt * ((i == 0 ? 0.0 : cl) + (i == nx - 1 ? 0.0 : cr)) / s.dr2;
  }
  return d;
}

// ---- CH block preconditioner pieces (kernels_prec.hip, kernels_stencil.hip) ----
// diagonal coefficient of op_double_gradient_Neum at (i,j) (TP/matrix_tool.f90:178-219);
// the neighbour mobilities are loaded unconditionally (clamped indices) and
// only the sums are selected, so a thread issues all its loads at once
// A cell field at cell (i, j) and its four neighbours, the neighbour index
// clamped to the cell at a wall (the value the reference's one-sided rules
// use there).  Loading the five first lets a thread issue all its loads
// before any arithmetic; the *_v forms below take the values.
template <class R>
struct Nb5T {
  R c, xl, xr, yl, yr;
};
using Nb5 = Nb5T<double>;
// (periodic x: the x neighbours wrap)
__device__ __forceinline__ Nb5 nb5_cell(const Grid& g, const double* __restrict__ a, int i, int j) {
  const int il = i > 0 ? i - 1 : (g.px ? g.nx - 1 : i), ir = i < g.nx - 1 ? i + 1 : (g.px ? 0 : i);
  const int jl = j > 0 ? j - 1 : j, jr = j < g.ny - 1 ? j + 1 : j;
  return Nb5{a[g.ci(i, j)], a[g.ci(il, j)], a[g.ci(ir, j)], a[g.ci(i, jl)], a[g.ci(i, jr)]};
}
// the same for a packed interior field x[i + nx j]
__device__ __forceinline__ Nb5 nb5_packed(const Grid& g, const double* x, int i, int j) {
  const int il = i > 0 ? i - 1 : (g.px ? g.nx - 1 : i), ir = i < g.nx - 1 ? i + 1 : (g.px ? 0 : i);
  const int jl = j > 0 ? j - 1 : j, jr = j < g.ny - 1 ? j + 1 : j;
  const int64_t nx = g.nx;
  return Nb5{x[i + nx * j], x[il + nx * j], x[ir + nx * j], x[i + nx * jl], x[i + nx * jr]};
}

// diagonal coefficient of op_double_gradient_Neum at (i,j) (TP/matrix_tool.f90:178-219)
// from the mobility at the cell and its neighbours; only the sums are selected.
// R = double is the reference's arithmetic; R = float (the FP32
// preconditioner only) computes in FP32 and multiplies by reciprocals where
// the reference divides.
template <class R = double>
__device__ __forceinline__ R dgn_diag_v(const Grid& g, const Scal& s, const Nb5T<R>& mo, int i, int j) {
  const R tl = (mo.xl + mo.c) / R(2.0), tr = (mo.c + mo.xr) / R(2.0);
  const R bl = (mo.yl + mo.c) / R(2.0), br = (mo.c + mo.yr) / R(2.0);
  R dx = R(0.0), dy = R(0.0);
  dx = (g.px || i > 0) ? dx + tl : dx;
  dx = (g.px || i < g.nx - 1) ? dx + tr : dx;
  dy = j > 0 ? dy + bl : dy;
  dy = j < g.ny - 1 ? dy + br : dy;
  if (sizeof(R) == sizeof(float)) return -(dx * R(s.idr2) + dy * R(s.idz2));
  return -(dx / R(s.dr2) + dy / R(s.dz2));
}

// The 2x2 block of cell (i, j) from bu at the cell and the two mobilities
// around it (R as for dgn_diag_v): its diagonal d11, d22 and q = 1 / det
// (FP32) or det (FP64).  The off-diagonal entries are -bv and bv + adj.
template <class R = double>
__device__ __forceinline__ void phmu_block_f(const Grid& g, const Scal& s, R bv, const Nb5T<R>& mi,
                                             const Nb5T<R>& m0, int i, int j, R& d11o, R& d22o, R& qo) {
  const R d11 = R(s.s_I) - R(s.I) * dgn_diag_v<R>(g, s, mi, i, j) +
                ((j == 0) ? R(s.I_dz2) * (R(1.0) - R(s.coef_bot)) : R(0.0)) +
                ((j == g.ny - 1) ? R(s.I_dz2) * (R(1.0) - R(s.coef_top)) : R(0.0));
  const R d12 = -bv;
  const R d21 = bv + R(s.adj);
  // mobility-free floor (s.floor22) keeps the block invertible where M_0 vanishes
  R d22 = -R(s.dt) * R(s.xld) * dgn_diag_v<R>(g, s, m0, i, j);
  d22 = d22 > R(s.floor22) ? d22 : R(s.floor22);
  const R det = d11 * d22 - d12 * d21;
  d11o = d11;
  d22o = d22;
  qo = sizeof(R) == sizeof(float) ? R(1.0) / det : det;
}
// x = B^-1 (r1, r2) from the block's factors (phmu_block_f)
template <class R = double>
__device__ __forceinline__ void phmu_block_x(const Scal& s, R bv, R d11, R d22, R q, R r1, R r2, R& x1, R& x2) {
  const R d12 = -bv;
  const R d21 = bv + R(s.adj);
  if (sizeof(R) == sizeof(float)) {
    x1 = (d22 * r1 - d12 * r2) * q;
    x2 = (-d21 * r1 + d11 * r2) * q;
  } else {
    x1 = (d22 * r1 - d12 * r2) / q;
    x2 = (-d21 * r1 + d11 * r2) / q;
  }
}
// x = B_q^-1 (r1, r2) for cell q (i, j), B_q the 2x2 block of the cell
template <class R = double>
__device__ __forceinline__ void phmu_block_v(const Grid& g, const Scal& s, R bv, const Nb5T<R>& mi,
                                             const Nb5T<R>& m0, int i, int j, R r1, R r2, R& x1, R& x2) {
  R d11, d22, q;
  phmu_block_f<R>(g, s, bv, mi, m0, i, j, d11, d22, q);
  phmu_block_x<R>(s, bv, d11, d22, q, r1, r2, x1, x2);
}
__device__ __forceinline__ void phmu_block_solve(const Grid& g, const Scal& s, const double* __restrict__ bu,
                                                 const double* __restrict__ minner,
                                                 const double* __restrict__ m0, int i, int j, double r1,
                                                 double r2, double& x1, double& x2) {
  phmu_block_v(g, s, bu[g.ci(i, j)], nb5_cell(g, minner, i, j), nb5_cell(g, m0, i, j), i, j, r1, r2, x1, x2);
}

// interpolate_coef TP/interpolate_coef.f90:4-115: fpd_ph, rho, eta from the
// particle and swimmer centres (pts: (x, y) pairs, np then nswm) and bdr_ph
// particle coupling: bodies are (x, y, radius, fx, fy); row chunks per body
// such that 3 * chunks * nbody <= cap; sums (u, v null: weights only) then
// fin[3b] = weight sum * dr * dz, fin[3b+1..2] = weighted velocity averages
int body_chunks(const Grid& g, int nbody, int64_t cap);
void launch_body_sums(hipStream_t st, const Grid& g, double dr, double dz, double I, double lx, double ly,
                      const double* bodies, int nbody, int chunks, const double* u, const double* v, double* part,
                      double* fin);
void launch_body_spread(hipStream_t st, const Grid& g, double dr, double dz, double I, double lx, double ly,
                        const double* bodies, int nbody, const double* fin, double* pf1, double* pf2);
void launch_interpolate_coef(hipStream_t st, const Grid& g, double dr, double dz, double I, double lrho,
                             double leta, const double* pts, int np, double radius, int nswm,
                             double swm_radius, const double* bdr, double* fpd, double* rho, double* eta);
// FP32 copies of the CH block coefficients (cell fields, ghosts included)
// FP32 copies of bu, M_inner, M_0 (ghosted cell layout) and each interior
// cell's block factors d11, d22, 1 / det (phmu_block_f<float>, same layout)
struct PhmuCoef32 {
  const float *bu, *minner, *m0, *d11, *d22, *q;
};
void launch_phmu_coef32(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const PhmuCoef32& c32);
// Event pairs for timing: group k (k < cap) of consecutive launches of the
// instrumented kernel is bracketed by ev[2k] and ev[2k+1]; *n counts the
// groups.
struct KTimer {
  hipEvent_t* ev;
  int cap;
  int* n;
};
constexpr int kRbNS = 2;        // red-black sweeps per launch
constexpr int kRbgsMiddle = 6;  // launches per red-black application besides the first and the last
// over-relaxation weight of each red-black cell update: x_c += w B_c^-1 (b - A x)_c
constexpr double kRbOmega = 1.25;
// c32 null: FP64 throughout, work = 2 ncell doubles; else FP32 storage,
// work = 6 ncell floats, and tm (optional) times the middle launches of each
// application as one group of kRbgsMiddle launches
void launch_phmu_rbgs(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const PhmuCoef32* c32,
                      const double* b, double* x, void* work, const KState* ks, const KTimer* tm = nullptr);
// one middle FP32 launch (kRbNS sweeps), xi -> xo with the FP32 b (stand-alone timing)
void launch_phmu_rb_sweep32(hipStream_t st, const Grid& g, const Scal& s, const PhmuCoef32& c32, const float* b,
                            const float* xi, float* xo, const KState* ks);

// ---- multigrid V-cycle for the BDF2 velocity operator (kernels_mg.hip) ----
enum { MG_FIRST = 0, MG_SWEEP = 1, MG_RES = 2, MG_SWEEP0 = 3 };
struct VeloMG;
// f32: the levels' coefficients, diagonals and vectors are stored in FP32
// and their stencils evaluated in FP32; the input b, the output z and the
// coarsest level stay FP64
// The multigrid's coarsest solve for any grid of at most 256 x 256 cells
// (kernels_coarse.hip): the fast-Poisson preconditioner with dense line
// transforms.  region: coarse_bytes(g) per instance, or nullptr.
struct CoarseSolve;
bool coarse_fits(const Grid& g);
size_t coarse_bytes(const Grid& g);
CoarseSolve* coarse_create(const Grid& g, const Scal& s, int* err, void* region);
void coarse_destroy(CoarseSolve* cs);
void coarse_setup(hipStream_t st, CoarseSolve* cs, const Grid& g, const Scal& s, const Coef& c);
void coarse_apply(hipStream_t st, CoarseSolve* cs, const Grid& g, const Scal& s, const double* r, double* z,
                  const KState* ks);
// region: where the buffers go (velo_mg_bytes bytes per instance; instance
// k's copy k * bat_bs() bytes further, as every per-instance buffer of a
// batch group), or nullptr for allocations of their own
size_t velo_mg_bytes(const Grid& g, bool f32);
VeloMG* velo_mg_create(const Grid& g, const Scal& s, bool f32, int* err, void* region = nullptr);
void velo_mg_destroy(VeloMG* mg);
bool velo_mg_f32(const VeloMG* mg);
void velo_mg_setup(hipStream_t st, VeloMG* mg, const Coef& c0);
void velo_mg_apply(hipStream_t st, VeloMG* mg, const double* b, double* z, const KState* ks);
int velo_mg_levels(const VeloMG* mg);
// on (the default): V(2,2) cycles by the LDS-tiled legs of kernels_mgf.hip;
// off: one launch per sweep and transfer (the same values bit for bit)
void velo_mg_set_fused(VeloMG* mg, bool on);
// its kernels (kernels_stencil.hip); a buffer's element type is FP32 where
// its flag says so, else FP64
struct MgCoef {
  const void *rho, *eta, *chi;
};
void launch_mg_convert(hipStream_t st, const Grid& g, const Coef& c, float* rho, float* eta, float* chi);
void launch_mg_coarsen(hipStream_t st, bool f32, bool o32, const Grid& gf, const Grid& gc, const MgCoef& f,
                       void* rc, void* ec, void* cc);
// dw = omega / the exact diagonal
void launch_mg_diag(hipStream_t st, bool f32, const Grid& g, const Scal& s, const MgCoef& c, void* dw,
                    double omega);
// coefficients, x and dw of the level's type (f32); b and out of their own;
// MG_FIRST and MG_SWEEP0 also write b to bc as FP32 if bc is given
void launch_mg_smooth(hipStream_t st, int kind, bool f32, bool b32, bool o32, const Grid& g, const Scal& s,
                      const MgCoef& c, const void* x, const void* b, const void* dw, void* out, float* bc,
                      const KState* ks);
// LDS-tiled legs (kernels_mgf.hip).  down: b (FP64 when !b32), dw, the
// level's coefficients -> x2 in xd, bcopy = FP32 b (optional), bc = R r
// (FP32 when bc32).  up: out = two sweeps from xd + P xc (xc FP32 when c32,
// out FP32 when o32).
void launch_mg_down(hipStream_t st, bool f32, bool b32, bool bc32, const Grid& g, const Grid& gc, const Scal& s,
                    const MgCoef& c, const void* b, const void* dw, void* xd, float* bcopy, void* bc,
                    const KState* ks);
void launch_mg_up(hipStream_t st, bool f32, bool c32, bool o32, const Grid& g, const Grid& gc, const Scal& s,
                  const MgCoef& c, const void* b, const void* dw, const void* xd, const void* xc, void* out,
                  const KState* ks);
void launch_mg_restrict(hipStream_t st, bool r32, bool b32, const Grid& gf, const Grid& gc, const void* r,
                        void* bc, const KState* ks);
void launch_mg_prolong_add(hipStream_t st, bool c32, bool f32, const Grid& gc, const Grid& gf, const void* xc,
                           void* x, const KState* ks);

// ---- reduction finish (kernels_krylov.hip) ----
// k_finalize runs as one workgroup after a reducing kernel: it adds the
// tile partials (layout [tile][segment][value]) in the canonical order, then
// either applies the BiCGSTAB decisions of `phase` (kfinalize) or, for
// PH_NONE, writes the nv totals to out.  `guard` repeats the producer's
// own skip test so a finished solve stays untouched.
enum { FIN_ALWAYS = 0, FIN_ACTIVE = 1 };
void launch_finalize(hipStream_t st, const double* part, int64_t ntiles, int nseg, int nv, double* out,
                     KState* ks, int phase, int guard);
inline unsigned tiles_of(int64_t n) { return (unsigned)((n + kTile - 1) / kTile); }
// the finish of a fused matvec sweep (nothing for MV_PLAIN)
inline void launch_mv_finalize(hipStream_t st, int mode, const double* part, int64_t ntiles, int nseg,
                               KState* ks) {
  if (mode == MV_DOT1) launch_finalize(st, part, ntiles, nseg, 1, nullptr, ks, PH_MV1, FIN_ACTIVE);
  if (mode == MV_DOT2) launch_finalize(st, part, ntiles, nseg, 2, nullptr, ks, PH_MV2, FIN_ACTIVE);
  if (mode == MV_INIT) launch_finalize(st, part, ntiles, nseg, 1, nullptr, ks, PH_INIT_R, FIN_ACTIVE);
}

// ---- launchers (kernels_stencil.hip) ----
// finish = false leaves the partial sums unreduced (kernel timing only);
// with events ev0, ev1 the sweep's own start and end are recorded in them
// (the dispatch's timestamps, not markers queued around it)
void launch_velo_matvec(hipStream_t st, int bdf, int mode, const Grid& g, const Scal& s, const Coef& c,
                        const double* x, double* y, double alpha, double beta, double* a0,
                        const double* b, double* part, KState* ks, bool finish = true,
                        hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void launch_phmu_matvec(hipStream_t st, int mode, const Grid& g, const Scal& s, const Coef& c,
                        const double* x, double* y, double alpha, double beta, double* a0,
                        const double* b, double* part, KState* ks, bool finish = true);
// mcoef: the mass coefficient, re (PSolve_uv_CG) or 1.5 re (the MF PSolve_uv_2_CG)
void launch_velo_jacobi_diag(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, double* diag,
                             double mcoef);
void launch_div_diag(hipStream_t st, int64_t n, const double* b, const double* d, double* x,
                     const KState* ks);
void launch_phmu_jacobi(hipStream_t st, const Grid& g, const Scal& s, const double* b, double* x,
                        const KState* ks);
void launch_velo_pack(hipStream_t st, const Grid& g, const double* u, const double* v, double* x);
void launch_velo_unpack(hipStream_t st, const Grid& g, const Scal& s, const double* x, const double* uL,
                        double* u, double* v);
// periodic x: the ghost columns of a field (kind 0 cell, 1 u, 2 v) from its interior
void launch_xghost_fill(hipStream_t st, const Grid& g, double* f, int kind);
void launch_velo_rhs(hipStream_t st, int bdf, const Grid& g, const Scal& s, const VeloRhsArgs& a,
                     double* rhs);
void launch_phmu_m0(hipStream_t st, const Grid& g, int mode, const double* ph, const double* bdr,
                    double* m0);
void launch_phmu_rhs(hipStream_t st, const Grid& g, const Scal& s, const PhmuRhsArgs& a, double* rhs,
                     double* x0, double* pn);
void launch_phmu_unpack(hipStream_t st, const Grid& g, const Scal& s, const double* x, const double* pn,
                        double* ph, double* mu);
void launch_phmu_pack(hipStream_t st, const Grid& g, const double* ph, const double* mu, double* x);

// ---- launchers (kernels_krylov.hip) ----
// Reducing kernels take the vector's segment count (1, or 2 for the phi|mu
// vector of the Cahn-Hilliard system; see tile_publish).
void launch_init_norms(hipStream_t st, int64_t n, int nseg, const double* x, const double* b, double* part,
                       KState* ks);
void launch_bicg_p(hipStream_t st, int64_t n, const double* r, const double* v, double* p, KState* ks);
void launch_bicg_s(hipStream_t st, int64_t n, int nseg, double* r, const double* v, double* part, KState* ks);
// the reductions of a fused matvec mode for a product y formed outside the library (one segment)
void launch_ext_mv(hipStream_t st, int mode, int64_t n, const double* y, double* a0, double* part, KState* ks);
void launch_bicg_xr(hipStream_t st, int64_t n, int nseg, double* x, double* r, const double* rt,
                    const double* ph, const double* sh, const double* t, double* part, KState* ks,
                    bool finish = true);
void launch_copy(hipStream_t st, int64_t n, const double* a, double* b);
void launch_axpby(hipStream_t st, int64_t n, double alpha, const double* x, double beta, double* y);
// sum of squares of (b - ax) and of b -> ks->red[0], red[1] (FP64, canonical order)
void launch_resid_norms(hipStream_t st, int64_t n, int nseg, const double* b, const double* ax, double* part,
                        KState* scratch_ks);
void launch_cell_axpy(hipStream_t st, int64_t n, const double* a, double* b);  // b = a + b
// FGMRES: res[i] = (w, V_i); w += sign * sum h_i V_i (and ||w||^2); b = a / s.
// part holds nv * nseg partials per tile of kTile elements.
void launch_multidot(hipStream_t st, int64_t n, int nseg, const double* w, const double* V, int64_t ldv, int nv,
                     double* part, double* res);
void launch_multiaxpy(hipStream_t st, int64_t n, int nseg, double* w, const double* V, int64_t ldv, int nv,
                      const double* h, double sign, double* part, double* res, int want_norm);
void launch_scale_into(hipStream_t st, int64_t n, const double* a, double s, double* b);

// ---- fast-path preconditioners (kernels_prec.hip) ----
void launch_phmu_block(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const double* b, double* x,
                       const KState* ks);
// BiCGSTAB p and s updates of the phi/mu solve with the block preconditioner
// fused in: p and phat = B^-1 p; s, ||s||^2 (then k_finalize) and shat = B^-1 s
void launch_bicg_p_block(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const double* r,
                         const double* v, double* p, double* ph, KState* ks);
void launch_bicg_s_block(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, double* r, const double* v,
                         double* sh, double* part, KState* ks);
struct VeloPrec;
// per-instance buffers of the preconditioner (Nuv doubles, Nuv floats and
// 2 Nuv doubles of transform work); null: the preconditioner allocates them
struct VeloPrecBufs {
  double* inv_s64;
  float* inv_s32;
  void* work;
};
VeloPrec* velo_prec_create(const Grid& g, const Scal& s, bool fp32, int* err, const VeloPrecBufs* ext = nullptr);
int velo_prec_update(VeloPrec* pc, const Scal& s);  // new dt, re, dx, dz
void velo_prec_destroy(VeloPrec* pc);
// diagonal scaling from the current coefficients (once per solve)
void velo_prec_setup(hipStream_t st, VeloPrec* pc, const Grid& g, const Scal& s, const Coef& c);
void velo_prec_apply(hipStream_t st, VeloPrec* pc, const Grid& g, const double* r, double* z, const KState* ks);
// picks the fastest transpose tile shape (32x32, 16x64, 64x16) by timing;
// ms[3] receives the mean time per application of each
int velo_prec_tune(hipStream_t st, VeloPrec* pc, const Grid& g, const double* r, double* z, int reps, double* ms);

// ---- PhSolve, VCOST and ESOLP (kernels_phsolve.hip) ----
struct Dct1Tab;
Dct1Tab* dct1_table_create(int n);  // VCOST of lines of n points
void dct1_table_destroy(Dct1Tab* t);
void launch_ph_dct_prec(hipStream_t s, const Grid& g, const Dct1Tab* tx, const Dct1Tab* ty, const double* ph_lcc,
                        const double* b, double* x, double* work, const KState* ks);
void launch_ph_rhs(hipStream_t st, const Grid& g, const Scal& s, const double* ph, const double* u, const double* v,
                   const double* bdr, const double* bg, const double* minner, const double* rmu, double cosa,
                   double sqrt2, double* f2, double* rhs, double* x0);
void launch_ph_matvec(hipStream_t st, int mode, const Grid& g, const Scal& s, const double* m0,
                      const double* minner, const double* x, double* t, double* y, double alpha, double beta,
                      double* a0, const double* b, double* part, KState* ks);
// PhSolve_ori TP/PhSolve_ori.f90 and the solid re-initialisation (kernels_phori.hip)
void launch_phori_rhs(hipStream_t st, const Grid& g, const Scal& s, const double* ph, const double* u,
                      const double* v, const double* angB, const double* angT, double alpha_s, double* rhs,
                      double* x0, double* pn);
void launch_phori_matvec(hipStream_t st, int mode, const Grid& g, const Scal& s, double cbot, double ctop,
                         const double* x, double* t, double* y, double alpha, double beta, double* a0,
                         const double* b, double* part, KState* ks);
void launch_phori_finish(hipStream_t st, const Grid& g, const Scal& s, double cbot, double ctop, const double* x,
                         const double* pn, double* ph, double* mu);
void launch_reinit_start(hipStream_t st, const Grid& g, const double* bdr, double* ph, double* m0, double* minner);
void launch_reinit_end(hipStream_t st, const Grid& g, const double* ph, double* bdr, double* pen);
void launch_ph_finish(hipStream_t st, const Grid& g, const Scal& s, const double* x, const double* f2,
                      const double* m0, double* mu, double* ph);
// tabs: [ua|va], [ub|vb], [uc|vc], each table (1:n, 0:ld-1) column-major
void launch_esolp(hipStream_t s, const Grid& g, const Dct1Tab* tu, const Dct1Tab* tv, const double* tabs,
                  const double* b, double* x, double* work, const KState* ks);

// ---- launchers (kernels_dct.hip) ----
struct DctPlan;
DctPlan* dct_plan_create(int nx, int ny);
void dct_plan_destroy(DctPlan* p);
// pressure projection: psi2 = coef * Poisson^-1 div(u,v); p = psi2 + p
void launch_pressure(hipStream_t st, DctPlan* pl, const Grid& g, const Scal& s, const double* u,
                     const double* v, const double* pre_lcc, double coef, double* psi, double* p,
                     double* work, double* part, KState* ks);
// 2-D FFTW-convention REDFT10 / REDFT01 of an nx x ny array (tests)
void launch_dct2_2d(hipStream_t st, DctPlan* pl, double* a, double* work);
void launch_dct3_2d(hipStream_t st, DctPlan* pl, double* a, double* work);

}  // namespace drd
