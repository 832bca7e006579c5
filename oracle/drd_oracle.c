/*
 * drd_oracle.c -- CPU oracle for the DRD-Micro implicit-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: tests/, the smoke test
 * and bench.py's cpu_baseline leg load it; the product never does.
 *
 * Plain C, FP64, serial.  Each function restates one reference routine and
 * cites it (paths relative to the reference checkout; TP/ is
 * two_phase_flow_complex_domain/, MF/ microfluidics_complex_domain/, LIB/
 * lib/).  Floating-point expressions keep the reference's evaluation order
 * (Fortran evaluates a chain of equal-precedence operators left to right),
 * so with -ffp-contract=off the stencil routines reproduce the Fortran
 * results bit for bit; reductions are plain sequential sums.
 *
 * How each part is pinned is listed in tests/ (see the test module headers):
 * operators and preconditioners against the reference Fortran compiled from
 * its own sources (oracle/_ref), transforms against their published
 * definitions and scipy, BiCGSTAB against dense LU solves and the
 * reference's documented control flow, right-hand sides against closed
 * forms of isolated terms.
 */
#include "drd_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* pi as the module parameter of TP/global.f90:39 (rounded to double)      */
/* OpenMP: loops whose iterations write disjoint elements run in parallel
 * when the grid is large enough to pay for it; every element's arithmetic
 * and every sum's order are those of the serial code, so results do not
 * depend on the thread count. */
#define PAR_MIN 16384
#define OMP_STR(x) #x
#define OMP_FOR(cond) _Pragma(OMP_STR(omp parallel for schedule(static) if(cond)))

static const double PI_G = 3.14159265358979323846264338327950288419716939937510;

struct oc_ctx {
  drd_params p;
  double* f[DRD_F_COUNT];
  double* pre_lcc; /* (0:nx,0:ny)       TP/setup.f90:122-131 */
  double* ph_lcc;  /* (0:nx-1,0:ny-1)   TP/setup.f90:149-155 */
  /* ESOLP tables TP/setup.f90:157-235: ua..uc (1:ny,0:nx), va..vc (1:ny-1,0:nx-1) */
  double *ua, *ub, *uc, *va, *vb, *vc;
  double coef_bot, coef_top; /* hidden outputs of ph_mu_solve     */
  double* uL;      /* (0:ny) hidden inflow profile TP/VeloSolve_2_combined.f90:289 */
  double* uR;
  /* scratch */
  double *w1, *w2, *w3, *w4, *w5, *w6;
  double* work;    /* Krylov WORK(N,7) */
  int guess;       /* initial guess of the BDF2 solves: 0 previous step, 1 2x^n - x^{n-1} */
  int64_t nmax;
};

/* ---- index helpers -------------------------------------------------- */
#define NX (c->p.nx)
#define NY (c->p.ny)
/* cell field (-1:nx,-1:ny) */
#define CI(i, k) ((size_t)((i) + 1) + (size_t)(NX + 2) * (size_t)((k) + 1))
/* u field (-1:nx+1,-1:ny) */
#define UI(i, k) ((size_t)((i) + 1) + (size_t)(NX + 3) * (size_t)((k) + 1))
/* v field (-1:nx,0:ny) */
#define VI(i, k) ((size_t)((i) + 1) + (size_t)(NX + 2) * (size_t)(k))

static int64_t ncell(const drd_params* p) { return (int64_t)(p->nx + 2) * (p->ny + 2); }
static int64_t nufield(const drd_params* p) { return (int64_t)(p->nx + 3) * (p->ny + 2); }
static int64_t nvfield(const drd_params* p) { return (int64_t)(p->nx + 2) * (p->ny + 1); }
/* Periodic x (lr_bd_type 'p'), an addition without a TP counterpart (TP's
 * sides are Neumann or a Dirichlet inflow; the periodic packing follows
 * AM/matrix_tool.f90:81-139): a u row has nx unknowns (u(nx) = u(0)), and
 * every ghost column, wall-ghost rows included, holds the periodic copy of
 * an interior column: f(-1) = f(nx-1), f(nx) = f(0), and for u f(nx+1) =
 * f(1). */
static int per_x(const drd_params* p) { return p->lr_bd_type == 'p'; }
static int nxu(const drd_params* p) { return p->nx + (per_x(p) ? 0 : 1); }
static int64_t nuv(const drd_params* p) {
  return (int64_t)nxu(p) * p->ny + (int64_t)p->nx * (p->ny - 1);
}

static void xghost_c(const oc_ctx* c, double* f) {
  for (int j = -1; j <= NY; ++j) {
    f[CI(-1, j)] = f[CI(NX - 1, j)];
    f[CI(NX, j)] = f[CI(0, j)];
  }
}
static void xghost_u(const oc_ctx* c, double* f) {
  for (int j = -1; j <= NY; ++j) {
    f[UI(-1, j)] = f[UI(NX - 1, j)];
    f[UI(NX, j)] = f[UI(0, j)];
    f[UI(NX + 1, j)] = f[UI(1, j)];
  }
}
static void xghost_v(const oc_ctx* c, double* f) {
  for (int j = 0; j <= NY; ++j) {
    f[VI(-1, j)] = f[VI(NX - 1, j)];
    f[VI(NX, j)] = f[VI(0, j)];
  }
}

int64_t oc_field_size(const drd_params* p, int32_t f) {
  switch (f) {
    case DRD_F_U:
    case DRD_F_U_PRE:
      return nufield(p);
    case DRD_F_V:
    case DRD_F_V_PRE:
      return nvfield(p);
    case DRD_F_ANGLEB:
    case DRD_F_ANGLET:
      return p->nx + 1;
    default:
      if (f >= 0 && f < DRD_F_COUNT) return ncell(p);
      return 0;
  }
}

void oc_params_default(drd_params* p, int32_t nx, int32_t ny) {
  memset(p, 0, sizeof(*p));
  /* TP/input and TP/mcl6.f90 / TP/finit.f90 defaults */
  p->nx = nx;
  p->ny = ny;
  p->variant = DRD_VARIANT_TP;
  p->phi_mode = 2;
  p->lr_bd_type = 'n';
  p->velo_prec = DRD_PREC_REFERENCE;
  p->maxit_velo1 = 800;
  p->maxit_velo2 = 2000;
  p->maxit_phmu = 5000;
  p->maxit_ph = 500;
  p->dx = 1.0 / nx;
  p->dz = 1.0 / ny;
  p->dt = 1.0e-4 / 2.0;
  p->re = 3;
  p->ca = 12;
  p->xld = 5.0e-2;
  p->s_implicit = 1.5;
  p->I_thickness = 0.01;
  p->ggx = 0;
  p->ggz = 0;
  p->uwT = 0;
  p->uwB = 0;
  p->xlsB = 0;
  p->tol = 1.0e-8;
  p->penalty_k = 1.0e8;
  p->angle = 60;
  p->lambda_rho = 1;
  p->bdr_ph_used_adjust = 1.0e-6;
  p->twall = 0.1;
  p->UU = 0;
  p->bdr_upp = 1;
  p->bdr_low = 0;
}

/* Ph_Lcc TP/setup.f90:137-155 for the step dt */
static void build_ph_lcc(oc_ctx* c, double dt) {
  const drd_params* p = &c->p;
  const int nx = p->nx, ny = p->ny;
  const double dr2 = p->dx * p->dx, dz2 = p->dz * p->dz;
  double* fmx = (double*)calloc((size_t)nx + 1, sizeof(double));
  double* fmy = (double*)calloc((size_t)ny + 1, sizeof(double));
  for (int i = 0; i < nx; ++i) {
    double t1 = PI_G * i / 2.0 / (nx - 1);
    double s = sin(t1);
    fmx[i] = 4.0 * (s * s) / dr2;
  }
  for (int j = 0; j < ny; ++j) {
    double t1 = PI_G * j / 2.0 / (ny - 1);
    double s = sin(t1);
    fmy[j] = 4.0 * (s * s) / dz2;
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double temp = fmx[i] + fmy[j];
      c->ph_lcc[i + (size_t)nx * j] = 1. + p->xld * dt * p->s_implicit * temp / p->I_thickness +
                                       p->xld * p->I_thickness * dt * (temp * temp);
    }
  free(fmx);
  free(fmy);
}

int oc_init(const drd_params* prm, oc_ctx** out) {
  if (!prm || !out || prm->nx < 2 || prm->ny < 2) return DRD_E_ARG;
  if (per_x(prm) && (prm->variant != DRD_VARIANT_TP || (prm->nx & (prm->nx - 1)) != 0)) return DRD_E_UNSUPPORTED;
  oc_ctx* c = (oc_ctx*)calloc(1, sizeof(oc_ctx));
  c->p = *prm;
  for (int f = 0; f < DRD_F_COUNT; ++f)
    c->f[f] = (double*)calloc((size_t)oc_field_size(prm, f), sizeof(double));
  c->nmax = nuv(prm) > 2 * (int64_t)prm->nx * prm->ny ? nuv(prm) : 2 * (int64_t)prm->nx * prm->ny;
  /* scratch holds a field with ghosts or a Krylov vector, whichever is larger */
  int64_t big = nufield(prm) > ncell(prm) ? nufield(prm) : ncell(prm);
  if (c->nmax > big) big = c->nmax;
  double** ws[] = {&c->w1, &c->w2, &c->w3, &c->w4, &c->w5, &c->w6};
  for (int i = 0; i < 6; ++i) *ws[i] = (double*)calloc((size_t)big, sizeof(double));
  c->work = (double*)calloc((size_t)c->nmax * 7, sizeof(double));
  c->uL = (double*)calloc((size_t)prm->ny + 1, sizeof(double));
  c->uR = (double*)calloc((size_t)prm->ny + 1, sizeof(double));
  c->coef_bot = c->coef_top = 1.0;

  /* setup: Pre_Lcc and Ph_Lcc, TP/setup.f90:108-155 */
  const drd_params* p = prm;
  int nx = p->nx, ny = p->ny;
  double dr2 = p->dx * p->dx, dz2 = p->dz * p->dz;
  c->pre_lcc = (double*)calloc((size_t)(nx + 1) * (ny + 1), sizeof(double));
  c->ph_lcc = (double*)calloc((size_t)nx * ny, sizeof(double));
  double* fmx = (double*)calloc((size_t)nx + 1, sizeof(double));
  double* fmy = (double*)calloc((size_t)ny + 1, sizeof(double));
  for (int i = 0; i <= nx; ++i) {
    double t1 = PI_G * i / 2.0 / nx;
    double s = sin(t1);
    fmx[i] = 4.0 * (s * s) / dr2;
  }
  for (int j = 0; j <= ny; ++j) {
    double t1 = PI_G * j / 2.0 / ny;
    double s = sin(t1);
    fmy[j] = 4.0 * (s * s) / dz2;
  }
  if (per_x(p)) {
    /* periodic x: the 5-point eigenvalue 4 sin^2(pi kx / nx) / dr2 of the
     * Fourier mode kx, stored at the halfcomplex index m of its cosine
     * (m = kx) and sine (m = nx - kx) parts; the mean mode as Pre_Lcc(0,0) */
    for (int m = 0; m < nx; ++m) {
      const int kx = m <= nx / 2 ? m : nx - m;
      double t1 = PI_G * kx / nx;
      double sn = sin(t1);
      fmx[m] = 4.0 * (sn * sn) / dr2;
    }
    fmx[nx] = 0.0;
  }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) c->pre_lcc[i + (size_t)(nx + 1) * j] = -(fmx[i] + fmy[j]);
  c->pre_lcc[0] = -1.0;
  build_ph_lcc(c, p->dt);
  /* ESOLP tables TP/setup.f90:157-235 (BDF1 set).  The mass shift t1 set at
   * :159 is overwritten by the loops at :167 and :203, so the tables use the
   * last loop value, (pi*nx/nx)/2 and (pi*(nx-1)/(nx-1))/2. */
  {
    double t1 = 0.0;
    double slip = 1.0; /* slipratio; xlsB = xlsT = 0 on the no-slip path */
    double xb = 0.0 * (1.0 + slip) / 2.0, xt = 0.0 * (1.0 + slip) / 2.0;
    xb = (2.0 * xb - p->dz) / (2.0 * xb + p->dz);
    xt = (2.0 * xt - p->dz) / (2.0 * xt + p->dz);
    c->ua = (double*)calloc((size_t)ny * (nx + 1), sizeof(double));
    c->ub = (double*)calloc((size_t)ny * (nx + 1), sizeof(double));
    c->uc = (double*)calloc((size_t)ny * (nx + 1), sizeof(double));
    c->va = (double*)calloc((size_t)(ny - 1) * nx, sizeof(double));
    c->vb = (double*)calloc((size_t)(ny - 1) * nx, sizeof(double));
    c->vc = (double*)calloc((size_t)(ny - 1) * nx, sizeof(double));
    double* fx = (double*)calloc((size_t)nx + 1, sizeof(double));
    for (int i = 0; i <= nx; ++i) {
      t1 = PI_G * i / nx / 2.;
      double s = sin(t1);
      fx[i] = 4.0 * p->dt * (s * s) / dr2;
    }
#define TU(a, i, k) a[((i) - 1) + (size_t)ny * (k)]
    for (int k = 0; k <= nx; ++k) {
      TU(c->ua, 1, k) = 0.0;
      TU(c->ub, 1, k) = t1 + ((2.0 - xb) * p->dt / dz2 + fx[k]);
      TU(c->uc, 1, k) = (-1.0 * p->dt / dz2);
      for (int i = 2; i <= ny - 1; ++i) {
        TU(c->ua, i, k) = (-p->dt / dz2);
        TU(c->ub, i, k) = t1 + (2.0 * p->dt / dz2 + fx[k]);
        TU(c->uc, i, k) = (-p->dt / dz2);
      }
      TU(c->ua, ny, k) = (-1.0 * p->dt / dz2);
      TU(c->ub, ny, k) = t1 + ((2.0 - xt) * p->dt / dz2 + fx[k]);
      TU(c->uc, ny, k) = 0.0;
    }
#undef TU
    for (int i = 0; i < nx; ++i) {
      t1 = PI_G * i / (nx - 1) / 2.;
      double s = sin(t1);
      fx[i] = 4.0 * p->dt * (s * s) / dr2;
    }
#define TV(a, i, k) a[((i) - 1) + (size_t)(ny - 1) * (k)]
    for (int k = 0; k < nx; ++k) {
      TV(c->va, 1, k) = 0.0;
      TV(c->vb, 1, k) = t1 + 2.0 * p->dt / dz2 + fx[k];
      TV(c->vc, 1, k) = -1.0 * p->dt / dz2;
      for (int i = 2; i <= ny - 2; ++i) {
        TV(c->va, i, k) = -p->dt / dz2;
        TV(c->vb, i, k) = t1 + 2.0 * p->dt / dz2 + fx[k];
        TV(c->vc, i, k) = -p->dt / dz2;
      }
      TV(c->va, ny - 1, k) = -1.0 * p->dt / dz2;
      TV(c->vb, ny - 1, k) = t1 + 2.0 * p->dt / dz2 + fx[k];
      TV(c->vc, ny - 1, k) = 0.0;
    }
#undef TV
    free(fx);
  }
  free(fmx);
  free(fmy);
  *out = c;
  return 0;
}

const double* oc_esolp_table(oc_ctx* c, int32_t which) {
  const double* t[6] = {c->ua, c->ub, c->uc, c->va, c->vb, c->vc};
  return (which >= 0 && which < 6) ? t[which] : NULL;
}

/* Periodic x: a field written from outside takes the periodic ghost columns
 * (the device library applies the same rule to drd_set_field). */
int oc_fill_ghosts(oc_ctx* c, int32_t f) {
  if (!c || f < 0 || f >= DRD_F_COUNT) return DRD_E_ARG;
  if (!per_x(&c->p)) return 0;
  const int64_t n = oc_field_size(&c->p, f);
  if (n == nufield(&c->p)) xghost_u(c, c->f[f]);
  else if (n == nvfield(&c->p)) xghost_v(c, c->f[f]);
  else if (n == ncell(&c->p)) xghost_c(c, c->f[f]);
  return 0;
}

int oc_finalize(oc_ctx* c) {
  if (!c) return 0;
  for (int f = 0; f < DRD_F_COUNT; ++f) free(c->f[f]);
  free(c->w1); free(c->w2); free(c->w3); free(c->w4); free(c->w5); free(c->w6);
  free(c->work); free(c->uL); free(c->uR); free(c->pre_lcc); free(c->ph_lcc);
  free(c->ua); free(c->ub); free(c->uc); free(c->va); free(c->vb); free(c->vc);
  free(c);
  return 0;
}

double* oc_field(oc_ctx* c, int32_t f) {
  if (!c || f < 0 || f >= DRD_F_COUNT) return NULL;
  return c->f[f];
}
const double* oc_pre_lcc(oc_ctx* c) { return c->pre_lcc; }
const double* oc_ph_lcc(oc_ctx* c) { return c->ph_lcc; }

/* ---- BLAS-1 in a fixed summation order ---------------------------------
 * The reference links MKL's DDOT/DNRM2, whose summation order is not
 * specified.  The oracle fixes one, the path's canonical order: the
 * pairwise sum over the element index, padded with zeros to a power of two,
 *   S[lo, lo+2L) = S[lo, lo+L) + S[lo+L, lo+2L),
 * which does not depend on how a parallel implementation splits the work.
 * A vector of nseg equal segments (2 for the phi|mu vector of ph_mu_solve)
 * sums each segment so and adds the segment sums in order.  DNRM2 is the
 * square root of that sum of squares (MKL's overflow scaling changes
 * nothing for values far from the overflow range).  Padding widths that
 * differ can change only the sign of an all-zero sum.                    */
#define PW_LEAF 16
static double pw_rec(const double* x, const double* y, int64_t lo, int64_t len, int64_t n) {
  if (lo >= n) return 0.0;
  if (len <= PW_LEAF) {
    double a[PW_LEAF];
    for (int64_t i = 0; i < len; ++i) {
      int64_t q = lo + i;
      a[i] = q < n ? (y ? x[q] * y[q] : x[q]) : 0.0;
    }
    for (int64_t w = len; w > 1; w >>= 1)
      for (int64_t i = 0; i < w / 2; ++i) a[i] = a[2 * i] + a[2 * i + 1];
    return a[0];
  }
  int64_t h = len >> 1;
  double l = pw_rec(x, y, lo, h, n);
  double r = pw_rec(x, y, lo + h, h, n);
  return l + r;
}
/* the same tree with the subtrees above 2^15 elements as OpenMP tasks */
static double pw_task(const double* x, const double* y, int64_t lo, int64_t len, int64_t n) {
  if (lo >= n || len <= ((int64_t)1 << 15)) return pw_rec(x, y, lo, len, n);
  const int64_t h = len >> 1;
  double l = 0.0, r = 0.0;
#pragma omp task shared(l)
  l = pw_task(x, y, lo, h, n);
  r = pw_task(x, y, lo + h, h, n);
#pragma omp taskwait
  return l + r;
}
/* canonical sum of x[i]*y[i] (or of x[i] when y is NULL), i < n */
double oc_pairwise(int64_t n, const double* x, const double* y) {
  if (n <= 0) return 0.0;
  int64_t P = 1;
  while (P < n) P <<= 1;
  if (n <= 2 * PAR_MIN || omp_get_max_threads() == 1) return pw_rec(x, y, 0, P, n);
  double s = 0.0;
#pragma omp parallel
#pragma omp single
  s = pw_task(x, y, 0, P, n);
  return s;
}

void oc_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
int oc_set_initial_guess(oc_ctx* c, int32_t mode) {
  if (mode != 0 && mode != 1) return -1;
  c->guess = mode;
  return 0;
}
int32_t oc_get_threads(void) { return omp_get_max_threads(); }
static double seg_dot(int64_t n, int nseg, const double* x, const double* y) {
  int64_t seg = n / nseg;
  double s = oc_pairwise(seg, x, y);
  for (int k = 1; k < nseg; ++k) s = s + oc_pairwise(seg, x + (size_t)k * seg, y + (size_t)k * seg);
  return s;
}
static double ddot_s(int64_t n, int nseg, const double* x, const double* y) { return seg_dot(n, nseg, x, y); }
static double dnrm2_s(int64_t n, int nseg, const double* x) { return sqrt(seg_dot(n, nseg, x, x)); }
static void daxpy(int64_t n, double a, const double* x, double* y) {
  OMP_FOR(n > PAR_MIN)
  for (int64_t i = 0; i < n; ++i) y[i] = y[i] + a * x[i];
}

/* ---- packers: Atob_uv / btoA_uv  TP/matrix_tool.f90:5-75 -------------- */
static void atob_uv(const oc_ctx* c, const double* f1, const double* f2, double* b) {
  int nx = NX, ny = NY, nu = nxu(&c->p);
  const int big = (int64_t)nx * ny > PAR_MIN;
  OMP_FOR(big)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nu; ++i) b[i + (size_t)nu * j] = f1[UI(i, j)];
  size_t mn = (size_t)nu * ny;
  OMP_FOR(big)
  for (int j = 1; j < ny; ++j)
    for (int i = 0; i < nx; ++i) b[i + (size_t)nx * (j - 1) + mn] = f2[VI(i, j)];
}

static void btoa_uv(const oc_ctx* c, double* f1, double* f2, const double* b) {
  int nx = NX, ny = NY, nu = nxu(&c->p);
  const int big = (int64_t)nx * ny > PAR_MIN;
  OMP_FOR(big)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nu; ++i) f1[UI(i, j)] = b[i + (size_t)nu * j];
  size_t mn = (size_t)nu * ny;
  OMP_FOR(big)
  for (int j = 1; j < ny; ++j)
    for (int i = 0; i < nx; ++i) f2[VI(i, j)] = b[i + (size_t)nx * (j - 1) + mn];
  for (int i = 0; i < nu; ++i) {
    f1[UI(i, -1)] = -f1[UI(i, 0)];
    f1[UI(i, ny)] = -f1[UI(i, ny - 1)];
  }
  if (per_x(&c->p)) {
    xghost_u(c, f1);
  } else {
    for (int j = -1; j <= ny; ++j) {
      f1[UI(-1, j)] = f1[UI(1, j)];
      f1[UI(nx + 1, j)] = f1[UI(nx - 1, j)];
    }
  }
  for (int i = 0; i < nx; ++i) {
    f2[VI(i, 0)] = 0;
    f2[VI(i, ny)] = 0;
  }
  if (per_x(&c->p)) {
    xghost_v(c, f2);
  } else {
    for (int j = 0; j <= ny; ++j) {
      f2[VI(-1, j)] = f2[VI(0, j)];
      f2[VI(nx, j)] = f2[VI(nx - 1, j)];
    }
  }
}

/* Atob_ph_mu / btoA_ph_mu  TP/matrix_tool.f90:82-136 */
static void atob_ph_mu(const oc_ctx* c, const double* f1, const double* f2, double* b) {
  int nx = NX, ny = NY;
  size_t mn = (size_t)nx * ny;
  OMP_FOR(mn > PAR_MIN)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      b[i + (size_t)nx * j] = f1[CI(i, j)];
      b[i + (size_t)nx * j + mn] = f2[CI(i, j)];
    }
}

static void btoa_ph_mu(const oc_ctx* c, double* f1, double* f2, const double* b) {
  int nx = NX, ny = NY;
  size_t mn = (size_t)nx * ny;
  OMP_FOR(mn > PAR_MIN)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      f1[CI(i, j)] = b[i + (size_t)nx * j];
      f2[CI(i, j)] = b[i + (size_t)nx * j + mn];
    }
  for (int i = 0; i < nx; ++i) {
    f1[CI(i, -1)] = f1[CI(i, 0)];
    f1[CI(i, ny)] = f1[CI(i, ny - 1)];
    f2[CI(i, -1)] = f2[CI(i, 0)];
    f2[CI(i, ny)] = f2[CI(i, ny - 1)];
  }
  if (per_x(&c->p)) {
    xghost_c(c, f1);
    xghost_c(c, f2);
    return;
  }
  for (int j = -1; j <= ny; ++j) {
    f1[CI(-1, j)] = f1[CI(0, j)];
    f1[CI(nx, j)] = f1[CI(nx - 1, j)];
    f2[CI(-1, j)] = f2[CI(0, j)];
    f2[CI(nx, j)] = f2[CI(nx - 1, j)];
  }
}

/* Atob / btoA on (0:m,0:n) of a (-1:m+1,-1:n+1) array TP/matrix_tool.f90:142-173 */
static void atob_c(const oc_ctx* c, const double* f, double* b) {
  for (int j = 0; j < NY; ++j)
    for (int i = 0; i < NX; ++i) b[i + (size_t)NX * j] = f[CI(i, j)];
}
static void btoa_c(const oc_ctx* c, double* f, const double* b) {
  for (int j = 0; j < NY; ++j)
    for (int i = 0; i < NX; ++i) f[CI(i, j)] = b[i + (size_t)NX * j];
}

/* op_double_gradient_Neum TP/matrix_tool.f90:178-219 */
/* per: periodic in x (the interior expression with the neighbour index
 * wrapped), Neumann in y as the reference */
static void dgn_x(int32_t nx, int32_t ny, double dx2, double dy2, const double* rmu, const double* mo,
                  double* rhs, int per) {
#define GI(i, k) ((size_t)((i) + 1) + (size_t)(nx + 2) * (size_t)((k) + 1))
  size_t n = (size_t)(nx + 2) * (ny + 2);
  for (size_t q = 0; q < n; ++q) rhs[q] = 0.0;
  OMP_FOR(n > PAR_MIN)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double moL, moR;
      if (per) {
        const int il = i == 0 ? nx - 1 : i - 1, ir = i == nx - 1 ? 0 : i + 1;
        moL = (mo[GI(il, j)] + mo[GI(i, j)]) / 2.0;
        moR = (mo[GI(i, j)] + mo[GI(ir, j)]) / 2.0;
        rhs[GI(i, j)] = rhs[GI(i, j)] + 1.0 / dx2 * (moL * rmu[GI(il, j)] + moR * rmu[GI(ir, j)] -
                                                      (moL + moR) * rmu[GI(i, j)]);
      } else if (i == 0) {
        moL = (mo[GI(i, j)] + mo[GI(i, j)]) / 2.0;
        moR = (mo[GI(i, j)] + mo[GI(i + 1, j)]) / 2.0;
        rhs[GI(i, j)] = rhs[GI(i, j)] + 1.0 / dx2 * (moL * rmu[GI(i, j)] + moR * rmu[GI(i + 1, j)] -
                                                      (moL + moR) * rmu[GI(i, j)]);
      } else if (i == nx - 1) {
        moL = (mo[GI(i - 1, j)] + mo[GI(i, j)]) / 2.0;
        moR = (mo[GI(i, j)] + mo[GI(i, j)]) / 2.0;
        rhs[GI(i, j)] = rhs[GI(i, j)] + 1.0 / dx2 * (moL * rmu[GI(i - 1, j)] + moR * rmu[GI(i, j)] -
                                                      (moL + moR) * rmu[GI(i, j)]);
      } else {
        moL = (mo[GI(i - 1, j)] + mo[GI(i, j)]) / 2.0;
        moR = (mo[GI(i, j)] + mo[GI(i + 1, j)]) / 2.0;
        rhs[GI(i, j)] = rhs[GI(i, j)] + 1.0 / dx2 * (moL * rmu[GI(i - 1, j)] + moR * rmu[GI(i + 1, j)] -
                                                      (moL + moR) * rmu[GI(i, j)]);
      }
      if (j == 0) {
        moL = (mo[GI(i, j)] + mo[GI(i, j)]) / 2.0;
        moR = (mo[GI(i, j)] + mo[GI(i, j + 1)]) / 2.0;
        rhs[GI(i, j)] = rhs[GI(i, j)] + 1.0 / dy2 * (moL * rmu[GI(i, j)] + moR * rmu[GI(i, j + 1)] -
                                                      (moL + moR) * rmu[GI(i, j)]);
      } else if (j == ny - 1) {
        moL = (mo[GI(i, j - 1)] + mo[GI(i, j)]) / 2.0;
        moR = (mo[GI(i, j)] + mo[GI(i, j)]) / 2.0;
        rhs[GI(i, j)] = rhs[GI(i, j)] + 1.0 / dy2 * (moL * rmu[GI(i, j - 1)] + moR * rmu[GI(i, j)] -
                                                      (moL + moR) * rmu[GI(i, j)]);
      } else {
        moL = (mo[GI(i, j - 1)] + mo[GI(i, j)]) / 2.0;
        moR = (mo[GI(i, j)] + mo[GI(i, j + 1)]) / 2.0;
        rhs[GI(i, j)] = rhs[GI(i, j)] + 1.0 / dy2 * (moL * rmu[GI(i, j - 1)] + moR * rmu[GI(i, j + 1)] -
                                                      (moL + moR) * rmu[GI(i, j)]);
      }
    }
#undef GI
}
void oc_op_double_gradient_neum(int32_t nx, int32_t ny, double dx2, double dy2,
                                const double* rmu, const double* mo, double* rhs) {
  dgn_x(nx, ny, dx2, dy2, rmu, mo, rhs, 0);
}

/* ---- velocity operator ------------------------------------------------
 * MatrixVector_uv_CG  TP/VeloSolve_combined.f90:537-675   (bdf = 1)
 * MatrixVector_uv_2_CG TP/VeloSolve_2_combined.f90:536-676 (bdf = 2)   */
/* The MF variant (MF/VeloSolve_combined.f90, MF/VeloSolve_2_combined.f90)
 * has no lr_bd_type switch: its inflow and u side ghosts are always the
 * Dirichlet ones; its BDF2 penalty mask is the smooth bdr_ph (:580, :613)
 * instead of bdr_ph_penalty. */
static int is_mf(const drd_params* p) { return p->variant == DRD_VARIANT_MF; }
static int lr_dirichlet(const drd_params* p) { return p->lr_bd_type == 'd' || is_mf(p); }

static void velo_matvec(oc_ctx* c, int bdf, double alpha, const double* xx, double beta,
                        double* yy) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  double *f1 = c->w1, *f2 = c->w2, *f11 = c->w3, *f22 = c->w4;
  const double *rho = c->f[DRD_F_RHO], *eta = c->f[DRD_F_ETA];
  const double* chi = is_mf(p) ? c->f[DRD_F_BDR_PH] : c->f[DRD_F_BDR_PH_PENALTY];
  double dt = p->dt, re = p->re, dr = p->dx, dz = p->dz, dr2 = dr * dr, dz2 = dz * dz;
  btoa_uv(c, f1, f2, xx);
  if (per_x(p)) {
    /* btoa_uv left the periodic side copies */
  } else if (lr_dirichlet(p)) {
    for (int k = 0; k < ny; ++k) {
      f1[UI(-1, k)] = -f1[UI(1, k)];
      f1[UI(nx + 1, k)] = -f1[UI(nx - 1, k)];
    }
  } else {
    for (int k = -1; k <= ny; ++k) {
      f1[UI(-1, k)] = f1[UI(1, k)];
      f1[UI(nx + 1, k)] = f1[UI(nx - 1, k)];
    }
  }
  if (!per_x(p))
    for (int k = 0; k <= ny; ++k) {
      f2[VI(-1, k)] = f2[VI(0, k)];
      f2[VI(nx, k)] = f2[VI(nx - 1, k)];
    }
  for (int i = -1; i <= nx + 1; ++i) {
    f1[UI(i, -1)] = -f1[UI(i, 0)];
    f1[UI(i, ny)] = -f1[UI(i, ny - 1)];
  }
  for (int i = -1; i <= nx; ++i) {
    f2[VI(i, 0)] = 0.0;
    f2[VI(i, ny)] = 0.0;
  }
  const int big = (int64_t)nx * ny > PAR_MIN;
  /* u rows */
  const int nu = nxu(p);
  OMP_FOR(big)
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nu; ++i) {
      double r;
      if (bdf == 2)
        r = f1[UI(i, k)] * 1.5 * re * (rho[CI(i, k)] + rho[CI(i - 1, k)]) / 2.0 +
            f1[UI(i, k)] * p->penalty_k * (chi[CI(i, k)] + chi[CI(i - 1, k)]) / 2.0 * dt;
      else
        r = f1[UI(i, k)] * re * (rho[CI(i, k)] + rho[CI(i - 1, k)]) / 2.0;
      double cr = eta[CI(i, k)], cl = eta[CI(i - 1, k)];
      r = r - 2.0 * dt * (cr * f1[UI(i + 1, k)] + cl * f1[UI(i - 1, k)] - (cl + cr) * f1[UI(i, k)]) / dr2;
      cr = (eta[CI(i, k + 1)] + eta[CI(i - 1, k + 1)] + eta[CI(i, k)] + eta[CI(i - 1, k)]) / 4.0;
      cl = (eta[CI(i, k)] + eta[CI(i - 1, k)] + eta[CI(i, k - 1)] + eta[CI(i - 1, k - 1)]) / 4.0;
      r = r - dt * (cr * f1[UI(i, k + 1)] + cl * f1[UI(i, k - 1)] - (cl + cr) * f1[UI(i, k)]) / dz2;
      r = r - dt * (cr * (f2[VI(i, k + 1)] - f2[VI(i - 1, k + 1)]) - cl * (f2[VI(i, k)] - f2[VI(i - 1, k)])) / dr / dz;
      f11[UI(i, k)] = r;
    }
  /* v rows */
  OMP_FOR(big)
  for (int k = 1; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      double r;
      if (bdf == 2)
        r = f2[VI(i, k)] * 1.5 * re * (rho[CI(i, k)] + rho[CI(i, k - 1)]) / 2.0 +
            f2[VI(i, k)] * p->penalty_k * (chi[CI(i, k)] + chi[CI(i, k - 1)]) / 2.0 * dt;
      else
        r = f2[VI(i, k)] * re * (rho[CI(i, k)] + rho[CI(i, k - 1)]) / 2.0;
      double cr = eta[CI(i, k)], cl = eta[CI(i, k - 1)];
      r = r - 2.0 * dt * (cr * f2[VI(i, k + 1)] + cl * f2[VI(i, k - 1)] - (cl + cr) * f2[VI(i, k)]) / dz2;
      cr = (eta[CI(i + 1, k)] + eta[CI(i, k)] + eta[CI(i + 1, k - 1)] + eta[CI(i, k - 1)]) / 4.0;
      cl = (eta[CI(i, k)] + eta[CI(i - 1, k)] + eta[CI(i, k - 1)] + eta[CI(i - 1, k - 1)]) / 4.0;
      r = r - dt * (cr * f2[VI(i + 1, k)] + cl * f2[VI(i - 1, k)] - (cl + cr) * f2[VI(i, k)]) / dr2;
      r = r - dt * (cr * (f1[UI(i + 1, k)] - f1[UI(i + 1, k - 1)]) - cl * (f1[UI(i, k)] - f1[UI(i, k - 1)])) / dr / dz;
      f22[VI(i, k)] = r;
    }
  int64_t n = nuv(p);
  double* zz = c->w5;
  atob_uv(c, f11, f22, zz);
  OMP_FOR(big)
  for (int64_t q = 0; q < n; ++q) yy[q] = beta == 0.0 ? alpha * zz[q] : alpha * zz[q] + beta * yy[q];
}

/* PSolve_uv_CG (Jacobi) TP/VeloSolve_combined.f90:679-741 (mass
 * coefficient re), and the MF BDF2 form PSolve_uv_2_CG
 * MF/VeloSolve_2_combined.f90:662-711 (1.5 re; no penalty term, as there).
 * The v rows average rho over (i-1,j) as the reference does (:720). */
static void velo_jacobi_m(oc_ctx* c, double mcoef, double* yy, const double* xx) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  double *f1 = c->w1, *f2 = c->w2, *f11 = c->w3, *f22 = c->w4;
  const double *rho = c->f[DRD_F_RHO], *eta = c->f[DRD_F_ETA];
  double dt = p->dt, dr2 = p->dx * p->dx, dz2 = p->dz * p->dz;
  btoa_uv(c, f1, f2, xx);
  const int nu = nxu(p), per = per_x(p);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nu; ++i) {
      double tmp = mcoef * (rho[CI(i, j)] + rho[CI(i - 1, j)]) / 2.0;
      double cr = eta[CI(i, j)], cl = eta[CI(i - 1, j)];
      tmp = tmp + 2.0 * dt * (cl + cr) / dr2;
      cr = (eta[CI(i, j + 1)] + eta[CI(i - 1, j + 1)] + eta[CI(i, j)] + eta[CI(i - 1, j)]) / 4.0;
      cl = (eta[CI(i, j)] + eta[CI(i - 1, j)] + eta[CI(i, j - 1)] + eta[CI(i - 1, j - 1)]) / 4.0;
      if (j == 0)
        tmp = tmp + dt * (cl + (cl + cr)) / dz2;
      else if (j == ny - 1)
        tmp = tmp + dt * (cr + (cl + cr)) / dz2;
      else
        tmp = tmp + dt * (cl + cr) / dz2;
      f11[UI(i, j)] = f1[UI(i, j)] / tmp;
    }
  for (int j = 1; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double tmp = mcoef * (rho[CI(i, j)] + rho[CI(i - 1, j)]) / 2.0;
      double cr = eta[CI(i, j)], cl = eta[CI(i, j - 1)];
      tmp = tmp + 2.0 * dt * (cl + cr) / dz2;
      cr = (eta[CI(i + 1, j)] + eta[CI(i, j)] + eta[CI(i + 1, j - 1)] + eta[CI(i, j - 1)]) / 4.0;
      cl = (eta[CI(i, j)] + eta[CI(i - 1, j)] + eta[CI(i, j - 1)] + eta[CI(i - 1, j - 1)]) / 4.0;
      if (!per && i == 0)
        tmp = tmp + dt * cr / dr2;
      else if (!per && i == nx - 1)
        tmp = tmp + dt * cl / dr2;
      else
        tmp = tmp + dt * (cl + cr) / dr2;
      f22[VI(i, j)] = f2[VI(i, j)] / tmp;
    }
  atob_uv(c, f11, f22, yy);
}
static void velo_jacobi(oc_ctx* c, double* yy, const double* xx) { velo_jacobi_m(c, c->p.re, yy, xx); }

/* ---- Cahn-Hilliard operators ------------------------------------------ */
/* MATRIX_VECTOR_PH_MU_CG TP/ph_mu_solve.f90:203-299 */
static void phmu_matvec(oc_ctx* c, double alpha, const double* xx, double beta, double* yy) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  double *f1 = c->w1, *f2 = c->w2, *ff1 = c->w3, *ff2 = c->w4;
  const double* bu = c->f[DRD_F_BDR_PH_USED];
  double I = p->I_thickness, s = p->s_implicit, dr2 = p->dx * p->dx, dz2 = p->dz * p->dz;
  btoa_ph_mu(c, f1, f2, xx);
  dgn_x(nx, ny, dr2, dz2, f1, c->f[DRD_F_PH_M_INNER], ff1, per_x(p));
  const int big = (int64_t)nx * ny > PAR_MIN;
  OMP_FOR(big)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      ff1[CI(i, j)] = s * f1[CI(i, j)] / I - I * ff1[CI(i, j)] - f2[CI(i, j)] * bu[CI(i, j)];
  for (int i = 0; i < nx; ++i) {
    int j = 0;
    ff1[CI(i, j)] = ff1[CI(i, j)] + I / dz2 * (f1[CI(i, j - 1)] - c->coef_bot * f1[CI(i, j)]);
    j = ny - 1;
    ff1[CI(i, j)] = ff1[CI(i, j)] + I / dz2 * (f1[CI(i, j + 1)] - c->coef_top * f1[CI(i, j)]);
  }
  dgn_x(nx, ny, dr2, dz2, f2, c->f[DRD_F_PH_M_0], ff2, per_x(p));
  OMP_FOR(big)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      ff2[CI(i, j)] = f1[CI(i, j)] * (bu[CI(i, j)] + p->bdr_ph_used_adjust) - p->dt * p->xld * ff2[CI(i, j)];
  int64_t n = 2 * (int64_t)nx * ny;
  double* zz = c->w5;
  atob_ph_mu(c, ff1, ff2, zz);
  OMP_FOR(big)
  for (int64_t q = 0; q < n; ++q) yy[q] = beta == 0.0 ? alpha * zz[q] : alpha * zz[q] + beta * yy[q];
}

/* PSolve_PH_MU_CG (Jacobi) TP/ph_mu_solve.f90:304-356.  The x test reads
 * `j == nr-1` and the z test `i == 0` as in the reference (:323,:328).
 * Periodic x (no reference counterpart): the unit-mobility diagonal, 2/dr2
 * in x everywhere and 1/dz2 on the wall rows j = 0, ny-1.  */
static void phmu_jacobi(oc_ctx* c, double* yy, const double* xx) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  const int per = per_x(p);
  double *f1 = c->w1, *f2 = c->w2, *ff1 = c->w3, *ff2 = c->w4;
  double I = p->I_thickness, s = p->s_implicit, dr2 = p->dx * p->dx, dz2 = p->dz * p->dz;
  btoa_ph_mu(c, f1, f2, xx);
  const int big = (int64_t)nx * ny > PAR_MIN;
  OMP_FOR(big)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double tmpx = (!per && (i == 0 || j == nx - 1)) ? 1.0 / dr2 : 2.0 / dr2;
      double tmpz = ((per ? j == 0 : i == 0) || j == ny - 1) ? 1.0 / dz2 : 2.0 / dz2;
      ff1[CI(i, j)] = f1[CI(i, j)] / ((tmpx + tmpz) * I + s / I);
    }
  OMP_FOR(big)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double tmpx = (!per && (i == 0 || j == nx - 1)) ? 1.0 / dr2 : 2.0 / dr2;
      double tmpz = ((per ? j == 0 : i == 0) || j == ny - 1) ? 1.0 / dz2 : 2.0 / dz2;
      ff2[CI(i, j)] = f2[CI(i, j)] / (p->dt * p->xld * (tmpx + tmpz));
    }
  atob_ph_mu(c, ff1, ff2, yy);
}

/* 2x2 block Jacobi of the phi/mu operator (an addition without a reference
 * counterpart): per cell the exact diagonal block of MATRIX_VECTOR_PH_MU_CG
 * (TP/ph_mu_solve.f90:203-299), inverted in closed form, with a floor on
 * the mu-mu entry where the mobility M_0 vanishes.                        */
static double dgn_diag(const oc_ctx* c, double dx2, double dy2, const double* mo, int i, int j) {
  int nx = NX, ny = NY;
  const double m0 = mo[CI(i, j)];
  const int per = per_x(&c->p), il = i == 0 ? nx - 1 : i - 1, ir = i == nx - 1 ? 0 : i + 1;
  double dx = 0.0, dy = 0.0;
  if (per || i > 0) dx += (mo[CI(il, j)] + m0) / 2.0;
  if (per || i < nx - 1) dx += (m0 + mo[CI(ir, j)]) / 2.0;
  if (j > 0) dy += (mo[CI(i, j - 1)] + m0) / 2.0;
  if (j < ny - 1) dy += (m0 + mo[CI(i, j + 1)]) / 2.0;
  return -(dx / dx2 + dy / dy2);
}
static void phmu_block(oc_ctx* c, double* yy, const double* xx) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  const double *bu = c->f[DRD_F_BDR_PH_USED], *mi = c->f[DRD_F_PH_M_INNER], *m0 = c->f[DRD_F_PH_M_0];
  double I = p->I_thickness, s = p->s_implicit, dr2 = p->dx * p->dx, dz2 = p->dz * p->dz;
  int64_t C = (int64_t)nx * ny;
  const double floor22 = 1e-3 * p->dt * p->xld * (2.0 / dr2 + 2.0 / dz2);
  OMP_FOR(C > PAR_MIN)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t q = i + (size_t)nx * j;
      const double bv = bu[CI(i, j)];
      const double d11 = s / I - I * dgn_diag(c, dr2, dz2, mi, i, j) +
                         ((j == 0) ? I / dz2 * (1.0 - c->coef_bot) : 0.0) +
                         ((j == ny - 1) ? I / dz2 * (1.0 - c->coef_top) : 0.0);
      const double d12 = -bv;
      const double d21 = bv + p->bdr_ph_used_adjust;
      double d22 = -p->dt * p->xld * dgn_diag(c, dr2, dz2, m0, i, j);
      d22 = d22 > floor22 ? d22 : floor22;
      const double det = d11 * d22 - d12 * d21;
      const double r1 = xx[q], r2 = xx[C + q];
      yy[q] = (d22 * r1 - d12 * r2) / det;
      yy[C + q] = (-d21 * r1 + d11 * r2) / det;
    }
}

/* MATRIX_VECTOR_CG MF/PhSolve.f90:240-292 */
static void ph_matvec(oc_ctx* c, double alpha, const double* xx, double beta, double* yy) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  double *f1 = c->w1, *ff1 = c->w3, *ff2 = c->w4;
  double I = p->I_thickness, s = p->s_implicit, dr2 = p->dx * p->dx, dz2 = p->dz * p->dz;
  int64_t n = ncell(&c->p);
  for (int64_t q = 0; q < n; ++q) f1[q] = 0.0;
  btoa_c(c, f1, xx);
  oc_op_double_gradient_neum(nx, ny, dr2, dz2, f1, c->f[DRD_F_PH_M_0], ff1);
  for (int64_t q = 0; q < n; ++q) ff1[q] = -p->dt * p->xld * ff1[q];
  oc_op_double_gradient_neum(nx, ny, dr2, dz2, ff1, c->f[DRD_F_PH_M_INNER], ff2);
  for (int64_t q = 0; q < n; ++q) ff2[q] = s * ff1[q] / I - I * ff2[q];
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t q = i + (size_t)nx * j;
      const double z = f1[CI(i, j)] + ff2[CI(i, j)];
      yy[q] = beta == 0.0 ? alpha * z : alpha * z + beta * yy[q];
    }
}

/* ---- transforms -------------------------------------------------------- */
/* FFTW REDFT10 along one dimension: Y_k = 2 sum_j x_j cos(pi (j+1/2) k / n) */
static void dct2_lines(int n, int m, double* a, int stride, int ld, double* tmp, const double* tab) {
  for (int l = 0; l < m; ++l) {
    double* x = a + (size_t)l * ld;
    for (int k = 0; k < n; ++k) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += x[(size_t)j * stride] * tab[(size_t)k * n + j];
      tmp[k] = 2.0 * s;
    }
    for (int k = 0; k < n; ++k) x[(size_t)k * stride] = tmp[k];
  }
}
/* FFTW REDFT01: Y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (k+1/2) / n)        */
static void dct3_lines(int n, int m, double* a, int stride, int ld, double* tmp, const double* tab) {
  for (int l = 0; l < m; ++l) {
    double* x = a + (size_t)l * ld;
    for (int k = 0; k < n; ++k) {
      double s = 0.0;
      for (int j = 1; j < n; ++j) s += x[(size_t)j * stride] * tab[(size_t)j * n + k];
      tmp[k] = x[0] + 2.0 * s;
    }
    for (int k = 0; k < n; ++k) x[(size_t)k * stride] = tmp[k];
  }
}
static double* dct_table(int n) {
  /* tab[k*n+j] = cos(pi (2j+1) k / (2n)) with the argument reduced mod 4n */
  double* t = (double*)malloc((size_t)n * n * sizeof(double));
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j) {
      long long q = ((long long)(2 * j + 1) * k) % (4LL * n);
      t[(size_t)k * n + j] = cos(M_PI * (double)q / (2.0 * n));
    }
  return t;
}
void oc_dct2_2d(int32_t nx, int32_t ny, double* a) {
  int mx = nx > ny ? nx : ny;
  double* tmp = (double*)malloc((size_t)mx * sizeof(double));
  double* tx = dct_table(nx);
  double* ty = dct_table(ny);
  dct2_lines(nx, ny, a, 1, nx, tmp, tx);
  dct2_lines(ny, nx, a, nx, 1, tmp, ty);
  free(tmp); free(tx); free(ty);
}
void oc_dct3_2d(int32_t nx, int32_t ny, double* a) {
  int mx = nx > ny ? nx : ny;
  double* tmp = (double*)malloc((size_t)mx * sizeof(double));
  double* tx = dct_table(nx);
  double* ty = dct_table(ny);
  dct3_lines(nx, ny, a, 1, nx, tmp, tx);
  dct3_lines(ny, nx, a, nx, 1, tmp, ty);
  free(tmp); free(tx); free(ty);
}

/* FFT-factored DCTs for the pressure solve.  dct_fftw runs FFTW's REDFT10 /
 * REDFT01, whose factorisation (and so rounding) is FFTW's own.  The path
 * restates them as Makhoul's even/odd reordering plus a radix-2 Stockham FFT
 * for power-of-two lengths up to 4096, and as the defining sums over a
 * cosine table otherwise; the 2-D transform runs x then y forward and y
 * then x inverse.  Written out here operation for operation (same twiddle
 * tables, same expression order), so pressure solves, and with them whole
 * time steps, compare bit for bit; the O(n^2) definitions above check it. */
typedef struct { double x, y; } cpx;
typedef struct {
  int L, fft;
  cpx *tw, *rot;
  double* cs;
} line_plan;
static void line_plan_init(line_plan* p, int L) {
  p->L = L;
  p->fft = L >= 2 && L <= 4096 && (L & (L - 1)) == 0;
  int nt = L / 2 > 0 ? L / 2 : 1;
  p->tw = (cpx*)calloc((size_t)nt, sizeof(cpx));
  p->rot = (cpx*)calloc((size_t)L, sizeof(cpx));
  p->cs = (double*)calloc(4 * (size_t)L, sizeof(double));
  for (int k = 0; k < L / 2; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)L;
    p->tw[k].x = cos(a);
    p->tw[k].y = sin(a);
  }
  for (int k = 0; k < L; ++k) {
    const double a = -M_PI * (double)k / (2.0 * L);
    p->rot[k].x = cos(a);
    p->rot[k].y = sin(a);
  }
  for (int m = 0; m < 4 * L; ++m) p->cs[m] = cos(M_PI * (double)m / (2.0 * L));
}
static void line_plan_free(line_plan* p) {
  free(p->tw);
  free(p->rot);
  free(p->cs);
}
/* radix-2 Stockham, sign -1 forward / +1 inverse, unnormalised; returns
 * the buffer holding the result */
static cpx* stockham(cpx* a, cpx* b, int N, const cpx* tw, int sign) {
  const int half = N >> 1;
  for (int ns = 1; ns < N; ns <<= 1) {
    const int stride = N / (2 * ns);
    for (int j = 0; j < half; ++j) {
      const int k = j & (ns - 1);
      cpx w = tw[k * stride];
      if (sign > 0) w.y = -w.y;
      const cpx v0 = a[j], h = a[j + half];
      const cpx v1 = {h.x * w.x - h.y * w.y, h.x * w.y + h.y * w.x};
      const int o = ((j - k) << 1) + k;
      b[o].x = v0.x + v1.x;
      b[o].y = v0.y + v1.y;
      b[o + ns].x = v0.x - v1.x;
      b[o + ns].y = v0.y - v1.y;
    }
    cpx* t = a;
    a = b;
    b = t;
  }
  return a;
}
static int makhoul_ix(int j, int N) { return (j & 1) ? (N - 1 - (j >> 1)) : (j >> 1); }
/* REDFT10 of x[0..L) in place; A, B hold L complex values each */
static void line_dct2(const line_plan* p, double* x, cpx* A, cpx* B) {
  const int N = p->L;
  if (p->fft) {
    for (int j = 0; j < N; ++j) {
      A[makhoul_ix(j, N)].x = x[j];
      A[makhoul_ix(j, N)].y = 0.0;
    }
    const cpx* V = stockham(A, B, N, p->tw, -1);
    double* out = (double*)malloc((size_t)N * sizeof(double));
    for (int k = 0; k < N; ++k) {
      const cpx w = p->rot[k], v = V[k];
      out[k] = 2.0 * (w.x * v.x - w.y * v.y);
    }
    memcpy(x, out, (size_t)N * sizeof(double));
    free(out);
  } else {
    double* out = (double*)malloc((size_t)N * sizeof(double));
    for (int k = 0; k < N; ++k) {
      double s = 0.0;
      for (int j = 0; j < N; ++j) s += x[j] * p->cs[((long)(2 * j + 1) * k) % (4L * N)];
      out[k] = 2.0 * s;
    }
    memcpy(x, out, (size_t)N * sizeof(double));
    free(out);
  }
}
/* REDFT01 of y[0..L) in place */
static void line_dct3(const line_plan* p, double* y, cpx* A, cpx* B) {
  const int N = p->L;
  double* out = (double*)malloc((size_t)N * sizeof(double));
  if (p->fft) {
    for (int k = 0; k < N; ++k) {
      const double a = y[k];
      const double b = (k == 0) ? 0.0 : -y[N - k];
      const cpx w = p->rot[k];
      A[k].x = w.x * a + w.y * b;
      A[k].y = -w.y * a + w.x * b;
    }
    const cpx* v = stockham(A, B, N, p->tw, +1);
    for (int j = 0; j < N; ++j) out[j] = v[makhoul_ix(j, N)].x;
  } else {
    for (int k = 0; k < N; ++k) {
      double s = 0.0;
      for (int j = 1; j < N; ++j) s += y[j] * p->cs[((long)j * (2 * k + 1)) % (4L * N)];
      out[k] = y[0] + 2.0 * s;
    }
  }
  memcpy(y, out, (size_t)N * sizeof(double));
  free(out);
}
/* a (nx x ny, x fastest): forward x then y; pre != NULL: divide by
 * pre[kx + (nx+1) ky] and transform back, y then x.  inverse_only: the
 * inverse pair alone. */
static void dct_2d_fast(int nx, int ny, double* a, const double* pre, int mode) {
  line_plan px, py;
  line_plan_init(&px, nx);
  line_plan_init(&py, ny);
  int mx = nx > ny ? nx : ny;
  cpx* A = (cpx*)malloc((size_t)mx * sizeof(cpx));
  cpx* B = (cpx*)malloc((size_t)mx * sizeof(cpx));
  double* col = (double*)malloc((size_t)ny * sizeof(double));
  const int fwd = mode != 1, inv = mode != 0;
  if (fwd)
    for (int k = 0; k < ny; ++k) line_dct2(&px, a + (size_t)nx * k, A, B);
  for (int i = 0; i < nx; ++i) {
    for (int k = 0; k < ny; ++k) col[k] = a[i + (size_t)nx * k];
    if (fwd) line_dct2(&py, col, A, B);
    if (pre)
      for (int k = 0; k < ny; ++k) col[k] = col[k] / pre[i + (size_t)(nx + 1) * k];
    if (inv) line_dct3(&py, col, A, B);
    for (int k = 0; k < ny; ++k) a[i + (size_t)nx * k] = col[k];
  }
  if (inv)
    for (int k = 0; k < ny; ++k) line_dct3(&px, a + (size_t)nx * k, A, B);
  free(A);
  free(B);
  free(col);
  line_plan_free(&px);
  line_plan_free(&py);
}
/* Real DFT of x[0..L) (L a power of two) in FFTW's halfcomplex order
 * (R2HC): r_k = Re X_k for k <= L/2 and r_{L-k} = Im X_k for 0 < k < L/2,
 * X_k = sum_j x_j exp(-2 pi i jk / L), as the complex Stockham FFT of
 * (x, 0). */
static void line_r2hc(const line_plan* p, double* x, cpx* A, cpx* B) {
  const int N = p->L;
  for (int j = 0; j < N; ++j) {
    A[j].x = x[j];
    A[j].y = 0.0;
  }
  const cpx* V = stockham(A, B, N, p->tw, -1);
  for (int k = 0; k <= N / 2; ++k) x[k] = V[k].x;
  for (int k = 1; k < N / 2; ++k) x[N - k] = V[k].y;
}
/* HC2R, the unnormalised inverse: x_j = sum_k X_k exp(+2 pi i jk / L) of
 * the Hermitian X the halfcomplex r holds */
static void line_hc2r(const line_plan* p, double* r, cpx* A, cpx* B) {
  const int N = p->L, h = N / 2;
  A[0].x = r[0];
  A[0].y = 0.0;
  for (int k = 1; k < h; ++k) {
    A[k].x = r[k];
    A[k].y = r[N - k];
    A[N - k].x = r[k];
    A[N - k].y = -r[N - k];
  }
  A[h].x = r[h];
  A[h].y = 0.0;
  const cpx* v = stockham(A, B, N, p->tw, +1);
  for (int j = 0; j < N; ++j) r[j] = v[j].x;
}
/* The pressure solve of a periodic-x, Neumann-y box: R2HC along x, REDFT10
 * along y, / pre[m + (nx+1) ky], REDFT01 along y, HC2R along x; the result
 * is 2 nx ny times the solution (the periodic counterpart of
 * TP/dct_fftw.f90:13-31; nx a power of two). */
static void per_solve_2d(int nx, int ny, double* a, const double* pre) {
  line_plan px, py;
  line_plan_init(&px, nx);
  line_plan_init(&py, ny);
  int mx = nx > ny ? nx : ny;
  cpx* A = (cpx*)malloc((size_t)mx * sizeof(cpx));
  cpx* B = (cpx*)malloc((size_t)mx * sizeof(cpx));
  double* col = (double*)malloc((size_t)ny * sizeof(double));
  for (int k = 0; k < ny; ++k) line_r2hc(&px, a + (size_t)nx * k, A, B);
  for (int i = 0; i < nx; ++i) {
    for (int k = 0; k < ny; ++k) col[k] = a[i + (size_t)nx * k];
    line_dct2(&py, col, A, B);
    for (int k = 0; k < ny; ++k) col[k] = col[k] / pre[i + (size_t)(nx + 1) * k];
    line_dct3(&py, col, A, B);
    for (int k = 0; k < ny; ++k) a[i + (size_t)nx * k] = col[k];
  }
  for (int k = 0; k < ny; ++k) line_hc2r(&px, a + (size_t)nx * k, A, B);
  free(A);
  free(B);
  free(col);
  line_plan_free(&px);
  line_plan_free(&py);
}
/* exported for tests: the 1-D halfcomplex transforms (dir 0 R2HC, 1 HC2R) */
void oc_rfft_line(int32_t n, double* x, int32_t dir) {
  line_plan pl;
  line_plan_init(&pl, n);
  cpx* A = (cpx*)malloc((size_t)n * sizeof(cpx));
  cpx* B = (cpx*)malloc((size_t)n * sizeof(cpx));
  if (dir == 0)
    line_r2hc(&pl, x, A, B);
  else
    line_hc2r(&pl, x, A, B);
  free(A);
  free(B);
  line_plan_free(&pl);
}

/* exported for tests: mode 0 = REDFT10 2-D, 1 = REDFT01 2-D (y then x) */
void oc_dct_2d_fast(int32_t nx, int32_t ny, double* a, int32_t mode) { dct_2d_fast(nx, ny, a, NULL, mode); }

/* VCOST LIB/vfftpack.f:643-704, by its definition:
 * X(i) = [X(1) + (-1)^(i-1) X(N) + sum_{k=2}^{N-1} 2 X(k) cos((k-1)(i-1) pi/(N-1))] / sqrt(2(N-1)) */
void oc_vcost(int32_t m, int32_t n, double* x, int32_t ld) {
  double* tmp = (double*)malloc((size_t)n * sizeof(double));
  double sc = 1.0 / sqrt(2.0 * (n - 1));
  for (int l = 0; l < m; ++l) {
    double* s = x + (size_t)l * ld;
    for (int i = 0; i < n; ++i) {
      double acc = s[0] + ((i & 1) ? -s[n - 1] : s[n - 1]);
      for (int k = 1; k < n - 1; ++k) {
        long long q = ((long long)k * i) % (2LL * (n - 1));
        acc += 2.0 * s[k] * cos(M_PI * (double)q / (double)(n - 1));
      }
      tmp[i] = acc * sc;
    }
    for (int i = 0; i < n; ++i) s[i] = tmp[i];
  }
  free(tmp);
}

/* DCT TP/DCT.f90:1-57.  fO,fN (-1:nr+1,-1:nz+1); cof (0:nr,0:nz).  The
 * VCOST passes along z run over the nr+1 columns, those along x over the
 * nz+1 rows (the reference transposes through `temp` to do the latter). */
void oc_dct_solve(int32_t nr, int32_t nz, double* fO, double* fN, const double* cof) {
  int ld = nr + 3;
#define DI(i, k) ((size_t)((i) + 1) + (size_t)ld * (size_t)((k) + 1))
  int n1 = nr + 1, n2 = nz + 1;
  double* col = (double*)malloc((size_t)(n1 > n2 ? n1 : n2) * sizeof(double));
  for (int pass = 0; pass < 2; ++pass) {
    /* along z for each x */
    for (int i = 0; i <= nr; ++i) {
      for (int k = 0; k <= nz; ++k) col[k] = fO[DI(i, k)];
      oc_vcost(1, n2, col, n2);
      for (int k = 0; k <= nz; ++k) fO[DI(i, k)] = col[k];
    }
    /* along x for each z */
    for (int k = 0; k <= nz; ++k) {
      for (int i = 0; i <= nr; ++i) col[i] = fO[DI(i, k)];
      oc_vcost(1, n1, col, n1);
      for (int i = 0; i <= nr; ++i) fO[DI(i, k)] = col[i];
    }
    if (pass == 0)
      for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nr; ++i) fO[DI(i, k)] = fO[DI(i, k)] / cof[i + (size_t)(nr + 1) * k];
  }
  for (int k = 0; k <= nz; ++k)
    for (int i = 0; i <= nr; ++i) fN[DI(i, k)] = fO[DI(i, k)];
  for (int i = 0; i <= nr; ++i) {
    fN[DI(i, -1)] = fN[DI(i, 1)];
    fN[DI(i, nz + 1)] = fN[DI(i, nz - 1)];
  }
  for (int k = 0; k <= nz; ++k) {
    fN[DI(-1, k)] = fN[DI(1, k)];
    fN[DI(nr + 1, k)] = fN[DI(nr - 1, k)];
  }
  free(col);
#undef DI
}

/* rfords TP/setup.f90:327-354 (Thomas algorithm, 1-based arrays)       */
void oc_rfords(int32_t n, double* ff, const double* aa, const double* bb, const double* cc) {
  double* f = (double*)malloc((size_t)n * sizeof(double));
  double* b = (double*)malloc((size_t)n * sizeof(double));
  for (int j = 0; j < n; ++j) {
    f[j] = ff[j];
    b[j] = bb[j];
  }
  for (int j = 1; j < n; ++j) {
    double r = -aa[j] / b[j - 1];
    b[j] = b[j] + r * cc[j - 1];
    f[j] = f[j] + r * f[j - 1];
  }
  if (fabs(b[n - 1]) <= 1.0e-12)
    ff[n - 1] = 0.0;
  else
    ff[n - 1] = f[n - 1] / b[n - 1];
  for (int j = n - 2; j >= 0; --j) ff[j] = (f[j] - cc[j] * ff[j + 1]) / b[j];
  free(f);
  free(b);
}

/* ESOLP_x TP/ESOLP.f90:1-82: VCOST along x per row, Thomas along z per
 * x-mode, VCOST along x again. */
void oc_esolp_x(int32_t nr, int32_t nz, const double* pa, const double* pb, const double* pc,
                const double* b, double* p) {
  int ld = nr + 3;
#define DI(i, k) ((size_t)((i) + 1) + (size_t)ld * (size_t)((k) + 1))
  int n1 = nr + 1, n2 = nz + 1;
  /* bb(i,k) = b(k,i): rows of b become rows of work arrays */
  double* w = (double*)malloc((size_t)n1 * n2 * sizeof(double)); /* w[x + n1*z] */
  double* line = (double*)malloc((size_t)(n1 > n2 ? n1 : n2) * sizeof(double));
  for (int z = 0; z <= nz; ++z)
    for (int x = 0; x <= nr; ++x) w[x + (size_t)n1 * z] = b[DI(x, z)];
  oc_vcost(n2, n1, w, n1);
  double *wa = (double*)malloc((size_t)n2 * sizeof(double)), *wb = (double*)malloc((size_t)n2 * sizeof(double)),
         *wc = (double*)malloc((size_t)n2 * sizeof(double));
  for (int k = 0; k <= nr; ++k) {
    for (int i = 0; i < n2; ++i) {
      line[i] = w[k + (size_t)n1 * i];
      wa[i] = pa[i + (size_t)n2 * k];
      wb[i] = pb[i + (size_t)n2 * k];
      wc[i] = pc[i + (size_t)n2 * k];
    }
    oc_rfords(n2, line, wa, wb, wc);
    for (int i = 0; i < n2; ++i) w[k + (size_t)n1 * i] = line[i];
  }
  oc_vcost(n2, n1, w, n1);
  for (int z = 0; z <= nz; ++z)
    for (int x = 0; x <= nr; ++x) p[DI(x, z)] = w[x + (size_t)n1 * z];
  free(w); free(line); free(wa); free(wb); free(wc);
#undef DI
}

/* PSolve_CG MF/PhSolve.f90:296-313: DCT(nr-1,nz-1,...,Ph_Lcc) */
static void ph_psolve(oc_ctx* c, double* yy, const double* xx) {
  double *f1 = c->w1, *f2 = c->w2;
  int64_t n = ncell(&c->p);
  for (int64_t q = 0; q < n; ++q) f1[q] = 0.0;
  btoa_c(c, f1, xx);
  oc_dct_solve(NX - 1, NY - 1, f1, f2, c->ph_lcc);
  atob_c(c, f2, yy);
}

static void phori_matvec(oc_ctx* c, double alpha, const double* xx, double beta, double* yy);
int oc_matvec(oc_ctx* c, int32_t op, double alpha, const double* x, double beta, double* y) {
  switch (op) {
    case DRD_OP_VELO1: velo_matvec(c, 1, alpha, x, beta, y); return 0;
    case DRD_OP_VELO2: velo_matvec(c, 2, alpha, x, beta, y); return 0;
    case DRD_OP_PHMU: phmu_matvec(c, alpha, x, beta, y); return 0;
    case DRD_OP_PH: ph_matvec(c, alpha, x, beta, y); return 0;
    case DRD_OP_PHORI: phori_matvec(c, alpha, x, beta, y); return 0;
  }
  return DRD_E_ARG;
}

static int64_t op_size(const oc_ctx* c, int32_t op) {
  switch (op) {
    case DRD_OP_VELO1:
    case DRD_OP_VELO2: return nuv(&c->p);
    case DRD_OP_PHMU: return 2 * (int64_t)c->p.nx * c->p.ny;
    case DRD_OP_PH:
    case DRD_OP_PHORI: return (int64_t)c->p.nx * c->p.ny;
  }
  return -1;
}

/* segments of the operator's Krylov vectors (the phi|mu vector is two) */
static int op_nseg(int32_t op) { return op == DRD_OP_PHMU ? 2 : 1; }

/* PSolve_uv_CG of PD/VeloSolve_combined.f90:697-807: ESOLP_x on the u block
 * (nr, nz-1; tables ua..uc) and on the v block (nr-1, nz-2; va..vc), the v
 * field (-1:nr,0:nz) read as (-1:nr,-1:nz-1). */
static void velo_esolp(oc_ctx* c, double* yy, const double* xx) {
  double *f1 = c->w1, *f2 = c->w2;
  btoa_uv(c, f1, f2, xx);
  oc_esolp_x(NX, NY - 1, c->ua, c->ub, c->uc, f1, f1);
  oc_esolp_x(NX - 1, NY - 2, c->va, c->vb, c->vc, f2, f2);
  atob_uv(c, f1, f2, yy);
}

int oc_psolve(oc_ctx* c, int32_t prec, int32_t op, double* x_out, const double* b_in) {
  int64_t n = op_size(c, op);
  if (n < 0) return DRD_E_ARG;
  if (per_x(&c->p) && (prec == DRD_PS_VELO_ESOLP || prec == DRD_PS_PH_DCT)) return DRD_E_UNSUPPORTED;
  switch (prec) {
    case DRD_PS_VELO_ESOLP: velo_esolp(c, x_out, b_in); return 0;
    case DRD_PS_VELO1_JACOBI: velo_jacobi(c, x_out, b_in); return 0;
    case DRD_PS_VELO2_JACOBI: velo_jacobi_m(c, 1.5 * c->p.re, x_out, b_in); return 0;
    case DRD_PS_IDENTITY: memcpy(x_out, b_in, (size_t)n * sizeof(double)); return 0;
    case DRD_PS_PHMU_JACOBI: phmu_jacobi(c, x_out, b_in); return 0;
    case DRD_PS_PHMU_BLOCK: phmu_block(c, x_out, b_in); return 0;
    case DRD_PS_PH_DCT: ph_psolve(c, x_out, b_in); return 0;
  }
  return DRD_E_UNSUPPORTED;
}

/* ---- BiCGSTAB ----------------------------------------------------------
 * LIB/BiCGSTAB.f:153-217 driving LIB/BiCGSTABREVCOM.f:150-418 with
 * LIB/STOPTEST2.f and LIB/getbreak.f, restated as one loop.  The control
 * flow, tests and reported ITER/RESID/INFO follow the reverse-communication
 * original exactly:
 *   - r0 = b - A x0 only if ||x0|| != 0; then exit if ||r0|| <= tol
 *     (absolute), leaving ITER and RESID as passed in   (REVCOM:229-249)
 *   - breakdown tolerance eps^2 with eps = DLAMCH('E') = 2^-53 (getbreak.f)
 *   - early exit when ||s|| <= tol (absolute)            (REVCOM:311-316)
 *   - convergence when ||r||/||b|| <= tol (STOPTEST2.f:13-16)
 *   - INFO = 1 at ITER == MAXIT, -10 (rho) / -11 (omega) breakdowns.    */
typedef void (*mv_fn)(void* u, double alpha, const double* x, double beta, double* y);
typedef void (*ps_fn)(void* u, double* x_out, const double* b_in);

static int bicgstab_core(int64_t n, int nseg, const double* b, double* x, double* work, int32_t* iter,
                         double* resid, mv_fn mv, ps_fn ps, void* u) {
#define dnrm2(n, x) dnrm2_s(n, nseg, x)
#define ddot(n, x, y) ddot_s(n, nseg, x, y)
  if (n < 0) return -1;
  if (*iter <= 0) return -3;
  int32_t maxit = *iter;
  double tol = *resid;
  double *R = work, *RT = work + n, *P = work + 2 * n, *V = work + 3 * n, *T = work + 4 * n,
         *PH = work + 5 * n, *SH = work + 6 * n, *S = R;
  const double eps = ldexp(1.0, -53);
  const double breaktol = eps * eps;
  double rho = 0, rho1 = 1, alpha = 0, omega = 1, bnrm2;
  memcpy(R, b, (size_t)n * sizeof(double));
  if (dnrm2(n, x) != 0.0) {
    mv(u, -1.0, x, 1.0, R);
    if (dnrm2(n, R) <= tol) return 0;
  }
  memcpy(RT, R, (size_t)n * sizeof(double));
  bnrm2 = dnrm2(n, b);
  if (bnrm2 == 0.0) bnrm2 = 1.0;
  int32_t it = 0;
  for (;;) {
    it = it + 1;
    *iter = it;
    rho = ddot(n, RT, R);
    if (fabs(rho) < breaktol) return -10;
    if (it > 1) {
      double beta = (rho / rho1) * (alpha / omega);
      daxpy(n, -omega, V, P);
      OMP_FOR(n > PAR_MIN)
      for (int64_t q = 0; q < n; ++q) P[q] = beta * P[q];
      daxpy(n, 1.0, R, P);
    } else {
      memcpy(P, R, (size_t)n * sizeof(double));
    }
    ps(u, PH, P);
    mv(u, 1.0, PH, 0.0, V);
    alpha = rho / ddot(n, RT, V);
    daxpy(n, -alpha, V, R);
    /* S aliases R (DCOPY of a column onto itself) */
    double snrm = dnrm2(n, S);
    if (snrm <= tol) {
      daxpy(n, alpha, PH, x);
      *resid = snrm / bnrm2;
      return 0;
    }
    ps(u, SH, S);
    mv(u, 1.0, SH, 0.0, T);
    omega = ddot(n, T, S) / ddot(n, T, T);
    daxpy(n, alpha, PH, x);
    daxpy(n, omega, SH, x);
    daxpy(n, -omega, T, R);
    *resid = dnrm2(n, R) / bnrm2;
    if (*resid <= tol) return 0;
    if (it == maxit) return 1;
    if (fabs(omega) < breaktol) return -11;
    rho1 = rho;
  }
#undef dnrm2
#undef ddot
}

typedef struct {
  oc_ctx* c;
  int32_t op, prec;
} op_bind;
static void mv_bound(void* u, double alpha, const double* x, double beta, double* y) {
  op_bind* b = (op_bind*)u;
  oc_matvec(b->c, b->op, alpha, x, beta, y);
}
static void ps_bound(void* u, double* xo, const double* bi) {
  op_bind* b = (op_bind*)u;
  oc_psolve(b->c, b->prec, b->op, xo, bi);
}

int oc_bicgstab(oc_ctx* c, int32_t op, int32_t prec, int32_t n, const double* b, double* x,
                int32_t* iter, double* resid, int32_t* info) {
  if (n != op_size(c, op)) return DRD_E_ARG;
  op_bind bd = {c, op, prec};
  *info = bicgstab_core(n, op_nseg(op), b, x, c->work, iter, resid, mv_bound, ps_bound, &bd);
  return 0;
}

typedef struct {
  int32_t n;
  const double* a;
  int32_t jacobi;
} dense_bind;
static void mv_dense(void* u, double alpha, const double* x, double beta, double* y) {
  dense_bind* d = (dense_bind*)u;
  for (int i = 0; i < d->n; ++i) {
    double s = 0;
    for (int j = 0; j < d->n; ++j) s += d->a[i + (size_t)d->n * j] * x[j];
    y[i] = beta == 0.0 ? alpha * s : alpha * s + beta * y[i];
  }
}
static void ps_dense(void* u, double* xo, const double* bi) {
  dense_bind* d = (dense_bind*)u;
  for (int i = 0; i < d->n; ++i) xo[i] = d->jacobi ? bi[i] / d->a[i + (size_t)d->n * i] : bi[i];
}
int oc_bicgstab_dense(int32_t n, const double* a, int32_t jacobi, const double* b, double* x,
                      int32_t* iter, double* resid, int32_t* info) {
  double* work = (double*)calloc((size_t)n * 7, sizeof(double));
  dense_bind d = {n, a, jacobi};
  *info = bicgstab_core(n, 1, b, x, work, iter, resid, mv_dense, ps_dense, &d);
  free(work);
  return 0;
}

/* ---- FGMRES(m) -----------------------------------------------------------
 * Restarted flexible GMRES with right preconditioning, Saad, SIAM J. Sci.
 * Comput. 14(2) 1993, Algorithm 2.2, with modified Gram-Schmidt and Givens
 * rotations.  LIB/fgmres.f wraps a closed library routine of this form (its
 * call sites are commented out); this is written from the published
 * algorithm.  Stops when the Arnoldi residual estimate |g_{j+1}| <=
 * tol*||b||; the solution is then formed and its true residual checked, and
 * the method restarts if that check fails and iterations remain.
 * ITER = inner iterations; RESID = true ||b-Ax||/||b||; INFO 0 converged,
 * 1 maximum iterations. */
static int fgmres_core(int64_t n, int nseg, const double* b, double* x, int32_t m, int32_t* iter, double* resid,
                       mv_fn mv, ps_fn ps, void* u) {
#define dnrm2(n, x) dnrm2_s(n, nseg, x)
#define ddot(n, x, y) ddot_s(n, nseg, x, y)
  int32_t maxit = *iter;
  double tol = *resid;
  double* V = (double*)malloc((size_t)n * (m + 1) * sizeof(double));
  double* Z = (double*)malloc((size_t)n * m * sizeof(double));
  double* H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double*)calloc(m, sizeof(double)), *sn = (double*)calloc(m, sizeof(double));
  double *g = (double*)calloc(m + 1, sizeof(double)), *y = (double*)calloc(m, sizeof(double));
  double* r = (double*)malloc((size_t)n * sizeof(double));
  double bn = dnrm2(n, b);
  int info = 1, it = 0;
  if (bn == 0.0) {
    for (int64_t q = 0; q < n; ++q) x[q] = 0.0;
    *iter = 0;
    *resid = 0.0;
    info = 0;
    goto done;
  }
  for (;;) {
    memcpy(r, b, (size_t)n * sizeof(double));
    mv(u, -1.0, x, 1.0, r);
    double beta = dnrm2(n, r);
    *resid = beta / bn;
    if (*resid <= tol) {
      info = 0;
      break;
    }
    if (it >= maxit) break;
    for (int64_t q = 0; q < n; ++q) V[q] = r[q] / beta;
    for (int k = 0; k <= m; ++k) g[k] = 0.0;
    g[0] = beta;
    int j, jdone = 0;
    for (j = 0; j < m && it < maxit; ++j) {
      double* vj = V + (size_t)j * n;
      double* zj = Z + (size_t)j * n;
      double* w = V + (size_t)(j + 1) * n;
      ps(u, zj, vj);
      mv(u, 1.0, zj, 0.0, w);
      for (int i = 0; i <= j; ++i) {
        double h = ddot(n, w, V + (size_t)i * n);
        H[i + (size_t)(m + 1) * j] = h;
        daxpy(n, -h, V + (size_t)i * n, w);
      }
      double hn = dnrm2(n, w);
      H[j + 1 + (size_t)(m + 1) * j] = hn;
      if (hn != 0.0)
        for (int64_t q = 0; q < n; ++q) w[q] = w[q] / hn;
      for (int i = 0; i < j; ++i) {  /* previous rotations */
        double* hc = H + (size_t)(m + 1) * j;
        double t1 = cs[i] * hc[i] + sn[i] * hc[i + 1];
        hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
        hc[i] = t1;
      }
      double* hc = H + (size_t)(m + 1) * j;
      double den = hypot(hc[j], hc[j + 1]);
      cs[j] = den == 0.0 ? 1.0 : hc[j] / den;
      sn[j] = den == 0.0 ? 0.0 : hc[j + 1] / den;
      hc[j] = den;
      hc[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++it;
      jdone = j + 1;
      if (fabs(g[j + 1]) <= tol * bn || hn == 0.0) break;
    }
    /* y = R^-1 g, x += Z y */
    for (int i = jdone - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < jdone; ++k) s -= H[i + (size_t)(m + 1) * k] * y[k];
      y[i] = s / H[i + (size_t)(m + 1) * i];
    }
    for (int i = 0; i < jdone; ++i) daxpy(n, y[i], Z + (size_t)i * n, x);
  }
  *iter = it;
done:
  free(V); free(Z); free(H); free(cs); free(sn); free(g); free(y); free(r);
  return info;
#undef dnrm2
#undef ddot
}

int oc_fgmres(oc_ctx* c, int32_t op, int32_t prec, int32_t n, int32_t restart, const double* b, double* x,
              int32_t* iter, double* resid, int32_t* info) {
  if (n != op_size(c, op) || restart < 1) return DRD_E_ARG;
  op_bind bd = {c, op, prec};
  *info = fgmres_core(n, op_nseg(op), b, x, restart, iter, resid, mv_bound, ps_bound, &bd);
  return 0;
}

int oc_fgmres_dense(int32_t n, const double* a, int32_t jacobi, int32_t restart, const double* b, double* x,
                    int32_t* iter, double* resid, int32_t* info) {
  dense_bind d = {n, a, jacobi};
  *info = fgmres_core(n, 1, b, x, restart, iter, resid, mv_dense, ps_dense, &d);
  return 0;
}

/* true residual ||b - A x|| / ||b|| (an addition; FP64) */
static double true_resid(oc_ctx* c, int32_t op, const double* b, const double* x) {
  int64_t n = op_size(c, op);
  double* r = (double*)malloc((size_t)n * sizeof(double));
  memcpy(r, b, (size_t)n * sizeof(double));
  oc_matvec(c, op, -1.0, x, 1.0, r);
  double bn = dnrm2_s(n, op_nseg(op), b);
  double v = dnrm2_s(n, op_nseg(op), r) / (bn == 0 ? 1.0 : bn);
  free(r);
  return v;
}

/* ---- velocity solve ----------------------------------------------------
 * RHS of VeloSolve_combined TP/VeloSolve_combined.f90:48-396 (bdf 1) and
 * VeloSolve_2_combined TP/VeloSolve_2_combined.f90:62-400 (bdf 2).      */
int oc_velo_rhs(oc_ctx* c, int32_t bdf, const double* psi1, const double* psi2,
                const double* u_pre, const double* v_pre, const double* ph_after,
                const double* rmu_after, double t, const double* pf1, const double* pf2,
                double* rhs) {
  const drd_params* p = &c->p;
  if (p->xlsB >= 1.0e-8) return DRD_E_UNSUPPORTED; /* GNBC branch is commented out */
  int nx = NX, ny = NY, nu_ = nxu(p);
  const double *u = c->f[DRD_F_U], *v = c->f[DRD_F_V], *pp = c->f[DRD_F_P];
  const double *rho = c->f[DRD_F_RHO], *eta = c->f[DRD_F_ETA], *bu = c->f[DRD_F_BDR_PH_USED];
  double dt = p->dt, re = p->re, ca = p->ca, dr = p->dx, dz = p->dz, dr2 = dr * dr, dz2 = dz * dz;
  double *fu = c->w1, *fv = c->w2, *ut = c->w3, *vt = c->w4, *tp = c->w6;
  for (int64_t q = 0; q < nufield(p); ++q) fu[q] = 0.0;
  for (int64_t q = 0; q < nvfield(p); ++q) fv[q] = 0.0;
  /* ca * mu * grad(phi * bdr_ph_used) :62-97 */
  if (is_mf(p)) { /* MF/VeloSolve_2_combined.f90:64-65, :78-79: the jump of phi, times 1 - mean bdr_ph */
    const double* bd = c->f[DRD_F_BDR_PH];
    for (int k = 0; k < ny; ++k)
      for (int i = 0; i < nu_; ++i)
        fu[UI(i, k)] = ca * (rmu_after[CI(i, k)] + rmu_after[CI(i - 1, k)]) / 2.0 *
                       (ph_after[CI(i, k)] - ph_after[CI(i - 1, k)]) / dr * (1.0 - (bd[CI(i, k)] + bd[CI(i - 1, k)]) / 2.0);
    for (int k = 1; k < ny; ++k)
      for (int i = 0; i < nx; ++i)
        fv[VI(i, k)] = ca * (rmu_after[CI(i, k)] + rmu_after[CI(i, k - 1)]) / 2.0 *
                       (ph_after[CI(i, k)] - ph_after[CI(i, k - 1)]) / dz * (1.0 - (bd[CI(i, k)] + bd[CI(i, k - 1)]) / 2.0);
  } else {
    for (int k = 0; k < ny; ++k)
      for (int i = 0; i < nu_; ++i)
        fu[UI(i, k)] = ca * (rmu_after[CI(i, k)] + rmu_after[CI(i - 1, k)]) / 2.0 *
                       (ph_after[CI(i, k)] * bu[CI(i, k)] - ph_after[CI(i - 1, k)] * bu[CI(i - 1, k)]) / dr;
    for (int k = 1; k < ny; ++k)
      for (int i = 0; i < nx; ++i)
        fv[VI(i, k)] = ca * (rmu_after[CI(i, k)] + rmu_after[CI(i, k - 1)]) / 2.0 *
                       (ph_after[CI(i, k)] * bu[CI(i, k)] - ph_after[CI(i, k - 1)] * bu[CI(i, k - 1)]) / dz;
  }
  /* convection re*rho*(u.grad)u with u (bdf1) or 2u - u_pre (bdf2) :108-124 */
  if (bdf == 2) {
    for (int64_t q = 0; q < nufield(p); ++q) ut[q] = 2.0 * u[q] - u_pre[q];
    for (int64_t q = 0; q < nvfield(p); ++q) vt[q] = 2.0 * v[q] - v_pre[q];
  } else {
    memcpy(ut, u, (size_t)nufield(p) * sizeof(double));
    memcpy(vt, v, (size_t)nvfield(p) * sizeof(double));
  }
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nu_; ++i)
      fu[UI(i, k)] = fu[UI(i, k)] - re * (rho[CI(i, k)] + rho[CI(i - 1, k)]) / 2.0 *
                                         (ut[UI(i, k)] * (ut[UI(i + 1, k)] - ut[UI(i - 1, k)]) / 2.0 / dr +
                                          (vt[VI(i, k)] + vt[VI(i - 1, k)] + vt[VI(i, k + 1)] + vt[VI(i - 1, k + 1)]) / 4.0 *
                                              (ut[UI(i, k + 1)] - ut[UI(i, k - 1)]) / 2.0 / dz);
  for (int k = 1; k < ny; ++k)
    for (int i = 0; i < nx; ++i)
      fv[VI(i, k)] = fv[VI(i, k)] - re * (rho[CI(i, k)] + rho[CI(i, k - 1)]) / 2.0 *
                                         ((ut[UI(i, k)] + ut[UI(i + 1, k)] + ut[UI(i, k - 1)] + ut[UI(i + 1, k - 1)]) / 4.0 *
                                              (vt[VI(i + 1, k)] - vt[VI(i - 1, k)]) / 2.0 / dr +
                                          vt[VI(i, k)] * (vt[VI(i, k + 1)] - vt[VI(i, k - 1)]) / 2.0 / dz);
  /* pressure, inertia and body force :198-221 (bdf2), :192-215 (bdf1) */
  int64_t nc = ncell(p);
  if (bdf == 2)
    for (int64_t q = 0; q < nc; ++q) tp[q] = pp[q] + 4 * psi2[q] / 3.0 - psi1[q] / 3.0;
  else
    memcpy(tp, pp, (size_t)nc * sizeof(double));
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nu_; ++i) {
      double rb = (rho[CI(i, k)] + rho[CI(i - 1, k)]) / 2.0;
      double in;
      if (bdf == 2 && is_mf(p)) /* MF/VeloSolve_2_combined.f90:201: plain ggx */
        in = 2.0 * u[UI(i, k)] * re / dt - u_pre[UI(i, k)] * re / 2.0 / dt + p->ggx;
      else if (bdf == 2)
        in = 2.0 * u[UI(i, k)] * re / dt - u_pre[UI(i, k)] * re / 2.0 / dt +
             (ph_after[CI(i, k)] + ph_after[CI(i - 1, k)]) / 2.0 * p->ggx;
      else
        in = u[UI(i, k)] * re / dt + p->ggx;
      fu[UI(i, k)] = fu[UI(i, k)] - (tp[CI(i, k)] - tp[CI(i - 1, k)]) / dr + rb * in;
    }
  for (int k = 1; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      double rb = (rho[CI(i, k)] + rho[CI(i, k - 1)]) / 2.0;
      double in;
      if (bdf == 2 && is_mf(p)) /* MF/VeloSolve_2_combined.f90:208 */
        in = 2.0 * v[VI(i, k)] * re / dt - v_pre[VI(i, k)] * re / 2.0 / dt + p->ggz;
      else if (bdf == 2)
        in = 2.0 * v[VI(i, k)] * re / dt - v_pre[VI(i, k)] * re / 2.0 / dt +
             (ph_after[CI(i, k)] + ph_after[CI(i - 1, k)]) / 2.0 * p->ggz;
      else
        in = v[VI(i, k)] * re / dt + p->ggz;
      fv[VI(i, k)] = fv[VI(i, k)] - (tp[CI(i, k)] - tp[CI(i, k - 1)]) / dz + rb * in;
    }
  /* particle force :240-250 */
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nu_; ++i) fu[UI(i, k)] = fu[UI(i, k)] + (pf1[CI(i, k)] + pf1[CI(i - 1, k)]) / 2.0;
  for (int k = 1; k < ny; ++k)
    for (int i = 0; i < nx; ++i) fv[VI(i, k)] = fv[VI(i, k)] + (pf2[CI(i, k)] + pf2[CI(i, k - 1)]) / 2.0;
  /* times dt :253-263 */
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nu_; ++i) fu[UI(i, k)] = fu[UI(i, k)] * dt;
  for (int k = 1; k < ny; ++k)
    for (int i = 0; i < nx; ++i) fv[VI(i, k)] = fv[VI(i, k)] * dt;
  /* Dirichlet inflow :272-310 (always in MF, MF/VeloSolve_2_combined.f90:261-295) */
  if (lr_dirichlet(p)) {
    double cf = 1.0 - exp(-t / p->twall);
    double tmp_hig = p->bdr_upp - p->bdr_low;
    for (int k = 0; k < ny; ++k) {
      double z0 = (k + 0.5) * dz; /* z(0,k) TP/finit.f90:34-35 */
      if (z0 > p->bdr_low && z0 < p->bdr_upp) {
        double d = z0 - p->bdr_low - tmp_hig / 2;
        c->uL[k] = p->UU * (-(d * d) / tmp_hig / tmp_hig * 4 + 1) * cf;
      } else {
        c->uL[k] = 0.0;
      }
      fu[UI(0, k)] = fu[UI(0, k)] + 2.0 * c->uL[k] * eta[CI(0, k)] * 2.0 * dt / dr2;
    }
    for (int k = 0; k < ny; ++k) {
      c->uR[k] = c->uL[k];
      fu[UI(nx, k)] = fu[UI(nx, k)] + 2.0 * c->uR[k] * eta[CI(nx, k)] * 2.0 * dt / dr2;
    }
  }
  /* wall speed terms :392-400 */
  for (int i = 0; i < nu_; ++i) {
    int k = 0;
    double etay = (eta[CI(i, k)] + eta[CI(i - 1, k)] + eta[CI(i, k - 1)] + eta[CI(i - 1, k - 1)]) / 4.0;
    fu[UI(i, k)] = fu[UI(i, k)] + dt * (2.0 * p->uwB / dz2 * etay);
    k = ny - 1;
    etay = (eta[CI(i, k + 1)] + eta[CI(i - 1, k + 1)] + eta[CI(i, k)] + eta[CI(i - 1, k)]) / 4.0;
    fu[UI(i, k)] = fu[UI(i, k)] + dt * (2.0 * p->uwT / dz2 * etay);
  }
  atob_uv(c, fu, fv, rhs);
  return 0;
}

/* unpack and ghost rows :426-457 */
static void velo_unpack(oc_ctx* c, const double* x, double* u_after, double* v_after) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  btoa_uv(c, u_after, v_after, x);
  if (per_x(p)) {
    /* btoa_uv left the periodic side copies */
  } else if (lr_dirichlet(p)) {
    for (int k = 0; k < ny; ++k) u_after[UI(-1, k)] = 2.0 * c->uL[k] - u_after[UI(1, k)];
    for (int k = 0; k < ny; ++k) u_after[UI(nx + 1, k)] = 2.0 * c->uR[k] - u_after[UI(nx - 1, k)];
  } else {
    for (int k = 0; k <= ny; ++k) {
      u_after[UI(-1, k)] = u_after[UI(1, k)];
      u_after[UI(nx + 1, k)] = u_after[UI(nx - 1, k)];
    }
  }
  for (int i = -1; i <= nx + 1; ++i) {
    u_after[UI(i, -1)] = p->uwB * 2.0 - u_after[UI(i, 0)];
    u_after[UI(i, ny)] = p->uwT * 2.0 - u_after[UI(i, ny - 1)];
  }
  for (int i = -1; i <= nx; ++i) {
    v_after[VI(i, 0)] = 0.0;
    v_after[VI(i, ny)] = 0.0;
  }
  if (!per_x(p))
    for (int k = 1; k < ny; ++k) {
      v_after[VI(-1, k)] = v_after[VI(0, k)];
      v_after[VI(nx, k)] = v_after[VI(nx - 1, k)];
    }
}

static int velo_prec_id(const oc_ctx* c, int bdf) {
  switch (c->p.velo_prec) {
    case DRD_PREC_IDENTITY: return DRD_PS_IDENTITY;
    case DRD_PREC_JACOBI: return DRD_PS_VELO1_JACOBI;
    default: /* the reference's: Jacobi for BDF1; identity (TP) or Jacobi (MF) for BDF2 */
      return bdf == 1 ? DRD_PS_VELO1_JACOBI : (is_mf(&c->p) ? DRD_PS_VELO2_JACOBI : DRD_PS_IDENTITY);
  }
}

static int velo_solve(oc_ctx* c, int bdf, const double* psi1, const double* psi2,
                      const double* u_pre, const double* v_pre, const double* ph_after,
                      const double* rmu_after, double t, const double* pf1, const double* pf2,
                      double* u_after, double* v_after, drd_solve_info* info) {
  int64_t n = nuv(&c->p);
  double* rhs = (double*)malloc((size_t)n * sizeof(double));
  double* x = (double*)malloc((size_t)n * sizeof(double));
  int rc = oc_velo_rhs(c, bdf, psi1, psi2, u_pre, v_pre, ph_after, rmu_after, t, pf1, pf2, rhs);
  if (rc) {
    free(rhs);
    free(x);
    return rc;
  }
  atob_uv(c, c->f[DRD_F_U], c->f[DRD_F_V], x);
  if (bdf == 2 && c->guess == 1 && u_pre && v_pre) {  /* 2 x^n - x^{n-1} */
    double* xp = (double*)malloc((size_t)n * sizeof(double));
    atob_uv(c, u_pre, v_pre, xp);
    for (int64_t q = 0; q < n; ++q) x[q] = 2.0 * x[q] - xp[q];
    free(xp);
  }
  int32_t it = bdf == 1 ? c->p.maxit_velo1 : c->p.maxit_velo2, inf = 0;
  double res = c->p.tol;
  int op = bdf == 1 ? DRD_OP_VELO1 : DRD_OP_VELO2;
  oc_bicgstab(c, op, velo_prec_id(c, bdf), (int32_t)n, rhs, x, &it, &res, &inf);
  if (info) {
    info->iter = it;
    info->info = inf;
    info->resid = res;
    info->true_resid = true_resid(c, op, rhs, x);
  }
  velo_unpack(c, x, u_after, v_after);
  free(rhs);
  free(x);
  return 0;
}

int oc_velosolve_combined(oc_ctx* c, const double* psi, double* u_after, double* v_after,
                          const double* rmu_after, const double* ph_after, double t,
                          const double* pf1, const double* pf2, drd_solve_info* info) {
  (void)psi; /* temp_p = p (the psi term is commented out, :192) */
  return velo_solve(c, 1, NULL, NULL, NULL, NULL, ph_after, rmu_after, t, pf1, pf2, u_after,
                    v_after, info);
}

int oc_velosolve_2_combined(oc_ctx* c, const double* psi1, const double* psi2,
                            const double* u_pre, double* u_after, const double* v_pre,
                            double* v_after, const double* ph_pre, const double* ph_after,
                            const double* rmu_pre, const double* rmu_after, double t,
                            const double* pf1, const double* pf2, drd_solve_info* info) {
  (void)ph_pre;
  (void)rmu_pre; /* unused by the active code */
  return velo_solve(c, 2, psi1, psi2, u_pre, v_pre, ph_after, rmu_after, t, pf1, pf2, u_after,
                    v_after, info);
}

/* ---- pressure projection  TP/PreSolve_2.f90:16-125 + TP/dct_fftw.f90 --- */
int oc_presolve_2(oc_ctx* c, double* psi, double coef, double* mean_div) {
  int nx = NX, ny = NY;
  const double *u = c->f[DRD_F_U], *v = c->f[DRD_F_V];
  double dr = c->p.dx, dz = c->p.dz;
  double* uin = (double*)malloc((size_t)nx * ny * sizeof(double));
  double sum = 0.0;
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      uin[i + (size_t)nx * k] = (u[UI(i + 1, k)] - u[UI(i, k)]) / dr + (v[VI(i, k + 1)] - v[VI(i, k)]) / dz;
    }
  {  /* canonical order: each row pairwise, then the row sums pairwise */
    double* rs = (double*)calloc((size_t)ny, sizeof(double));
    for (int k = 0; k < ny; ++k) rs[k] = oc_pairwise(nx, uin + (size_t)nx * k, NULL);
    sum = oc_pairwise(ny, rs, NULL);
    free(rs);
  }
  if (mean_div) *mean_div = sum / nx / ny;
  /* dct_fftw(nr,nz,plan,plan_back,uin,psi,Pre_Lcc,4.0*nr*nz): REDFT10,
   * rhs / Pre_Lcc, REDFT01 (dct_fftw.f90:13-31) */
  const int per = per_x(&c->p);
  if (per)
    per_solve_2d(nx, ny, uin, c->pre_lcc);
  else
    dct_2d_fast(nx, ny, uin, c->pre_lcc, 2);
  double sc = per ? 2.0 * nx * ny : 4.0 * nx * ny;
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) psi[CI(i, k)] = uin[i + (size_t)nx * k] / sc;
  for (int k = 0; k < ny; ++k) {
    psi[CI(-1, k)] = per ? psi[CI(nx - 1, k)] : psi[CI(0, k)];
    psi[CI(nx, k)] = per ? psi[CI(0, k)] : psi[CI(nx - 1, k)];
  }
  for (int i = -1; i <= nx; ++i) {
    psi[CI(i, -1)] = psi[CI(i, 0)];
    psi[CI(i, ny)] = psi[CI(i, ny - 1)];
  }
  int64_t nc = ncell(&c->p);
  for (int64_t q = 0; q < nc; ++q) psi[q] = psi[q] * coef;
  free(uin);
  return 0;
}

/* ---- Cahn-Hilliard solves --------------------------------------------- */
/* ch_bulk_energy TP/ch_bulk_energy.f90:4-39.  phi**3.0 is evaluated as
 * phi*phi*phi, which is how the reference build lowers a real power with an
 * integral constant exponent (pinned against oracle/_ref). */
static double ch_bulk_energy(double phi, double psi, int mode) {
  if (mode == 1) return 0.5 * phi * (1.0 - phi) * (1.0 - 2.0 * phi) * (1 - psi);
  return (phi * phi * phi - phi) * (1 - psi);
}
/* ch_bdr_condition TP/ch_bdr_condition.f90:43-67 */
static double ch_bdr_condition(double angle, double phi_val, int mode) {
  double cf = cos(angle * PI_G / 180.0) / sqrt(2.0);
  if (mode == 1) return cf * (1.0 - phi_val) * phi_val;
  return cf * (1.0 - phi_val * phi_val);
}

/* RHS of ph_mu_solve TP/ph_mu_solve.f90:31-153; writes ph_M_0 and the
 * wall values pnb/pnt (returned through c->w5/c->w6 first nx entries). */
/* interpolate_coef TP/interpolate_coef.f90:54-99: fpd_ph at the interior
 * cells from dist (TP/dist.f90:63 -> fpd_dist TP/particles_move.f90:361-378)
 * at the cell centres (TP/finit.f90:44-45), rescaled, plus bdr_ph; the
 * ghost copies (columns, then rows); rho and eta from fpd_ph.  In the
 * reference's loop order. */
int oc_interpolate_coef(oc_ctx* c, double lambda_eta, const double* px, const double* py, int32_t np,
                        double radius, const double* sx, const double* sy, int32_t nswm, double swm_radius) {
  const drd_params* p = &c->p;
  const int nx = NX, ny = NY;
  double *fpd = c->f[DRD_F_FPD_PH], *rho = c->f[DRD_F_RHO], *eta = c->f[DRD_F_ETA];
  const double* bdr = c->f[DRD_F_BDR_PH];
  const int64_t nc = (int64_t)(nx + 2) * (ny + 2);
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      const double x = (i + 0.5) * p->dx, y = (k + 0.5) * p->dz;
      double d = 1.0E10;
      for (int32_t j = 0; j < np; ++j) {
        const double e = sqrt((x - px[j]) * (x - px[j]) + (y - py[j]) * (y - py[j])) - radius;
        d = e < d ? e : d;
      }
      for (int32_t j = 0; j < nswm; ++j) {
        const double e = sqrt((x - sx[j]) * (x - sx[j]) + (y - sy[j]) * (y - sy[j])) - swm_radius;
        d = e < d ? e : d;
      }
      fpd[CI(i, k)] = tanh(-1.0 * d / sqrt(2.0) / p->I_thickness);
    }
  for (int64_t q = 0; q < nc; ++q) fpd[q] = (fpd[q] + 1.0) / 2.0;
  for (int64_t q = 0; q < nc; ++q) fpd[q] = fpd[q] + bdr[q];
  const int per = per_x(p);
  for (int k = 0; k < ny; ++k) {
    fpd[CI(-1, k)] = per ? fpd[CI(nx - 1, k)] : fpd[CI(0, k)];
    fpd[CI(nx, k)] = per ? fpd[CI(0, k)] : fpd[CI(nx - 1, k)];
  }
  for (int i = -1; i <= nx; ++i) {
    fpd[CI(i, -1)] = fpd[CI(i, 0)];
    fpd[CI(i, ny)] = fpd[CI(i, ny - 1)];
  }
  for (int64_t q = 0; q < nc; ++q) {
    rho[q] = 1.0 + (p->lambda_rho - 1.0) * fpd[q];
    eta[q] = 1.0 + (lambda_eta - 1.0) * fpd[q];
  }
  return 0;
}

/* fpd_dist_periodic TP/particles_move.f90:326-357, then the weight of
 * force_caculation_vel / update_position_vel (:786-787, :863-865). */
static double oc_body_weight(double x, double y, double lx, double ly, double rx, double ry, double r, double I) {
  double xd = 0.0, yd = 0.0;
  if (rx < r) xd = -lx;
  else if (lx - rx < r) xd = lx;
  if (ry < r) yd = -ly;
  else if (ly - ry < r) yd = ly;
  double d = sqrt((x - rx) * (x - rx) + (y - ry) * (y - ry)) - r;
  double e = sqrt((x + xd - rx) * (x + xd - rx) + (y - ry) * (y - ry)) - r;
  d = e < d ? e : d;
  e = sqrt((x - rx) * (x - rx) + (y + yd - ry) * (y + yd - ry)) - r;
  d = e < d ? e : d;
  e = sqrt((x + xd - rx) * (x + xd - rx) + (y + yd - ry) * (y + yd - ry)) - r;
  d = e < d ? e : d;
  return (tanh(-1.0 * d / sqrt(2.0) / I) + 1.0) / 2.0;
}

/* force_caculation_vel TP/particles_move.f90:758-839: the bodies' loops in
 * order (particles, swimmers, phantoms as one list), each adding
 * force / sum_val with force = F * w and sum_val = (sum of w) * dr * dz. */
int oc_particle_force(oc_ctx* c, const double* bx, const double* by, const double* br, const double* fx,
                      const double* fy, int32_t nbody, double lx, double ly) {
  const drd_params* p = &c->p;
  const int nx = NX, ny = NY;
  const double dr = p->dx, dz = p->dz, I = p->I_thickness;
  double *pf1 = c->f[DRD_F_PFORCE1], *pf2 = c->f[DRD_F_PFORCE2];
  const size_t nc = (size_t)(nx + 2) * (ny + 2);
  double* w = (double*)malloc(nc * sizeof(double));
  if (!w) return -1;
  for (size_t q = 0; q < nc; ++q) pf1[q] = pf2[q] = 0.0;
  for (int32_t b = 0; b < nbody; ++b) {
    double sum_val = 0.0;
    for (int k = 0; k < ny; ++k)
      for (int i = 0; i < nx; ++i) {
        const double dd = oc_body_weight((i + 0.5) * dr, (k + 0.5) * dz, lx, ly, bx[b], by[b], br[b], I);
        w[CI(i, k)] = dd;
        sum_val = sum_val + dd;
      }
    sum_val = sum_val * dr * dz;
    for (int k = 0; k < ny; ++k)
      for (int i = 0; i < nx; ++i) {
        pf1[CI(i, k)] = pf1[CI(i, k)] + (0.0 + fx[b] * w[CI(i, k)]) / sum_val;
        pf2[CI(i, k)] = pf2[CI(i, k)] + (0.0 + fy[b] * w[CI(i, k)]) / sum_val;
      }
  }
  if (per_x(p)) { /* periodic x: the ghost columns the momentum rows read */
    xghost_c(c, pf1);
    xghost_c(c, pf2);
  }
  free(w);
  return 0;
}

/* update_position_vel TP/particles_move.f90:842-920, the averages only. */
int oc_particle_velocity(oc_ctx* c, const double* bx, const double* by, const double* br, int32_t nbody, double lx,
                         double ly, double* u_aver, double* v_aver) {
  const drd_params* p = &c->p;
  const int nx = NX, ny = NY;
  const double dr = p->dx, dz = p->dz, I = p->I_thickness;
  const double *u = c->f[DRD_F_U], *v = c->f[DRD_F_V];
  for (int32_t b = 0; b < nbody; ++b) {
    double ua = 0.0, va = 0.0, sum_val = 0.0;
    for (int k = 0; k < ny; ++k)
      for (int i = 0; i < nx; ++i) {
        const double dd = oc_body_weight((i + 0.5) * dr, (k + 0.5) * dz, lx, ly, bx[b], by[b], br[b], I);
        ua = ua + (u[UI(i, k)] + u[UI(i, k + 1)]) / 2.0 * dd;
        va = va + (v[VI(i, k)] + v[VI(i, k + 1)]) / 2.0 * dd;
        sum_val = sum_val + dd;
      }
    u_aver[b] = ua / sum_val;
    v_aver[b] = va / sum_val;
  }
  return 0;
}

int oc_ph_mu_rhs(oc_ctx* c, double* rhs) {
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  const double *ph = c->f[DRD_F_PH], *u = c->f[DRD_F_U], *v = c->f[DRD_F_V];
  const double *bdr = c->f[DRD_F_BDR_PH], *bu = c->f[DRD_F_BDR_PH_USED], *bg = c->f[DRD_F_BDR_PH_GRAD];
  double* m0 = c->f[DRD_F_PH_M_0];
  double dt = p->dt, dr = p->dx, dz = p->dz, dz2 = dz * dz;
  double I = p->I_thickness, s = p->s_implicit;
  int64_t nc = ncell(p);
  double *f1 = c->w1, *f2 = c->w2;
  for (int64_t q = 0; q < nc; ++q) {
    if (p->phi_mode == 1)
      m0[q] = (1.0 - bdr[q]) * (ph[q] * ph[q]) * ((ph[q] - 1.0) * (ph[q] - 1.0));
    else
      m0[q] = (1.0 - bdr[q]) * ((ph[q] * ph[q] - 1.0) * (ph[q] * ph[q] - 1.0));
  }
  for (int64_t q = 0; q < nc; ++q) f2[q] = ph[q] * (bu[q] + p->bdr_ph_used_adjust);
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i)
      f2[CI(i, k)] = f2[CI(i, k)] -
                     dt * (u[UI(i + 1, k)] + u[UI(i, k)]) / 2.0 *
                         (ph[CI(i + 1, k)] * bu[CI(i + 1, k)] - ph[CI(i - 1, k)] * bu[CI(i - 1, k)]) / 2.0 / dr -
                     dt * (v[VI(i, k + 1)] + v[VI(i, k)]) / 2.0 *
                         (ph[CI(i, k + 1)] * bu[CI(i, k + 1)] - ph[CI(i, k - 1)] * bu[CI(i, k - 1)]) / 2.0 / dz;
  double cosa = cos(p->angle * PI_G / 180);
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      double phv = ph[CI(i, k)];
      if (p->phi_mode == 1)
        f1[CI(i, k)] = s * phv / I - ch_bulk_energy(phv, bdr[CI(i, k)], 1) / I -
                       phv * (phv - 1.0) * bg[CI(i, k)] * cosa / sqrt(2.0);
      else
        f1[CI(i, k)] = s * phv / I - ch_bulk_energy(phv, bdr[CI(i, k)], 2) / I -
                       (phv * phv - 1.0) * bg[CI(i, k)] * cosa / sqrt(2.0);
    }
  double* pnb = c->w5;
  double* pnt = c->w6;
  const double *aB = c->f[DRD_F_ANGLEB], *aT = c->f[DRD_F_ANGLET];
  for (int i = 0; i < nx; ++i) {
    double tmp = (ph[CI(i, -1)] + ph[CI(i, 0)]) / 2.0;
    pnb[i] = ch_bdr_condition(aB[i], tmp, p->phi_mode);
  }
  for (int i = 0; i < nx; ++i) {
    double tmp = I / dz;
    c->coef_bot = (I / dz) / tmp;
    pnb[i] = pnb[i] / tmp;
    f1[CI(i, 0)] = f1[CI(i, 0)] + pnb[i] * I / dz2;
  }
  for (int i = 0; i < nx; ++i) {
    double tmp = (ph[CI(i, ny - 1)] + ph[CI(i, ny)]) / 2.0;
    pnt[i] = ch_bdr_condition(aT[i], tmp, p->phi_mode);
  }
  for (int i = 0; i < nx; ++i) {
    double tmp = I / dz;
    c->coef_top = (I / dz) / tmp;
    pnt[i] = pnt[i] / tmp;
    f1[CI(i, ny - 1)] = f1[CI(i, ny - 1)] + pnt[i] * I / dz2;
  }
  atob_ph_mu(c, f1, f2, rhs);
  return 0;
}

static int ph_mu_solve_x(oc_ctx* c, double* ph_after, double* rmu_after, drd_solve_info* info, int extrapolate);
int oc_ph_mu_solve(oc_ctx* c, double* ph_after, double* rmu_after, double t, drd_solve_info* info) {
  (void)t;
  return ph_mu_solve_x(c, ph_after, rmu_after, info, 0);
}
static int ph_mu_solve_x(oc_ctx* c, double* ph_after, double* rmu_after, drd_solve_info* info, int extrapolate) {
  int nx = NX, ny = NY;
  int64_t n = 2 * (int64_t)nx * ny;
  double* rhs = (double*)malloc((size_t)n * sizeof(double));
  double* x = (double*)malloc((size_t)n * sizeof(double));
  double* pn = (double*)malloc((size_t)2 * nx * sizeof(double));
  oc_ph_mu_rhs(c, rhs);
  memcpy(pn, c->w5, (size_t)nx * sizeof(double));
  memcpy(pn + nx, c->w6, (size_t)nx * sizeof(double));
  atob_ph_mu(c, c->f[DRD_F_PH], c->f[DRD_F_RMU], x);
  if (extrapolate) { /* 2 (phi, mu)^n - (phi, mu)^{n-1} */
    double* xp = (double*)malloc((size_t)n * sizeof(double));
    atob_ph_mu(c, c->f[DRD_F_PH_PRE], c->f[DRD_F_RMU_PRE], xp);
    for (int64_t q = 0; q < n; ++q) x[q] = 2.0 * x[q] - xp[q];
    free(xp);
  }
  int32_t it = c->p.maxit_phmu, inf = 0;
  double res = c->p.tol;
  oc_bicgstab(c, DRD_OP_PHMU, c->p.ch_prec == DRD_CH_PREC_BLOCK ? DRD_PS_PHMU_BLOCK : DRD_PS_PHMU_JACOBI,
              (int32_t)n, rhs, x, &it, &res, &inf);
  if (info) {
    info->iter = it;
    info->info = inf;
    info->resid = res;
    info->true_resid = true_resid(c, DRD_OP_PHMU, rhs, x);
  }
  btoa_ph_mu(c, ph_after, rmu_after, x);
  for (int i = 0; i < nx; ++i) {
    ph_after[CI(i, -1)] = ph_after[CI(i, 0)] * c->coef_bot + pn[i];
    ph_after[CI(i, ny)] = ph_after[CI(i, ny - 1)] * c->coef_top + pn[nx + i];
  }
  if (per_x(&c->p)) xghost_c(c, ph_after); /* corners: copies of the wall ghosts */
  free(rhs);
  free(x);
  free(pn);
  return 0;
}

/* ch_bulk_energy MF/ch_bulk_energy.f90:4-28 (localised, phi in [0,1]) */
static double ch_bulk_energy_mf(double phi, double psi) {
  double aa = pow(phi, 2.0) * pow(1.0 - phi, 2.0);
  aa = 1.0 - sqrt(aa / (aa + 1.0e-2)) * psi;
  return 0.5 * phi * (1.0 - phi) * (1.0 - 2.0 * phi) * aa;
}

/* PhSolve MF/PhSolve.f90:1-126 */
int oc_phsolve(oc_ctx* c, double* rmu_after, double* ph_after, double t, drd_solve_info* info) {
  (void)t;
  const drd_params* p = &c->p;
  int nx = NX, ny = NY;
  const double *ph = c->f[DRD_F_PH], *u = c->f[DRD_F_U], *v = c->f[DRD_F_V];
  const double *bdr = c->f[DRD_F_BDR_PH], *bg = c->f[DRD_F_BDR_PH_GRAD];
  double dt = p->dt, dr = p->dx, dz = p->dz, dr2 = dr * dr, dz2 = dz * dz;
  double I = p->I_thickness, s = p->s_implicit;
  int64_t nc = ncell(p), n = (int64_t)nx * ny;
  double* f1 = (double*)calloc((size_t)nc, sizeof(double));
  double* f2 = (double*)calloc((size_t)nc, sizeof(double));
  double* ff = (double*)calloc((size_t)nc, sizeof(double));
  double* rhs = (double*)malloc((size_t)n * sizeof(double));
  double* x = (double*)malloc((size_t)n * sizeof(double));
  memcpy(f2, ph, (size_t)nc * sizeof(double));
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i)
      f2[CI(i, k)] = f2[CI(i, k)] - dt * (u[UI(i + 1, k)] + u[UI(i, k)]) / 2.0 * (ph[CI(i + 1, k)] - ph[CI(i - 1, k)]) / 2.0 / dr -
                     dt * (v[VI(i, k + 1)] + v[VI(i, k)]) / 2.0 * (ph[CI(i, k + 1)] - ph[CI(i, k - 1)]) / 2.0 / dz;
  double cosa = cos(p->angle * PI_G / 180);
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      double phv = ph[CI(i, k)];
      f1[CI(i, k)] = s * phv / I - ch_bulk_energy_mf(phv, bdr[CI(i, k)]) / I -
                     phv * (phv - 1.0) * bg[CI(i, k)] * cosa / sqrt(2.0);
    }
  oc_op_double_gradient_neum(nx, ny, dr2, dz2, f2, c->f[DRD_F_PH_M_INNER], ff);
  for (int64_t q = 0; q < nc; ++q) f1[q] = -f1[q] + s * f2[q] / I - I * ff[q];
  atob_c(c, f1, rhs);
  atob_c(c, c->f[DRD_F_RMU], x);
  int32_t it = c->p.maxit_ph, inf = 0;
  double res = c->p.tol;
  oc_bicgstab(c, DRD_OP_PH, DRD_PS_PH_DCT, (int32_t)n, rhs, x, &it, &res, &inf);
  if (info) {
    info->iter = it;
    info->info = inf;
    info->resid = res;
    info->true_resid = true_resid(c, DRD_OP_PH, rhs, x);
  }
  btoa_c(c, rmu_after, x);
  for (int i = 0; i < nx; ++i) {
    rmu_after[CI(i, -1)] = rmu_after[CI(i, 0)];
    rmu_after[CI(i, ny)] = rmu_after[CI(i, ny - 1)];
  }
  for (int k = -1; k <= ny; ++k) {
    rmu_after[CI(-1, k)] = rmu_after[CI(0, k)];
    rmu_after[CI(nx, k)] = rmu_after[CI(nx - 1, k)];
  }
  oc_op_double_gradient_neum(nx, ny, dr2, dz2, rmu_after, c->f[DRD_F_PH_M_0], ff);
  for (int64_t q = 0; q < nc; ++q) ph_after[q] = f2[q] + dt * p->xld * ff[q];
  for (int i = 0; i < nx; ++i) ph_after[CI(i, -1)] = ph_after[CI(i, 0)];
  for (int i = 0; i < nx; ++i) ph_after[CI(i, ny)] = ph_after[CI(i, ny - 1)];
  for (int k = -1; k <= ny; ++k) {
    ph_after[CI(-1, k)] = ph_after[CI(0, k)];
    ph_after[CI(nx, k)] = ph_after[CI(nx - 1, k)];
  }
  free(f1); free(f2); free(ff); free(rhs); free(x);
  return 0;
}

/* ---- PhSolve_ori and the CH re-initialisation of solid domains ---------- */
/* laplace_nn TP/PhSolve_ori.f90:463-487: the 5-point Laplacian of the
 * interior with the ghosts set to copies of their neighbour first (rows,
 * then columns, corners included), as the reference sets them in a. */
static void laplace_nn(const oc_ctx* c, double* a, double* b) {
  const int nx = NX, ny = NY;
  const double dr2 = c->p.dx * c->p.dx, dz2 = c->p.dz * c->p.dz;
  for (int i = 0; i < nx; ++i) {
    a[CI(i, -1)] = a[CI(i, 0)];
    a[CI(i, ny)] = a[CI(i, ny - 1)];
  }
  for (int j = -1; j <= ny; ++j) {
    a[CI(-1, j)] = a[CI(0, j)];
    a[CI(nx, j)] = a[CI(nx - 1, j)];
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      b[CI(i, j)] = (a[CI(i + 1, j)] + a[CI(i - 1, j)] - 2 * a[CI(i, j)]) / dr2 +
                    (a[CI(i, j + 1)] + a[CI(i, j - 1)] - 2 * a[CI(i, j)]) / dz2;
}

/* MATRIX_VECTOR_CG_ori TP/PhSolve_ori.f90:348-397 (MF/PhSolve_ori.f90 is
 * the same): y = alpha (x - dt M lap(s/eps x - eps lap x + walls)) + beta y,
 * the wall rows taking I/dz2 (x(i,-1) - coef_bot x(i,0)) with the ghost
 * x(i,-1) = x(i,0) laplace_nn leaves.  Neumann sides (the re-initialisation
 * runs with lr_bd_type 'n'). */
static void phori_matvec(oc_ctx* c, double alpha, const double* xx, double beta, double* yy) {
  const drd_params* p = &c->p;
  const int nx = NX, ny = NY;
  const double I = p->I_thickness, dz2 = p->dz * p->dz;
  double *f1 = c->w1, *ff1 = c->w3, *ff2 = c->w4;
  const int64_t nc = ncell(p);
  for (int64_t q = 0; q < nc; ++q) f1[q] = ff1[q] = ff2[q] = 0.0;
  btoa_c(c, f1, xx);
  laplace_nn(c, f1, ff1);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) ff1[CI(i, j)] = -I * ff1[CI(i, j)];
  for (int i = 0; i < nx; ++i) {
    int j = 0;
    ff1[CI(i, j)] = ff1[CI(i, j)] + I / dz2 * (f1[CI(i, j - 1)] - c->coef_bot * f1[CI(i, j)]);
    j = ny - 1;
    ff1[CI(i, j)] = ff1[CI(i, j)] + I / dz2 * (f1[CI(i, j + 1)] - c->coef_top * f1[CI(i, j)]);
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) ff1[CI(i, j)] = ff1[CI(i, j)] + f1[CI(i, j)] * p->s_implicit / I;
  laplace_nn(c, ff1, ff2);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double z = f1[CI(i, j)] + -p->dt * p->xld * ff2[CI(i, j)];
      const size_t q = i + (size_t)nx * j;
      yy[q] = beta == 0.0 ? alpha * z : alpha * z + beta * yy[q];
    }
}

/* PhSolve_ori TP/PhSolve_ori.f90:1-234: one implicit CH step for phi with
 * unit mobility, wall ('n' type) boundary terms with alpha_s and the contact
 * angles angleB/angleT, the DCT preconditioner of PhSolve (PSolve_CG_ori
 * :402-416 = PSolve_CG), then mu from the new phi. */
int oc_phsolve_ori(oc_ctx* c, double alpha_s, double* rmu_after, double* ph_after, drd_solve_info* info) {
  const drd_params* p = &c->p;
  if (per_x(p) || lr_dirichlet(p)) return DRD_E_UNSUPPORTED;
  const int nx = NX, ny = NY;
  const double *ph = c->f[DRD_F_PH], *u = c->f[DRD_F_U], *v = c->f[DRD_F_V];
  const double *aB = c->f[DRD_F_ANGLEB], *aT = c->f[DRD_F_ANGLET];
  const double dt = p->dt, dr = p->dx, dz = p->dz, dr2 = dr * dr, dz2 = dz * dz;
  const double I = p->I_thickness, s = p->s_implicit;
  const int64_t nc = ncell(p), n = (int64_t)nx * ny;
  double* f1 = (double*)calloc((size_t)nc, sizeof(double));
  double* f2 = (double*)calloc((size_t)nc, sizeof(double));
  double* ff = (double*)calloc((size_t)nc, sizeof(double));
  double* pnb = (double*)calloc((size_t)nx, sizeof(double));
  double* pnt = (double*)calloc((size_t)nx, sizeof(double));
  double* rhs = (double*)malloc((size_t)n * sizeof(double));
  double* x = (double*)malloc((size_t)n * sizeof(double));
  memcpy(f2, ph, (size_t)nc * sizeof(double));
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i)
      f2[CI(i, k)] = f2[CI(i, k)] - dt * (u[UI(i + 1, k)] + u[UI(i, k)]) / 2.0 * (ph[CI(i + 1, k)] - ph[CI(i - 1, k)]) / 2.0 / dr -
                     dt * (v[VI(i, k + 1)] + v[VI(i, k)]) / 2.0 * (ph[CI(i, k + 1)] - ph[CI(i, k - 1)]) / 2.0 / dz;
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      const double phv = ph[CI(i, k)];
      f1[CI(i, k)] = ((1. + s) * phv - phv * phv * phv) / I;
    }
  for (int i = 0; i < nx; ++i) { /* bottom wall, 'n' type (:48-80) */
    const double cf = sqrt(2.0) * cos(aB[i] * PI_G / 180.0) / 3.0 * PI_G / 2.;
    const double tmp = (ph[CI(i, -1)] + ph[CI(i, 0)]) / 2.0;
    pnb[i] = cf * cos(tmp * PI_G / 2) + alpha_s * tmp;
    const double tmp1 = I / dz + 0.5 * alpha_s;
    c->coef_bot = (I / dz - 0.5 * alpha_s) / tmp1;
    pnb[i] = pnb[i] / tmp1;
    f1[CI(i, 0)] = f1[CI(i, 0)] + pnb[i] * I / dz2;
  }
  for (int i = 0; i < nx; ++i) { /* top wall (:83-118) */
    const double cf = sqrt(2.0) * cos(aT[i] * PI_G / 180.0) / 3.0 * PI_G / 2.;
    const double tmp = (ph[CI(i, ny - 1)] + ph[CI(i, ny)]) / 2.0;
    pnt[i] = cf * cos(tmp * PI_G / 2) + alpha_s * tmp;
    const double tmp1 = I / dz + 0.5 * alpha_s;
    c->coef_top = (I / dz - 0.5 * alpha_s) / tmp1;
    pnt[i] = pnt[i] / tmp1;
    f1[CI(i, ny - 1)] = f1[CI(i, ny - 1)] + pnt[i] * I / dz2;
  }
  laplace_nn(c, f1, ff);
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) f2[CI(i, k)] = f2[CI(i, k)] - dt * p->xld * ff[CI(i, k)];
  atob_c(c, f2, rhs);
  atob_c(c, ph, x);
  int32_t it = 500, inf = 0; /* :148 */
  double res = p->tol;
  oc_bicgstab(c, DRD_OP_PHORI, DRD_PS_PH_DCT, (int32_t)n, rhs, x, &it, &res, &inf);
  if (info) {
    info->iter = it;
    info->info = inf;
    info->resid = res;
    info->true_resid = true_resid(c, DRD_OP_PHORI, rhs, x);
  }
  for (int64_t q = 0; q < nc; ++q) ph_after[q] = 0.0;
  btoa_c(c, ph_after, x);
  for (int i = 0; i < nx; ++i) {
    ph_after[CI(i, -1)] = ph_after[CI(i, 0)] * c->coef_bot + pnb[i];
    ph_after[CI(i, ny)] = ph_after[CI(i, ny - 1)] * c->coef_top + pnt[i];
  }
  for (int k = -1; k <= ny; ++k) {
    ph_after[CI(-1, k)] = ph_after[CI(0, k)];
    ph_after[CI(nx, k)] = ph_after[CI(nx - 1, k)];
  }
  for (int k = 0; k < ny; ++k)
    for (int i = 0; i < nx; ++i) {
      const double a = ph_after[CI(i, k)];
      rmu_after[CI(i, k)] = -I * ((ph_after[CI(i + 1, k)] + ph_after[CI(i - 1, k)] - 2 * a) / dr2 +
                                  (ph_after[CI(i, k + 1)] + ph_after[CI(i, k - 1)] - 2 * a) / dz2) -
                            a / I + a * a * a / I;
    }
  for (int i = 0; i < nx; ++i) {
    rmu_after[CI(i, -1)] = rmu_after[CI(i, 0)];
    rmu_after[CI(i, ny)] = rmu_after[CI(i, ny - 1)];
  }
  for (int k = -1; k <= ny; ++k) {
    rmu_after[CI(-1, k)] = rmu_after[CI(0, k)];
    rmu_after[CI(nx, k)] = rmu_after[CI(nx - 1, k)];
  }
  free(f1); free(f2); free(ff); free(pnb); free(pnt); free(rhs); free(x);
  return 0;
}

/* The CH re-initialisation of a solid domain, MF/finit.f90:132-188 (=
 * TP/reinint_by_PhSolve.f90:1-91): phi := bdr_ph with copied ghosts, unit
 * mobilities, Ph_Lcc rebuilt for the step dt, `steps` PhSolve_ori steps
 * (the reference takes 51 at dt = 1e-4), then bdr_ph = (phi + 1) / 2 (over
 * the whole field; the interior clamped to [0, 1]) and bdr_ph_penalty =
 * [bdr_ph > 0.5].  Left as the reference leaves them: phi and mu the last
 * step's, ph_M_0 = ph_M_inner = 1; the context's own dt and Ph_Lcc are
 * restored.  info: the last step's solve. */
static void build_ph_lcc(oc_ctx* c, double dt);
int oc_reinit_solid(oc_ctx* c, int32_t steps, double dt, double alpha_s, drd_solve_info* info) {
  const drd_params* p = &c->p;
  if (per_x(p) || lr_dirichlet(p)) return DRD_E_UNSUPPORTED;
  const int nx = NX, ny = NY;
  const int64_t nc = ncell(p);
  double *ph = c->f[DRD_F_PH], *rmu = c->f[DRD_F_RMU], *bdr = c->f[DRD_F_BDR_PH];
  memcpy(ph, bdr, (size_t)nc * sizeof(double));
  for (int j = 0; j < ny; ++j) {
    ph[CI(-1, j)] = ph[CI(0, j)];
    ph[CI(nx, j)] = ph[CI(nx - 1, j)];
  }
  for (int i = -1; i <= nx; ++i) {
    ph[CI(i, -1)] = ph[CI(i, 0)];
    ph[CI(i, ny)] = ph[CI(i, ny - 1)];
  }
  for (int64_t q = 0; q < nc; ++q) c->f[DRD_F_PH_M_0][q] = c->f[DRD_F_PH_M_INNER][q] = 1.0;
  const double dt0 = c->p.dt;
  c->p.dt = dt;
  build_ph_lcc(c, dt);
  double* pa = (double*)calloc((size_t)nc, sizeof(double));
  double* ma = (double*)calloc((size_t)nc, sizeof(double));
  drd_solve_info loc;
  memset(&loc, 0, sizeof(loc));
  int rc = 0;
  for (int32_t s = 0; s < steps && !rc; ++s) {
    rc = oc_phsolve_ori(c, alpha_s, ma, pa, &loc);
    memcpy(rmu, ma, (size_t)nc * sizeof(double));
    memcpy(ph, pa, (size_t)nc * sizeof(double));
  }
  c->p.dt = dt0;
  build_ph_lcc(c, dt0);
  for (int64_t q = 0; q < nc; ++q) bdr[q] = (ph[q] + 1.0) / 2.0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (bdr[CI(i, j)] >= 1.0)
        bdr[CI(i, j)] = 1.0;
      else if (bdr[CI(i, j)] <= 0.0)
        bdr[CI(i, j)] = 0.0;
    }
  for (int64_t q = 0; q < nc; ++q) c->f[DRD_F_BDR_PH_PENALTY][q] = bdr[q] > 0.5 ? 1.0 : 0.0;
  if (info) *info = loc;
  free(pa);
  free(ma);
  return rc;
}

/* ---- one time step  TP/mcl6.f90:257-303 (bdf 1) and :395-502 (bdf 2); the MF
 * variant takes PhSolve for the CH solve (MF/mcl6.f90:239-300, :388-440) -- */
int oc_step(oc_ctx* c, int32_t bdf, double t, drd_step_info* info) {
  const drd_params* p = &c->p;
  int64_t nc = ncell(p), nu = nufield(p), nv = nvfield(p);
  double* ph_after = (double*)calloc((size_t)nc, sizeof(double));
  double* rmu_after = (double*)calloc((size_t)nc, sizeof(double));
  double* u_after = (double*)calloc((size_t)nu, sizeof(double));
  double* v_after = (double*)calloc((size_t)nv, sizeof(double));
  double* zero = (double*)calloc((size_t)nc, sizeof(double));
  drd_step_info loc;
  memset(&loc, 0, sizeof(loc));
  double gamma1 = p->lambda_rho < 1.0 ? p->lambda_rho : 1.0;
  double** F = c->f;
  int rc;
  if (bdf == 1) {
    if (is_mf(p)) /* MF/mcl6.f90:239 */
      oc_phsolve(c, rmu_after, ph_after, t + p->dt, &loc.phmu);
    else
      oc_ph_mu_solve(c, ph_after, rmu_after, t, &loc.phmu);
    rc = oc_velosolve_combined(c, F[DRD_F_PSI1], u_after, v_after, rmu_after, ph_after, t + p->dt,
                               c->f[DRD_F_PFORCE1], c->f[DRD_F_PFORCE2], &loc.velo);
    if (rc) goto out;
    memcpy(F[DRD_F_U_PRE], F[DRD_F_U], (size_t)nu * sizeof(double));
    memcpy(F[DRD_F_V_PRE], F[DRD_F_V], (size_t)nv * sizeof(double));
    memcpy(F[DRD_F_U], u_after, (size_t)nu * sizeof(double));
    memcpy(F[DRD_F_V], v_after, (size_t)nv * sizeof(double));
    oc_presolve_2(c, F[DRD_F_PSI2], p->re * gamma1 / p->dt, &loc.div_before);
    for (int64_t q = 0; q < nc; ++q) F[DRD_F_P][q] = F[DRD_F_PSI2][q] + F[DRD_F_P][q];
    memcpy(F[DRD_F_PH_PRE], F[DRD_F_PH], (size_t)nc * sizeof(double));
    memcpy(F[DRD_F_RMU_PRE], F[DRD_F_RMU], (size_t)nc * sizeof(double));
    memcpy(F[DRD_F_RMU], rmu_after, (size_t)nc * sizeof(double));
    memcpy(F[DRD_F_PH], ph_after, (size_t)nc * sizeof(double));
  } else {
    if (is_mf(p)) /* MF/mcl6.f90:388 */
      oc_phsolve(c, rmu_after, ph_after, t + p->dt, &loc.phmu);
    else
      ph_mu_solve_x(c, ph_after, rmu_after, &loc.phmu, c->guess == 1);
    rc = oc_velosolve_2_combined(c, F[DRD_F_PSI1], F[DRD_F_PSI2], F[DRD_F_U_PRE], u_after,
                                 F[DRD_F_V_PRE], v_after, F[DRD_F_PH_PRE], ph_after,
                                 F[DRD_F_RMU_PRE], rmu_after, t + p->dt, c->f[DRD_F_PFORCE1], c->f[DRD_F_PFORCE2], &loc.velo);
    if (rc) goto out;
    memcpy(F[DRD_F_U_PRE], F[DRD_F_U], (size_t)nu * sizeof(double));
    memcpy(F[DRD_F_V_PRE], F[DRD_F_V], (size_t)nv * sizeof(double));
    memcpy(F[DRD_F_U], u_after, (size_t)nu * sizeof(double));
    memcpy(F[DRD_F_V], v_after, (size_t)nv * sizeof(double));
    memcpy(F[DRD_F_PH_PRE], F[DRD_F_PH], (size_t)nc * sizeof(double));
    memcpy(F[DRD_F_RMU_PRE], F[DRD_F_RMU], (size_t)nc * sizeof(double));
    memcpy(F[DRD_F_PH], ph_after, (size_t)nc * sizeof(double));
    memcpy(F[DRD_F_RMU], rmu_after, (size_t)nc * sizeof(double));
    memcpy(F[DRD_F_PSI1], F[DRD_F_PSI2], (size_t)nc * sizeof(double));
    oc_presolve_2(c, F[DRD_F_PSI2], 3 * p->re * gamma1 / 2. / p->dt, &loc.div_before);
    for (int64_t q = 0; q < nc; ++q) F[DRD_F_P][q] = F[DRD_F_PSI2][q] + F[DRD_F_P][q];
  }
  rc = 0;
out:
  if (info) *info = loc;
  free(ph_after); free(rmu_after); free(u_after); free(v_after); free(zero);
  return rc;
}

/* scalar helpers exported for pinning against the reference build */
double oc_ch_bulk_energy(double phi, double psi, int32_t mode) { return ch_bulk_energy(phi, psi, mode); }
double oc_ch_bulk_energy_mf(double phi, double psi) { return ch_bulk_energy_mf(phi, psi); }
double oc_ch_bdr_condition(double angle, double phi, int32_t mode) {
  return ch_bdr_condition(angle, phi, mode);
}
