// kernels_mg.hip -- geometric multigrid V-cycle preconditioner for the BDF2
// velocity operator (fast path; no reference counterpart).
//
// Level 0 is the solve's own grid; level l+1 halves level l in x and y while
// the result keeps nx >= 64 and ny >= 32.  The coarse operators are
// rediscretised: the same MAC stencil (velo_row) with doubled spacings, on
// coefficients rho, eta, chi averaged over 2x2 cells.  One V-cycle:
//   down:  x = omega b / d, one more damped Jacobi sweep, r = b - A x,
//          b_{l+1} = R r
//   coarsest: x = the FP64 DCT/DST fast-Poisson solve of its b
//   up:    x += P x_{l+1}, two damped Jacobi sweeps
// with omega = 0.7 and d the exact diagonal of each level's operator.  The
// cycle is a fixed linear map of b, so BiCGSTAB accepts it.  Every kernel
// skips once the solve it serves is done or has taken its early exit.
#include <vector>

#include "kernels.hpp"

namespace drd {

struct MGLevel {
  Grid g;
  Scal s;
  Coef c{};                   // rho, eta, chi of the level (level 0: the context's fields)
  double* fields = nullptr;   // own coarse rho, eta, chi (levels >= 1)
  double *d = nullptr, *x = nullptr, *t = nullptr, *b = nullptr, *r = nullptr;
};

struct VeloMG {
  std::vector<MGLevel> lv;
  VeloPrec* coarse = nullptr;
  double omega = 0.7;
};

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int velo_mg_levels(const VeloMG* mg) { return mg ? (int)mg->lv.size() : 0; }

void velo_mg_destroy(VeloMG* mg) {
  if (!mg) return;
  for (auto& L : mg->lv) {
    (void)hipFree(L.fields);
    (void)hipFree(L.d);
    (void)hipFree(L.x);
    (void)hipFree(L.t);
    (void)hipFree(L.b);
    (void)hipFree(L.r);
  }
  velo_prec_destroy(mg->coarse);
  delete mg;
}

VeloMG* velo_mg_create(const Grid& g, const Scal& s, int* err) {
  *err = 0;
  if (!is_pow2(g.nx) || !is_pow2(g.ny) || g.nx < 128 || g.ny < 64) {
    *err = DRD_E_UNSUPPORTED;
    return nullptr;
  }
  VeloMG* mg = new VeloMG();
  int nx = g.nx, ny = g.ny;
  Scal sl = s;
  for (;;) {
    MGLevel L;
    L.g = Grid::make(nx, ny);
    L.s = sl;
    mg->lv.push_back(L);
    if (nx / 2 < 64 || ny / 2 < 32) break;
    nx /= 2;
    ny /= 2;
    sl.dr *= 2.0;
    sl.dz *= 2.0;
    sl.dr2 = sl.dr * sl.dr;
    sl.dz2 = sl.dz * sl.dz;
  }
  bool ok = true;
  for (size_t l = 0; l < mg->lv.size() && ok; ++l) {
    MGLevel& L = mg->lv[l];
    const size_t n = (size_t)L.g.nuv() * sizeof(double);
    ok = hipMalloc(&L.d, n) == hipSuccess && hipMalloc(&L.t, n) == hipSuccess;
    if (ok && l + 1 < mg->lv.size()) ok = hipMalloc(&L.r, n) == hipSuccess;
    if (ok && l > 0) {
      const size_t nc = (size_t)L.g.nc() * sizeof(double);
      ok = hipMalloc(&L.fields, 3 * nc) == hipSuccess && hipMalloc(&L.x, n) == hipSuccess &&
           hipMalloc(&L.b, n) == hipSuccess;
      if (ok) {
        L.c.rho = L.fields;
        L.c.eta = L.fields + L.g.nc();
        L.c.chi = L.fields + 2 * L.g.nc();
      }
    }
  }
  if (ok) {
    const MGLevel& C = mg->lv.back();
    mg->coarse = velo_prec_create(C.g, C.s, false, err, nullptr);
    ok = mg->coarse != nullptr;
  }
  if (!ok || mg->lv.size() < 2) {
    velo_mg_destroy(mg);
    if (!*err) *err = DRD_E_HIP;
    return nullptr;
  }
  return mg;
}

// Coarse coefficients, every level's diagonal and the coarsest level's
// fast-Poisson scaling from the current fields (once per solve).
void velo_mg_setup(hipStream_t st, VeloMG* mg, const Coef& c0) {
  mg->lv[0].c = c0;
  for (size_t l = 1; l < mg->lv.size(); ++l) {
    MGLevel& L = mg->lv[l];
    launch_mg_coarsen(st, mg->lv[l - 1].g, L.g, mg->lv[l - 1].c, const_cast<double*>(L.c.rho),
                      const_cast<double*>(L.c.eta), const_cast<double*>(L.c.chi));
  }
  for (auto& L : mg->lv) launch_mg_diag(st, L.g, L.s, L.c, L.d);
  const MGLevel& C = mg->lv.back();
  velo_prec_setup(st, mg->coarse, C.g, C.s, C.c);
}

// z = V-cycle(b).  Level l smooths in two buffers: the first sweep writes
// A, the next B, the residual and the coarse correction act on B, and the
// two sweeps after it return to B (level 0: A = its temporary, B = z).
void velo_mg_apply(hipStream_t st, VeloMG* mg, const double* b, double* z, const KState* ks) {
  const int nl = (int)mg->lv.size();
  const double w = mg->omega;
  auto bufA = [&](int l) { return mg->lv[l].t; };
  auto bufB = [&](int l) { return l == 0 ? z : mg->lv[l].x; };
  auto rhs = [&](int l) { return l == 0 ? b : (const double*)mg->lv[l].b; };
  for (int l = 0; l + 1 < nl; ++l) {
    MGLevel& L = mg->lv[l];
    launch_mg_smooth(st, MG_FIRST, L.g, L.s, L.c, nullptr, rhs(l), L.d, bufA(l), w, ks);
    launch_mg_smooth(st, MG_SWEEP, L.g, L.s, L.c, bufA(l), rhs(l), L.d, bufB(l), w, ks);
    launch_mg_smooth(st, MG_RES, L.g, L.s, L.c, bufB(l), rhs(l), L.d, L.r, w, ks);
    launch_mg_restrict(st, L.g, mg->lv[l + 1].g, L.r, mg->lv[l + 1].b, ks);
  }
  {
    const MGLevel& C = mg->lv[nl - 1];
    velo_prec_apply(st, mg->coarse, C.g, C.b, C.x, ks);
  }
  for (int l = nl - 2; l >= 0; --l) {
    MGLevel& L = mg->lv[l];
    launch_mg_prolong_add(st, mg->lv[l + 1].g, L.g, mg->lv[l + 1].x, bufB(l), ks);
    launch_mg_smooth(st, MG_SWEEP, L.g, L.s, L.c, bufB(l), rhs(l), L.d, bufA(l), w, ks);
    launch_mg_smooth(st, MG_SWEEP, L.g, L.s, L.c, bufA(l), rhs(l), L.d, bufB(l), w, ks);
  }
}

}  // namespace drd
