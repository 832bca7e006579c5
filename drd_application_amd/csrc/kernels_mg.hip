// kernels_mg.hip -- geometric multigrid V-cycle preconditioner for the BDF2
// velocity operator (fast path; no reference counterpart).
//
// Level 0 is the solve's own grid; level l+1 halves level l in x and y while
// level l's sizes are even and the result keeps nx >= 64 and ny >= 32.  The coarse operators are
// rediscretised: the same MAC stencil (velo_row) with doubled spacings, on
// coefficients rho, eta, chi averaged over 2x2 cells.  One V-cycle:
//   down:  x = omega b / d, one more weighted Jacobi sweep, r = b - A x,
//          b_{l+1} = R r
//   coarsest: x = the FP64 DCT/DST fast-Poisson solve of its b (the FFT form
//             of kernels_prec.hip on a power-of-two coarsest level, the
//             dense-transform form of kernels_coarse.hip on any other of at
//             most 256 x 256 cells)
//   up:    x += P x_{l+1}, two weighted Jacobi sweeps
// with d the exact diagonal of each level's operator and omega = 0.7 on the
// finest level, 1.2 on the coarser ones: against the mass term, the viscous
// part of a level's diagonal falls fourfold per level, and over-relaxed
// sweeps there took the 2048 x 1024 two-phase channel solve from 13.7 to
// 10.6 BiCGSTAB iterations.  The cycle is a fixed map of b, so BiCGSTAB accepts it.  Every kernel skips
// once the solve it serves is done or has taken its early exit.  V(2,2)
// runs each leg of a level as one LDS-tiled launch (kernels_mgf.hip); other
// sweep counts, and V(2,2) when asked, one launch per sweep and transfer.
//
// FP32 storage: every level but the coarsest keeps its coefficients,
// diagonal and vectors in FP32 (level 0 an FP32 copy of the context's
// fields), which halves the bytes of each sweep, and evaluates its stencil
// in FP32; the input b, the output z, each sweep's update and the coarsest
// level stay FP64.
#include <vector>

#include "kernels.hpp"

namespace drd {

struct MGLevel {
  Grid g;
  Scal s;
  bool f32 = false;            // storage type of the level's buffers
  MgCoef c{};                  // rho, eta, chi of the level (FP64 level 0: the context's fields)
  void* fields = nullptr;      // own rho, eta, chi
  void *d = nullptr, *x = nullptr, *t = nullptr, *b = nullptr, *r = nullptr;  // d: omega / diagonal
};

struct VeloMG {
  std::vector<MGLevel> lv;
  CoarseSolve* csolve = nullptr;  // the coarsest solve, or
  VeloPrec* coarse = nullptr;     // its FFT form
  bool own = true;  // buffers from hipMalloc (else carved from a caller's region)
  bool f32 = false;
  double omega = 0.7;         // damped Jacobi weight on the finest level
  double omega_coarse = 1.2;  // and on the coarser levels
  int nu_pre = 2, nu_post = 2;  // Jacobi sweeps before / after the coarse correction (>= 1 each)
  bool fused = true;            // V(2,2) by the LDS-tiled legs of kernels_mgf.hip
};

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int velo_mg_levels(const VeloMG* mg) { return mg ? (int)mg->lv.size() : 0; }
bool velo_mg_f32(const VeloMG* mg) { return mg && mg->f32; }
void velo_mg_set_fused(VeloMG* mg, bool on) {
  if (mg) mg->fused = on;
}

void velo_mg_destroy(VeloMG* mg) {
  if (!mg) return;
  for (auto& L : mg->lv) {
    if (!mg->own) break;
    (void)hipFree(L.fields);
    (void)hipFree(L.d);
    (void)hipFree(L.x);
    (void)hipFree(L.t);
    (void)hipFree(L.b);
    (void)hipFree(L.r);
  }
  velo_prec_destroy(mg->coarse);
  coarse_destroy(mg->csolve);
  delete mg;
}

// Levels of grid g: level l+1 halves level l while level l's sizes are even
// and the halves keep nx >= 64 and ny >= 32.
static bool is_pow2(int n);
// the coarsest level takes the dense-transform solve (else the FFT form)
static bool mg_dense(const Grid& C) { return !(is_pow2(C.nx) && is_pow2(C.ny)) && coarse_fits(C); }
static std::vector<Grid> mg_grids(const Grid& g) {
  std::vector<Grid> v;
  int nx = g.nx, ny = g.ny;
  for (;;) {
    v.push_back(Grid::make(nx, ny));
    if (nx / 2 < 64 || ny / 2 < 32 || (nx & 1) || (ny & 1)) break;
    nx /= 2;
    ny /= 2;
  }
  return v;
}
static size_t mg_round(size_t n) { return (n + 255) / 256 * 256; }
// Bytes of every buffer of a hierarchy (levels, then the coarsest level's
// fast-Poisson buffers), each rounded to 256 B, in the order velo_mg_create
// lays them out in a caller's region.
static size_t mg_layout(const Grid& g, bool f32, const std::vector<Grid>& gs, std::vector<size_t>* sizes) {
  size_t S = 0;
  auto add = [&](size_t n) {
    if (sizes) sizes->push_back(n);
    S += mg_round(n);
  };
  const size_t nl = gs.size();
  for (size_t l = 0; l < nl; ++l) {
    const Grid& G = gs[l];
    const bool lf = f32 && l + 1 < nl;
    const size_t es = lf ? sizeof(float) : sizeof(double);
    const size_t n = (size_t)G.nuv() * es;
    add(n);                                          // d
    add(n);                                          // t
    add(l + 1 < nl ? n : 0);                         // r
    add((l > 0 || lf) ? n : 0);                      // x
    add((l > 0 || lf) ? n : 0);                      // b
    add((l > 0 || lf) ? 3 * (size_t)G.nc() * es : 0);  // fields
  }
  const Grid& C = gs.back();
  if (mg_dense(C)) {
    add(coarse_bytes(C));  // dense-transform coarsest solve: inv_s, work
  } else {
    add((size_t)C.nuv() * sizeof(double));      // fast-Poisson inv_s64
    add((size_t)C.nuv() * sizeof(float));       // inv_s32
    add(2 * (size_t)C.nuv() * sizeof(double));  // work
  }
  return S;
}

// Grids the hierarchy takes: at least two levels, and a coarsest level the
// dense-transform solve fits or a power-of-two one for the FFT form.
static bool mg_ok(const Grid& g) {
  if (g.nx < 128 || g.ny < 64 || (g.nx & 1) || (g.ny & 1)) return false;
  const Grid C = mg_grids(g).back();
  return coarse_fits(C) || (is_pow2(C.nx) && is_pow2(C.ny));
}

size_t velo_mg_bytes(const Grid& g, bool f32) {
  if (!mg_ok(g)) return 0;
  return mg_layout(g, f32, mg_grids(g), nullptr);
}

VeloMG* velo_mg_create(const Grid& g, const Scal& s, bool f32, int* err, void* region) {
  *err = 0;
  if (!mg_ok(g)) {
    *err = DRD_E_UNSUPPORTED;
    return nullptr;
  }
  VeloMG* mg = new VeloMG();
  mg->f32 = f32;
  mg->own = region == nullptr;
  const std::vector<Grid> gs = mg_grids(g);
  Scal sl = s;
  for (size_t l = 0; l < gs.size(); ++l) {
    if (l > 0) {
      sl.dr *= 2.0;
      sl.dz *= 2.0;
      sl.dr2 = sl.dr * sl.dr;
      sl.dz2 = sl.dz * sl.dz;
      scal_derive(sl);
    }
    MGLevel L;
    L.g = gs[l];
    L.s = sl;
    mg->lv.push_back(L);
  }
  const size_t nl = mg->lv.size();
  std::vector<size_t> sizes;
  mg_layout(g, f32, gs, &sizes);
  size_t k = 0;
  char* at = static_cast<char*>(region);
  bool ok = true;
  // the next buffer of the layout: from the region, or its own allocation
  auto take = [&](void** p) {
    const size_t n = sizes[k++];
    if (!n || !ok) return;
    if (region) {
      *p = at;
      at += mg_round(n);
    } else {
      ok = hipMalloc(p, n) == hipSuccess;
    }
  };
  for (size_t l = 0; l < nl && ok; ++l) {
    MGLevel& L = mg->lv[l];
    L.f32 = f32 && l + 1 < nl;
    const size_t es = L.f32 ? sizeof(float) : sizeof(double);
    take(&L.d);
    take(&L.t);
    take(&L.r);  // FP64 level 0: x is z, b is the input
    take(&L.x);
    take(&L.b);
    take(&L.fields);
    if (ok && L.fields) {
      const size_t nc = (size_t)L.g.nc() * es;
      char* f = (char*)L.fields;
      L.c = MgCoef{f, f + nc, f + 2 * nc};
    }
  }
  if (ok) {
    const MGLevel& C = mg->lv.back();
    if (mg_dense(C.g)) {
      void* cr = nullptr;
      if (region) take(&cr);
      mg->csolve = coarse_create(C.g, C.s, err, cr);
      ok = mg->csolve != nullptr;
    } else if (region) {
      VeloPrecBufs bufs{};
      take((void**)&bufs.inv_s64);
      take((void**)&bufs.inv_s32);
      take(&bufs.work);
      mg->coarse = velo_prec_create(C.g, C.s, false, err, &bufs);
      ok = mg->coarse != nullptr;
    } else {
      mg->coarse = velo_prec_create(C.g, C.s, false, err, nullptr);
      ok = mg->coarse != nullptr;
    }
  }
  if (!ok || nl < 2) {
    velo_mg_destroy(mg);
    if (!*err) *err = DRD_E_HIP;
    return nullptr;
  }
  return mg;
}

// Coarse coefficients, every level's diagonal and the coarsest level's
// fast-Poisson scaling from the current fields (once per solve).
void velo_mg_setup(hipStream_t st, VeloMG* mg, const Coef& c0) {
  MGLevel& L0 = mg->lv[0];
  if (L0.f32)
    launch_mg_convert(st, L0.g, c0, (float*)L0.c.rho, (float*)L0.c.eta, (float*)L0.c.chi);
  else
    L0.c = MgCoef{c0.rho, c0.eta, c0.chi};
  for (size_t l = 1; l < mg->lv.size(); ++l) {
    MGLevel& L = mg->lv[l];
    const MGLevel& F = mg->lv[l - 1];
    launch_mg_coarsen(st, F.f32, L.f32, F.g, L.g, F.c, const_cast<void*>(L.c.rho), const_cast<void*>(L.c.eta),
                      const_cast<void*>(L.c.chi));
  }
  for (size_t l = 0; l < mg->lv.size(); ++l) {
    MGLevel& L = mg->lv[l];
    launch_mg_diag(st, L.f32, L.g, L.s, L.c, L.d, l == 0 ? mg->omega : mg->omega_coarse);
  }
  const MGLevel& C = mg->lv.back();  // FP64
  const Coef cc{(const double*)C.c.rho, (const double*)C.c.eta, (const double*)C.c.chi};
  if (mg->csolve)
    coarse_setup(st, mg->csolve, C.g, C.s, cc);
  else
    velo_prec_setup(st, mg->coarse, C.g, C.s, cc);
}

static void mg_coarsest(hipStream_t st, VeloMG* mg, const KState* ks) {
  const MGLevel& C = mg->lv.back();  // FP64
  if (mg->csolve)
    coarse_apply(st, mg->csolve, C.g, C.s, (const double*)C.b, (double*)C.x, ks);
  else
    velo_prec_apply(st, mg->coarse, C.g, (const double*)C.b, (double*)C.x, ks);
}

// V(2,2) by one launch per level and leg (kernels_mgf.hip): the down leg of
// level l leaves its x2 in lv[l].t and the coarse b in lv[l+1].b; the up leg
// reads lv[l].t and the coarse level's x (lv[l+1].x) and writes lv[l].x, or
// z on level 0.  The FP32 hierarchy's level 0 reads the FP64 b going down
// and keeps an FP32 copy in lv[0].b for the way up.
static void velo_mg_apply_fused(hipStream_t st, VeloMG* mg, const double* b, double* z, const KState* ks) {
  const int nl = (int)mg->lv.size();
  for (int l = 0; l + 1 < nl; ++l) {
    MGLevel& L = mg->lv[l];
    const MGLevel& N = mg->lv[l + 1];
    const bool top = l == 0;
    const void* bl = top ? (const void*)b : (const void*)L.b;
    float* bcopy = top && L.f32 ? (float*)L.b : nullptr;
    launch_mg_down(st, L.f32, L.f32 && !top, N.f32, L.g, N.g, L.s, L.c, bl, L.d, L.t, bcopy, N.b, ks);
  }
  mg_coarsest(st, mg, ks);
  for (int l = nl - 2; l >= 0; --l) {
    MGLevel& L = mg->lv[l];
    const MGLevel& N = mg->lv[l + 1];
    const bool top = l == 0;
    const void* bl = (top && !L.f32) ? (const void*)b : (const void*)L.b;
    void* out = top ? (void*)z : L.x;
    launch_mg_up(st, L.f32, N.f32, L.f32 && !top, L.g, N.g, L.s, L.c, bl, L.d, L.t, N.x, out, ks);
  }
}

// z = V-cycle(b).  Level l smooths between two buffers, P (its temporary)
// and Q (its x, which the level above prolongs from; FP64 level 0: z).  The
// nu_post sweeps after the coarse correction alternate from D and must end
// in Q, so D = P for odd nu_post and Q for even; the nu_pre sweeps before it
// alternate so as to end in D, where the residual is taken and the coarse
// correction added.  Level 0 of an FP32 hierarchy ends in z, a third
// buffer, so its D is free (Q).  No sweep reads the buffer it writes.
void velo_mg_apply(hipStream_t st, VeloMG* mg, const double* b, double* z, const KState* ks) {
  if (mg->fused && mg->nu_pre == 2 && mg->nu_post == 2) {
    velo_mg_apply_fused(st, mg, b, z, ks);
    return;
  }
  const int nl = (int)mg->lv.size();
  const int n1 = mg->nu_pre, n2 = mg->nu_post;
  auto P = [&](int l) { return mg->lv[l].t; };
  auto Q = [&](int l) { return (l == 0 && !mg->lv[0].f32) ? (void*)z : mg->lv[l].x; };
  auto other = [&](int l, void* u) { return u == P(l) ? Q(l) : P(l); };
  auto D = [&](int l) { return (l == 0 && mg->lv[0].f32) || n2 % 2 == 0 ? Q(l) : P(l); };
  // level 0 of an FP32 hierarchy: its first sweep reads the FP64 b and keeps
  // an FP32 copy in lv[0].b, which the later sweeps read
  auto rhs = [&](int l) { return (l == 0 && !mg->lv[0].f32) ? (const void*)b : (const void*)mg->lv[l].b; };
  auto b32 = [&](int l) { return mg->lv[l].f32; };
  for (int l = 0; l + 1 < nl; ++l) {
    MGLevel& L = mg->lv[l];
    const bool f = L.f32;
    const bool copy = l == 0 && f;
    void* cur = (n1 - 1) % 2 == 0 ? D(l) : other(l, D(l));  // the first sweep's output
    const void* b0 = copy ? (const void*)b : rhs(l);
    float* bc = copy ? (float*)L.b : nullptr;
    int k = 1;
    // the first two sweeps in one pass, except on the FP32 hierarchy's
    // level 0, where the pass would read the FP64 b at every stencil point
    if (n1 >= 2 && !copy) {
      cur = other(l, cur);
      launch_mg_smooth(st, MG_SWEEP0, f, f && !copy, f, L.g, L.s, L.c, nullptr, b0, L.d, cur, bc, ks);
      k = 2;
    } else {
      launch_mg_smooth(st, MG_FIRST, f, f && !copy, f, L.g, L.s, L.c, nullptr, b0, L.d, cur, bc, ks);
    }
    for (; k < n1; ++k) {
      void* nxt = other(l, cur);
      launch_mg_smooth(st, MG_SWEEP, f, b32(l), f, L.g, L.s, L.c, cur, rhs(l), L.d, nxt, nullptr, ks);
      cur = nxt;
    }
    launch_mg_smooth(st, MG_RES, f, b32(l), f, L.g, L.s, L.c, cur, rhs(l), L.d, L.r, nullptr, ks);
    launch_mg_restrict(st, f, mg->lv[l + 1].f32, L.g, mg->lv[l + 1].g, L.r, mg->lv[l + 1].b, ks);
  }
  mg_coarsest(st, mg, ks);
  for (int l = nl - 2; l >= 0; --l) {
    MGLevel& L = mg->lv[l];
    const bool f = L.f32;
    void* cur = D(l);
    launch_mg_prolong_add(st, mg->lv[l + 1].f32, f, mg->lv[l + 1].g, L.g, mg->lv[l + 1].x, cur, ks);
    for (int k = 1; k <= n2; ++k) {
      const bool last = k == n2;
      void* out = last ? (l == 0 ? (void*)z : Q(l)) : other(l, cur);
      launch_mg_smooth(st, MG_SWEEP, f, b32(l), f && !(last && l == 0), L.g, L.s, L.c, cur, rhs(l), L.d, out,
                       nullptr, ks);
      cur = out;
    }
  }
}

}  // namespace drd
