// kernels_coarse.hip -- the multigrid's coarsest solve, for grids of any size.
//
// The fast-Poisson preconditioner of kernels_prec.hip, M^-1 = S^-1 A0^-1 S^-1
// with A0 the constant-coefficient velocity operator under the reference's
// ghost rules (u: DCT-I in x, DST-II in y; v: DCT-II in x, DST-I in y) and
// S = sqrt(diag A / diag A0), with its line transforms written as dense
// products with the transform matrices (scipy's unnormalised DCT/DST
// conventions; the inverses are the matching type's matrix over its scale).
// On the coarsest level (at most 256 x 256 cells) a line has at most a few
// hundred points, so O(n^2) products are cheap, and they take any length,
// where the FFTs of kernels_prec.hip need powers of two.  Three launches:
// rows forward (x), columns (y forward, divide by the eigenvalue, y
// inverse), rows inverse.  FP64 throughout.
#include <cmath>
#include <vector>

#include "kernels.hpp"

namespace drd {

struct CoarseSolve {
  int nx = 0, ny = 0;
  double a0 = 0;
  bool own = true;              // inv_s and work from hipMalloc (else a caller's region)
  double* tabs = nullptr;       // matrices and 1-D eigenvalue parts (shared by a batch group)
  double* inv_s = nullptr;      // 1 / S per unknown (per instance)
  double* work = nullptr;       // the row-transformed vector (per instance)
  // offsets into tabs
  size_t cxu_f, cxu_i, syu_f, syu_i, cxv_f, cxv_i, syv_f, syv_i, lxu, lyu, lxv, lyv;
};

struct CoarseTabs {
  const double *cxu_f, *cxu_i, *syu_f, *syu_i, *cxv_f, *cxv_i, *syv_f, *syv_i, *lxu, *lyu, *lxv, *lyv;
  double a0, dt;
};
static CoarseTabs coarse_tabs(const CoarseSolve* cs, double dt) {
  const double* t = cs->tabs;
  return CoarseTabs{t + cs->cxu_f, t + cs->cxu_i, t + cs->syu_f, t + cs->syu_i, t + cs->cxv_f, t + cs->cxv_i,
                    t + cs->syv_f, t + cs->syv_i, t + cs->lxu, t + cs->lyu, t + cs->lxv, t + cs->lyv,
                    cs->a0, dt};
}

size_t coarse_bytes(const Grid& g) { return 2 * (size_t)g.nuv() * sizeof(double); }
bool coarse_fits(const Grid& g) { return g.nx >= 2 && g.ny >= 2 && g.nx <= 256 && g.ny <= 256; }

// Transform matrices, row = output index, column = input index.
//   DCT-I  (N): y_j = x_0 + (-1)^j x_{N-1} + 2 sum_{0<i<N-1} x_i cos(pi i j / (N-1)); inverse = same / (2(N-1))
//   DCT-II (N): y_k = 2 sum_n x_n cos(pi k (2n+1) / 2N); inverse = DCT-III / 2N,
//               DCT-III: y_n = x_0 + 2 sum_{k>0} x_k cos(pi k (2n+1) / 2N)
//   DST-II (N): y_m = 2 sum_k x_k sin(pi (2k+1)(m+1) / 2N); inverse = DST-III / 2N,
//               DST-III: y_k = (-1)^k x_{N-1} + 2 sum_{m<N-1} x_m sin(pi (2k+1)(m+1) / 2N)
//   DST-I  (N): y_m = 2 sum_k x_k sin(pi (k+1)(m+1) / (N+1)); inverse = same / (2(N+1))
static void dct1_mats(int N, std::vector<double>& f, std::vector<double>& inv) {
  f.assign((size_t)N * N, 0.0);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i)
      f[(size_t)j * N + i] = i == 0 ? 1.0 : (i == N - 1 ? ((j & 1) ? -1.0 : 1.0) : 2.0 * std::cos(M_PI * i * j / (N - 1.0)));
  inv = f;
  for (double& v : inv) v /= 2.0 * (N - 1);
}
static void dct2_mats(int N, std::vector<double>& f, std::vector<double>& inv) {
  f.assign((size_t)N * N, 0.0);
  inv.assign((size_t)N * N, 0.0);
  for (int k = 0; k < N; ++k)
    for (int n = 0; n < N; ++n) {
      const double c = std::cos(M_PI * k * (2.0 * n + 1.0) / (2.0 * N));
      f[(size_t)k * N + n] = 2.0 * c;
      inv[(size_t)n * N + k] = (k == 0 ? 1.0 : 2.0 * c) / (2.0 * N);
    }
}
static void dst2_mats(int N, std::vector<double>& f, std::vector<double>& inv) {
  f.assign((size_t)N * N, 0.0);
  inv.assign((size_t)N * N, 0.0);
  for (int m = 0; m < N; ++m)
    for (int k = 0; k < N; ++k) {
      const double sn = std::sin(M_PI * (2.0 * k + 1.0) * (m + 1.0) / (2.0 * N));
      f[(size_t)m * N + k] = 2.0 * sn;
      inv[(size_t)k * N + m] = (m == N - 1 ? ((k & 1) ? -1.0 : 1.0) : 2.0 * sn) / (2.0 * N);
    }
}
static void dst1_mats(int N, std::vector<double>& f, std::vector<double>& inv) {
  f.assign((size_t)N * N, 0.0);
  for (int m = 0; m < N; ++m)
    for (int k = 0; k < N; ++k) f[(size_t)m * N + k] = 2.0 * std::sin(M_PI * (k + 1.0) * (m + 1.0) / (N + 1.0));
  inv = f;
  for (double& v : inv) v /= 2.0 * (N + 1);
}

void coarse_destroy(CoarseSolve* cs) {
  if (!cs) return;
  (void)hipFree(cs->tabs);
  if (cs->own) {
    (void)hipFree(cs->inv_s);
    (void)hipFree(cs->work);
  }
  delete cs;
}

// region: coarse_bytes(g) per instance (inv_s, then work), or nullptr
CoarseSolve* coarse_create(const Grid& g, const Scal& s, int* err, void* region) {
  *err = 0;
  if (!coarse_fits(g)) {
    *err = DRD_E_UNSUPPORTED;
    return nullptr;
  }
  const int nx = g.nx, ny = g.ny;
  CoarseSolve* cs = new CoarseSolve();
  cs->nx = nx;
  cs->ny = ny;
  cs->a0 = 1.5 * s.re;  // unit density, unit viscosity, no penalty
  std::vector<double> all, f, inv;
  auto put = [&](const std::vector<double>& v) {
    const size_t o = all.size();
    all.insert(all.end(), v.begin(), v.end());
    return o;
  };
  dct1_mats(nx + 1, f, inv);
  cs->cxu_f = put(f);
  cs->cxu_i = put(inv);
  dst2_mats(ny, f, inv);
  cs->syu_f = put(f);
  cs->syu_i = put(inv);
  dct2_mats(nx, f, inv);
  cs->cxv_f = put(f);
  cs->cxv_i = put(inv);
  dst1_mats(ny - 1, f, inv);
  cs->syv_f = put(f);
  cs->syv_i = put(inv);
  // 1-D eigenvalue parts of A0: 4 sin^2(pi k / 2n) / h^2 (u: k = 0..nx in x,
  // 1..ny in y; v: k = 0..nx-1 in x, 1..ny-1 in y)
  std::vector<double> e;
  for (int k = 0; k <= nx; ++k) e.push_back(4.0 * std::pow(std::sin(M_PI * k / (2.0 * nx)), 2) / s.dr2);
  cs->lxu = put(e);
  e.clear();
  for (int k = 1; k <= ny; ++k) e.push_back(4.0 * std::pow(std::sin(M_PI * k / (2.0 * ny)), 2) / s.dz2);
  cs->lyu = put(e);
  e.clear();
  for (int k = 0; k < nx; ++k) e.push_back(4.0 * std::pow(std::sin(M_PI * k / (2.0 * nx)), 2) / s.dr2);
  cs->lxv = put(e);
  e.clear();
  for (int k = 1; k < ny; ++k) e.push_back(4.0 * std::pow(std::sin(M_PI * k / (2.0 * ny)), 2) / s.dz2);
  cs->lyv = put(e);
  bool ok = hipMalloc(&cs->tabs, all.size() * sizeof(double)) == hipSuccess &&
            hipMemcpy(cs->tabs, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  cs->own = region == nullptr;
  const size_t n = (size_t)g.nuv() * sizeof(double);
  if (ok && region) {
    cs->inv_s = static_cast<double*>(region);
    cs->work = reinterpret_cast<double*>(static_cast<char*>(region) + n);
  } else if (ok) {
    ok = hipMalloc(&cs->inv_s, n) == hipSuccess && hipMalloc(&cs->work, n) == hipSuccess;
  }
  if (!ok) {
    coarse_destroy(cs);
    *err = DRD_E_HIP;
    return nullptr;
  }
  return cs;
}

// 1 / S = sqrt(diag(A0) / diag(A)) per unknown, diag(A) the exact diagonal
// of the level's BDF2 operator
__global__ __launch_bounds__(kBlock) void k_coarse_scale(Grid g, Scal s, const double* __restrict__ rho,
                                                         const double* __restrict__ eta,
                                                         const double* __restrict__ chi, double a0,
                                                         double* __restrict__ inv_s, int64_t bs) {
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  chi = bofs(chi, bs);
  inv_s = bofs(inv_s, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nuv()) return;
  const double d = velo_diag2(g, s, rho, eta, chi, q);
  const double d0 = q < g.mn() ? a0 + s.dt * (4.0 / s.dr2 + 2.0 / s.dz2) : a0 + s.dt * (2.0 / s.dr2 + 4.0 / s.dz2);
  inv_s[q] = sqrt(d0 / d);
}

void coarse_setup(hipStream_t st, CoarseSolve* cs, const Grid& g, const Scal& s, const Coef& c) {
  hipLaunchKernelGGL(k_coarse_scale, dim3((unsigned)((g.nuv() + kBlock - 1) / kBlock), 1, bat_n()), dim3(kBlock), 0,
                     st, g, s, c.rho, c.eta, c.chi, cs->a0, cs->inv_s, bat_bs());
}

// rows: FWD: w[row][j] = sum_i (r inv_s)[row][i] F[j][i]; else
// z[row][i] = inv_s[row][i] sum_j w[row][j] I[i][j]; one thread per output
template <bool FWD>
__global__ __launch_bounds__(kBlock) void k_coarse_rows(Grid g, CoarseTabs t, const double* __restrict__ r,
                                                        const double* __restrict__ inv_s, double* __restrict__ w,
                                                        double* __restrict__ z, const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (st && (st->done || st->early)) return;
  r = bofs(r, bs);
  inv_s = bofs(inv_s, bs);
  w = bofs(w, bs);
  z = bofs(z, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nuv()) return;
  const bool u = q < g.mn();
  const int n = u ? g.nx + 1 : g.nx;
  const int64_t e = u ? q : q - g.mn();
  const int row = (int)(e / n), col = (int)(e - (int64_t)row * n);
  const int64_t base = (u ? 0 : g.mn()) + (int64_t)row * n;
  const double* M = FWD ? (u ? t.cxu_f : t.cxv_f) : (u ? t.cxu_i : t.cxv_i);
  double acc = 0.0;
  if (FWD) {
    for (int i = 0; i < n; ++i) acc += (r[base + i] * inv_s[base + i]) * M[(int64_t)col * n + i];
    w[q] = acc;
  } else {
    for (int j = 0; j < n; ++j) acc += w[base + j] * M[(int64_t)col * n + j];
    z[q] = acc * inv_s[q];
  }
}

// columns: one workgroup per column of the row-transformed u (nx + 1 of ny
// values) and v (nx of ny - 1): forward along y, divide by
// a0 + dt (2 lx + ly) (u) or a0 + dt (lx + 2 ly) (v), inverse, in place
__global__ __launch_bounds__(kBlock) void k_coarse_cols(Grid g, CoarseTabs t, double* __restrict__ w,
                                                        const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (st && (st->done || st->early)) return;
  w = bofs(w, bs);
  __shared__ double col[256], spec[256];
  const int jc = (int)blockIdx.x;
  const bool u = jc <= g.nx;
  const int j = u ? jc : jc - (g.nx + 1);
  const int n = u ? g.ny : g.ny - 1;      // points in the column
  const int ld = u ? g.nx + 1 : g.nx;     // row length
  const int64_t base = u ? 0 : g.mn();
  for (int k = threadIdx.x; k < n; k += kBlock) col[k] = w[base + (int64_t)k * ld + j];
  __syncthreads();
  const double* F = u ? t.syu_f : t.syv_f;
  const double* I = u ? t.syu_i : t.syv_i;
  const double lx = u ? t.lxu[j] : t.lxv[j];
  for (int m = threadIdx.x; m < n; m += kBlock) {
    double acc = 0.0;
    for (int k = 0; k < n; ++k) acc += F[(int64_t)m * n + k] * col[k];
    const double ly = u ? t.lyu[m] : t.lyv[m];
    const double lam = u ? t.a0 + t.dt * (2.0 * lx + ly) : t.a0 + t.dt * (lx + 2.0 * ly);
    spec[m] = acc / lam;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += kBlock) {
    double acc = 0.0;
    for (int m = 0; m < n; ++m) acc += I[(int64_t)k * n + m] * spec[m];
    w[base + (int64_t)k * ld + j] = acc;
  }
}

void coarse_apply(hipStream_t st, CoarseSolve* cs, const Grid& g, const Scal& s, const double* r, double* z,
                  const KState* ks) {
  const CoarseTabs t = coarse_tabs(cs, s.dt);
  const dim3 rows((unsigned)((g.nuv() + kBlock - 1) / kBlock), 1, bat_n());
  hipLaunchKernelGGL(k_coarse_rows<true>, rows, dim3(kBlock), 0, st, g, t, r, cs->inv_s, cs->work, (double*)nullptr,
                     ks, bat_bs());
  hipLaunchKernelGGL(k_coarse_cols, dim3((unsigned)(2 * g.nx + 1), 1, bat_n()), dim3(kBlock), 0, st, g, t, cs->work,
                     ks, bat_bs());
  hipLaunchKernelGGL(k_coarse_rows<false>, rows, dim3(kBlock), 0, st, g, t, (const double*)nullptr, cs->inv_s,
                     cs->work, z, ks, bat_bs());
}

}  // namespace drd
