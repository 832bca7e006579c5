// kernels_phsolve.hip -- mu-only Cahn-Hilliard solve (MF/PhSolve.f90), its
// VCOST-based DCT-I preconditioner (TP/DCT.f90) and the ESOLP_x line
// preconditioner (TP/ESOLP.f90, used by PD/VeloSolve_combined.f90).
//
// VCOST (LIB/vfftpack.f:643-704) is REDFT00 / sqrt(2(n-1)).  A line of n
// points is transformed by an FFT of length 2(n-1) when that is a power of
// two, otherwise by its definition from a cosine table.  The stencil kernels
// keep the reference's FP64 evaluation order (bit-exact with the oracle).
#include <cmath>
#include <vector>

#include "fft_lines.hpp"
#include "kernels.hpp"

namespace drd {

struct Dct1Tab {
  int n = 0;           // points per line
  bool fft = false;    // 2(n-1) is a power of two
  Cx<double>* tw = nullptr;  // FFT twiddles, L = 2(n-1)
  double* cs = nullptr;      // cos(pi m/(n-1)), m < 2(n-1)
  double scale = 0;          // 1/sqrt(2(n-1))
};

Dct1Tab* dct1_table_create(int n) {
  Dct1Tab* t = new Dct1Tab();
  t->n = n;
  const int L = 2 * (n - 1);
  t->fft = L >= 2 && (L & (L - 1)) == 0;
  t->scale = 1.0 / std::sqrt(2.0 * (n - 1));
  std::vector<Cx<double>> tw(L / 2 > 0 ? L / 2 : 1);
  for (int k = 0; k < L / 2; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)L;
    tw[k] = Cx<double>{std::cos(a), std::sin(a)};
  }
  std::vector<double> cs(L > 0 ? L : 1);
  for (int m = 0; m < L; ++m) cs[m] = std::cos(M_PI * (double)m / (double)(n - 1));
  if (hipMalloc(&t->tw, tw.size() * sizeof(Cx<double>)) != hipSuccess ||
      hipMalloc(&t->cs, cs.size() * sizeof(double)) != hipSuccess ||
      hipMemcpy(t->tw, tw.data(), tw.size() * sizeof(Cx<double>), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(t->cs, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    dct1_table_destroy(t);
    return nullptr;
  }
  return t;
}

void dct1_table_destroy(Dct1Tab* t) {
  if (!t) return;
  (void)hipFree(t->tw);
  (void)hipFree(t->cs);
  delete t;
}

struct D1 {  // by-value view for kernels
  int n;
  bool fft;
  const Cx<double>* tw;
  const double* cs;
  double scale;
};
static D1 view(const Dct1Tab* t) { return D1{t->n, t->fft, t->tw, t->cs, t->scale}; }

// VCOST of the line (LDS) in place; A, B complex scratch of 2(n-1)
__device__ void vcost_line(double* line, const D1& t, Cx<double>* A, Cx<double>* B) {
  const int n = t.n;
  if (t.fft) {
    line_transform<double>(TR_DCT1, line, n, A, B, t.tw, nullptr);
  } else {
    // X_i = x_0 + (-1)^i x_{n-1} + 2 sum_{k=1}^{n-2} x_k cos(k i pi/(n-1))
    double* out = reinterpret_cast<double*>(A);
    const int L = 2 * (n - 1);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      double acc = line[0] + ((i & 1) ? -line[n - 1] : line[n - 1]);
      for (int k = 1; k < n - 1; ++k) acc += 2.0 * line[k] * t.cs[((long)k * i) % L];
      out[i] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = out[i];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = line[i] * t.scale;
  __syncthreads();
}

static size_t vcost_lds(int n) {
  return 2 * (size_t)(2 * n) * sizeof(Cx<double>) + (size_t)(n + 1) * sizeof(double);
}

// lines l < nlines of length t.n: in[l*ld_in + i] -> VCOST -> out[l*ld_out + i],
// one line per workgroup of kVcostThreads threads (each radix-2 pass of a
// line up to 2048 points is one butterfly per thread; the same butterflies in
// the same order as with fewer threads)
constexpr int kVcostThreads = 1024;
__global__ __launch_bounds__(kVcostThreads) void k_vcost_lines(D1 t, int nlines, const double* in, int64_t ld_in,
                                                        double* out, int64_t ld_out, const double* cof,
                                                        int64_t ld_cof, const KState* st) {
  if (st && (st->done || st->early)) return;
  extern __shared__ unsigned char smem_raw[];
  const int n = t.n;
  Cx<double>* A = reinterpret_cast<Cx<double>*>(smem_raw);
  Cx<double>* B = A + 2 * n;
  double* line = reinterpret_cast<double*>(B + 2 * n);
  const int l = blockIdx.x;
  for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = in[(int64_t)l * ld_in + i];
  __syncthreads();
  vcost_line(line, t, A, B);
  if (cof) {  // DCT TP/DCT.f90:22-26 between the two pairs of passes
    for (int i = threadIdx.x; i < n; i += blockDim.x) line[i] = line[i] / cof[(int64_t)l + ld_cof * i];
    __syncthreads();
    vcost_line(line, t, A, B);
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[(int64_t)l * ld_out + i] = line[i];
}

__global__ __launch_bounds__(kBlock) void k_transpose_d(int rows, int cols, const double* __restrict__ in,
                                                        double* __restrict__ out, const KState* st) {
  if (st && (st->done || st->early)) return;
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    if (r < rows && c < cols) tile[j][tx] = in[(size_t)r * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = bx + j, c = by + tx;
    if (r < cols && c < rows) out[(size_t)r * rows + c] = tile[tx][j];
  }
}

static void vcost_launch(hipStream_t s, const Dct1Tab* t, int nlines, const double* in, int64_t ld_in, double* out,
                         int64_t ld_out, const double* cof, int64_t ld_cof, const KState* ks) {
  static const bool raised = [] {  // once; concurrent first callers wait for it
    (void)hipFuncSetAttribute((const void*)k_vcost_lines, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    return true;
  }();
  (void)raised;
  hipLaunchKernelGGL(k_vcost_lines, dim3(nlines), dim3(kVcostThreads), vcost_lds(t->n), s, view(t), nlines, in, ld_in, out,
                     ld_out, cof, ld_cof, ks);
}

static void transpose_launch(hipStream_t s, int rows, int cols, const double* in, double* out, const KState* ks) {
  hipLaunchKernelGGL(k_transpose_d, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(kBlock), 0, s, rows, cols, in,
                     out, ks);
}

// PSolve_CG MF/PhSolve.f90:296-313 = DCT(nr-1,nz-1,...,Ph_Lcc) TP/DCT.f90:1-57
// on the packed nx*ny vector: VCOST along x and y, divide by Ph_Lcc, again.
void launch_ph_dct_prec(hipStream_t s, const Grid& g, const Dct1Tab* tx, const Dct1Tab* ty, const double* ph_lcc,
                        const double* b, double* x, double* work, const KState* ks) {
  const int nx = g.nx, ny = g.ny;
  double* w1 = work;
  double* w2 = work + (size_t)nx * ny;
  vcost_launch(s, tx, ny, b, nx, w1, nx, nullptr, 0, ks);        // rows along x
  transpose_launch(s, ny, nx, w1, w2, ks);                        // -> column kx is contiguous
  vcost_launch(s, ty, nx, w2, ny, w2, ny, ph_lcc, nx, ks);        // y, / Ph_Lcc(kx,ky), y
  transpose_launch(s, nx, ny, w2, w1, ks);
  vcost_launch(s, tx, ny, w1, nx, x, nx, nullptr, 0, ks);         // rows along x
}

// ---- PhSolve stencils -----------------------------------------------------------
// op_double_gradient_Neum TP/matrix_tool.f90:178-219 reading interior values
template <class F>
__device__ __forceinline__ double dgn(const Grid& g, double dx2, double dy2, const double* __restrict__ mo, int i,
                                      int j, F get) {
  const int nx = g.nx, ny = g.ny;
  const double m0 = mo[g.ci(i, j)];
  const double r0 = get(i, j);
  double moL, moR, rhs = 0.0;
  if (i == 0) {
    moL = (m0 + m0) / 2.0;
    moR = (m0 + mo[g.ci(i + 1, j)]) / 2.0;
    rhs = rhs + 1.0 / dx2 * (moL * r0 + moR * get(i + 1, j) - (moL + moR) * r0);
  } else if (i == nx - 1) {
    moL = (mo[g.ci(i - 1, j)] + m0) / 2.0;
    moR = (m0 + m0) / 2.0;
    rhs = rhs + 1.0 / dx2 * (moL * get(i - 1, j) + moR * r0 - (moL + moR) * r0);
  } else {
    moL = (mo[g.ci(i - 1, j)] + m0) / 2.0;
    moR = (m0 + mo[g.ci(i + 1, j)]) / 2.0;
    rhs = rhs + 1.0 / dx2 * (moL * get(i - 1, j) + moR * get(i + 1, j) - (moL + moR) * r0);
  }
  if (j == 0) {
    moL = (m0 + m0) / 2.0;
    moR = (m0 + mo[g.ci(i, j + 1)]) / 2.0;
    rhs = rhs + 1.0 / dy2 * (moL * r0 + moR * get(i, j + 1) - (moL + moR) * r0);
  } else if (j == ny - 1) {
    moL = (mo[g.ci(i, j - 1)] + m0) / 2.0;
    moR = (m0 + m0) / 2.0;
    rhs = rhs + 1.0 / dy2 * (moL * get(i, j - 1) + moR * r0 - (moL + moR) * r0);
  } else {
    moL = (mo[g.ci(i, j - 1)] + m0) / 2.0;
    moR = (m0 + mo[g.ci(i, j + 1)]) / 2.0;
    rhs = rhs + 1.0 / dy2 * (moL * get(i, j - 1) + moR * get(i, j + 1) - (moL + moR) * r0);
  }
  return rhs;
}

// ch_bulk_energy MF/ch_bulk_energy.f90:4-28 (phi**2.0 evaluated as phi*phi)
__device__ __forceinline__ double chb_mf(double phi, double psi) {
  double aa = (phi * phi) * ((1.0 - phi) * (1.0 - phi));
  aa = 1.0 - sqrt(aa / (aa + 1.0e-2)) * psi;
  return 0.5 * phi * (1.0 - phi) * (1.0 - 2.0 * phi) * aa;
}

// f2 = ph - dt u.grad(ph) over the cell field (MF/PhSolve.f90:26-34)
__global__ __launch_bounds__(kBlock) void k_ph_f2(Grid g, Scal s, const double* __restrict__ ph,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  double* __restrict__ f2) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nc()) return;
  const int nx = g.nx, ny = g.ny;
  const int k = (int)(q / (nx + 2)) - 1, i = (int)(q - (int64_t)(k + 1) * (nx + 2)) - 1;
  double f = ph[q];
  if (i >= 0 && i < nx && k >= 0 && k < ny)
    f = f - s.dt * (u[g.ui(i + 1, k)] + u[g.ui(i, k)]) / 2.0 * (ph[g.ci(i + 1, k)] - ph[g.ci(i - 1, k)]) / 2.0 / s.dr -
        s.dt * (v[g.vi(i, k + 1)] + v[g.vi(i, k)]) / 2.0 * (ph[g.ci(i, k + 1)] - ph[g.ci(i, k - 1)]) / 2.0 / s.dz;
  f2[q] = f;
}

// rhs = -f1 + s f2/eps - eps L_{M_in} f2 ; x0 = rmu   (MF/PhSolve.f90:37-63)
__global__ __launch_bounds__(kBlock) void k_ph_rhs(Grid g, Scal s, const double* __restrict__ ph,
                                                   const double* __restrict__ f2, const double* __restrict__ bdr,
                                                   const double* __restrict__ bg, const double* __restrict__ minner,
                                                   const double* __restrict__ rmu, double cosa, double sqrt2,
                                                   double* __restrict__ rhs, double* __restrict__ x0) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.ncell()) return;
  const int nx = g.nx;
  const int k = g.by_nx.div(q), i = (int)(q - (int64_t)k * nx);
  const int64_t c = g.ci(i, k);
  const double phv = ph[c];
  const double f1 = s.s * phv / s.I - chb_mf(phv, bdr[c]) / s.I - phv * (phv - 1.0) * bg[c] * cosa / sqrt2;
  const double ff = dgn(g, s.dr2, s.dz2, minner, i, k, [&](int a, int b) { return f2[g.ci(a, b)]; });
  rhs[q] = -f1 + s.s * f2[c] / s.I - s.I * ff;
  x0[q] = rmu[c];
}

// MATRIX_VECTOR_CG MF/PhSolve.f90:249-277, stage 1: t = -dt M L_{M_0} x
__global__ __launch_bounds__(kBlock) void k_ph_mv1(Grid g, Scal s, const double* __restrict__ m0,
                                                   const double* __restrict__ x, double* __restrict__ t,
                                                   const KState* st) {
  if (st && (st->done || st->early)) return;
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.ncell()) return;
  const int nx = g.nx;
  const int k = g.by_nx.div(q), i = (int)(q - (int64_t)k * nx);
  const double ff = dgn(g, s.dr2, s.dz2, m0, i, k, [&](int a, int b) { return x[a + (int64_t)nx * b]; });
  t[q] = -s.dt * s.xld * ff;
}

// stage 2: y = alpha (x + s t/eps - eps L_{M_in} t) + beta y, with the
// BiCGSTAB fusions of k_velo_matvec
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_ph_mv2(Grid g, Scal s, const double* __restrict__ minner,
                                                   const double* __restrict__ x, const double* __restrict__ t,
                                                   double* __restrict__ y, double alpha, double beta,
                                                   double* __restrict__ a0, const double* __restrict__ b,
                                                   double* part, KState* st) {
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
    if (MODE != MV_INIT || xnz) {
      double ff2 = dgn(g, s.dr2, s.dz2, minner, i, k, [&](int a, int bb) { return t[a + (int64_t)nx * bb]; });
      ff2 = s.s * t[q] / s.I - s.I * ff2;
      z = x[q] + ff2;
    }
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

// rmu_after with Neumann ghosts and ph_after = f2 + dt M L_{M_0} rmu_after,
// ghosts copied from the nearest interior cell (MF/PhSolve.f90:85-124)
__global__ __launch_bounds__(kBlock) void k_ph_finish(Grid g, Scal s, const double* __restrict__ x,
                                                      const double* __restrict__ f2, const double* __restrict__ m0,
                                                      double* __restrict__ mu, double* __restrict__ ph) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nc()) return;
  const int nx = g.nx, ny = g.ny;
  const int k = (int)(q / (nx + 2)) - 1, i = (int)(q - (int64_t)(k + 1) * (nx + 2)) - 1;
  const int ic = i < 0 ? 0 : (i >= nx ? nx - 1 : i);
  const int kc = k < 0 ? 0 : (k >= ny ? ny - 1 : k);
  mu[q] = x[ic + (int64_t)nx * kc];
  const double ff = dgn(g, s.dr2, s.dz2, m0, ic, kc, [&](int a, int b) { return x[a + (int64_t)nx * b]; });
  ph[q] = f2[g.ci(ic, kc)] + s.dt * s.xld * ff;
}

static inline unsigned nbk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

void launch_ph_rhs(hipStream_t st, const Grid& g, const Scal& s, const double* ph, const double* u, const double* v,
                   const double* bdr, const double* bg, const double* minner, const double* rmu, double cosa,
                   double sqrt2, double* f2, double* rhs, double* x0) {
  hipLaunchKernelGGL(k_ph_f2, dim3(nbk(g.nc(), kBlock)), dim3(kBlock), 0, st, g, s, ph, u, v, f2);
  hipLaunchKernelGGL(k_ph_rhs, dim3(nbk(g.ncell(), kBlock)), dim3(kBlock), 0, st, g, s, ph, f2, bdr, bg, minner,
                     rmu, cosa, sqrt2, rhs, x0);
}

void launch_ph_matvec(hipStream_t st, int mode, const Grid& g, const Scal& s, const double* m0,
                      const double* minner, const double* x, double* t, double* y, double alpha, double beta,
                      double* a0, const double* b, double* part, KState* ks) {
  hipLaunchKernelGGL(k_ph_mv1, dim3(nbk(g.ncell(), kBlock)), dim3(kBlock), 0, st, g, s, m0, x, t,
                     mode == MV_PLAIN ? nullptr : ks);
  dim3 grid(nbk(g.ncell(), kTile));
#define PM2(M) \
  hipLaunchKernelGGL((k_ph_mv2<M>), grid, dim3(kBlock), 0, st, g, s, minner, x, t, y, alpha, beta, a0, b, part, ks)
  switch (mode) {
    case MV_PLAIN: PM2(MV_PLAIN); break;
    case MV_DOT1: PM2(MV_DOT1); break;
    case MV_DOT2: PM2(MV_DOT2); break;
    case MV_INIT: PM2(MV_INIT); break;
  }
#undef PM2
  launch_mv_finalize(st, mode, part, grid.x, 1, ks);
}

void launch_ph_finish(hipStream_t st, const Grid& g, const Scal& s, const double* x, const double* f2,
                      const double* m0, double* mu, double* ph) {
  hipLaunchKernelGGL(k_ph_finish, dim3(nbk(g.nc(), kBlock)), dim3(kBlock), 0, st, g, s, x, f2, m0, mu, ph);
}

// ---- ESOLP_x (TP/ESOLP.f90:1-82) on both velocity blocks --------------------------
// Tridiagonal solve along y per x-mode (rfords TP/setup.f90:327-354); one
// thread per mode, rows read coalesced across modes.  bw, fw: scratch of
// the same layout as w.
__global__ __launch_bounds__(kBlock) void k_esolp_thomas(Grid g, double* __restrict__ w,
                                                         const double* __restrict__ ta, const double* __restrict__ tb,
                                                         const double* __restrict__ tc, double* __restrict__ bw,
                                                         double* __restrict__ fw, const KState* st) {
  if (st && (st->done || st->early)) return;
  const int nx = g.nx, ny = g.ny;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const bool is_u = t <= nx;
  const int kx = is_u ? t : t - (nx + 1);
  if (!is_u && kx >= nx) return;
  const int n = is_u ? ny : ny - 1;      // system size
  const int ld = is_u ? nx + 1 : nx;     // row stride of this block in the packed vector
  const int64_t base = is_u ? 0 : g.mn();
  // tables are (1:n, 0:ld-1) column-major: entry (j, kx) at (j-1) + n*kx
  const int tab_off = is_u ? 0 : (ny) * (nx + 1);
  const double* a = ta + tab_off + (int64_t)n * kx;
  const double* b = tb + tab_off + (int64_t)n * kx;
  const double* c = tc + tab_off + (int64_t)n * kx;
  auto W = [&](int j) -> int64_t { return base + (int64_t)j * ld + kx; };
  double bp = b[0], fp = w[W(0)];
  bw[W(0)] = bp;
  fw[W(0)] = fp;
  for (int j = 1; j < n; ++j) {
    const double r = -a[j] / bp;
    const double bj = b[j] + r * c[j - 1];
    const double fj = w[W(j)] + r * fp;
    bw[W(j)] = bj;
    fw[W(j)] = fj;
    bp = bj;
    fp = fj;
  }
  double xn = fabs(bp) <= 1.0e-12 ? 0.0 : fp / bp;
  w[W(n - 1)] = xn;
  for (int j = n - 2; j >= 0; --j) {
    xn = (fw[W(j)] - c[j] * xn) / bw[W(j)];
    w[W(j)] = xn;
  }
}

// b (packed Krylov vector) -> VCOST_x -> Thomas_y -> VCOST_x -> x
void launch_esolp(hipStream_t s, const Grid& g, const Dct1Tab* tu, const Dct1Tab* tv, const double* tabs,
                  const double* b, double* x, double* work, const KState* ks) {
  const int nx = g.nx, ny = g.ny;
  const int64_t N = g.nuv();
  double* w = work;
  double* bw = work + N;
  double* fw = work + 2 * N;
  const int64_t ntab = (int64_t)ny * (nx + 1) + (int64_t)(ny - 1) * nx;
  const double *ta = tabs, *tb = tabs + ntab, *tc = tabs + 2 * ntab;
  vcost_launch(s, tu, ny, b, nx + 1, w, nx + 1, nullptr, 0, ks);
  vcost_launch(s, tv, ny - 1, b + g.mn(), nx, w + g.mn(), nx, nullptr, 0, ks);
  hipLaunchKernelGGL(k_esolp_thomas, dim3(nbk(2 * nx + 1, kBlock)), dim3(kBlock), 0, s, g, w, ta, tb, tc, bw, fw, ks);
  vcost_launch(s, tu, ny, w, nx + 1, x, nx + 1, nullptr, 0, ks);
  vcost_launch(s, tv, ny - 1, w + g.mn(), nx, x + g.mn(), nx, nullptr, 0, ks);
}

}  // namespace drd
