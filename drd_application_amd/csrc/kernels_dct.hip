// kernels_dct.hip -- DCT-II / DCT-III (FFTW REDFT10 / REDFT01 convention)
// and the pressure projection of TP/PreSolve_2.f90 + TP/dct_fftw.f90.
//
// A 2-D transform is done as line transforms along contiguous rows with a
// tiled LDS transpose in between:
//   rows(x): div(u,v) -> DCT-II          [k_rows<R_DIV_DCT2>]
//   transpose
//   rows(y): DCT-II -> / Pre_Lcc -> DCT-III   (one LDS round trip) [R_SOLVE]
//   transpose
//   rows(x): DCT-III -> /(4 nx ny) -> *coef -> ghosts, p += psi   [R_DCT3_FINISH]
// A line of length N (power of two, N <= 4096) is transformed in LDS by a
// complex radix-2 Stockham FFT of length N after Makhoul's even/odd
// reordering; other lengths use the O(N^2) definition from a cosine table.
#include <cmath>
#include <vector>

#include "kernels.hpp"

namespace drd {

struct DctPlan {
  int nx, ny;
  // per length L in {nx, ny}: FFT twiddles exp(-2 pi i t/L), t < L/2, and
  // Makhoul rotations exp(-i pi k/(2L)), k < L; cos table cos(pi m/(2L)), m < 4L
  double2 *tw_x, *tw_y, *rot_x, *rot_y;
  double *cos_x, *cos_y;
  bool pow2_x, pow2_y;
};

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void make_tables(int L, double2** tw, double2** rot, double** cs) {
  std::vector<double2> t(L / 2 > 0 ? L / 2 : 1), r(L);
  std::vector<double> c(4 * (size_t)L);
  for (int k = 0; k < L / 2; ++k) {
    // angle reduced exactly before the libm call
    const double a = -2.0 * M_PI * (double)k / (double)L;
    t[k] = make_double2(cos(a), sin(a));
  }
  for (int k = 0; k < L; ++k) {
    const double a = -M_PI * (double)k / (2.0 * L);
    r[k] = make_double2(cos(a), sin(a));
  }
  for (int m = 0; m < 4 * L; ++m) c[m] = cos(M_PI * (double)m / (2.0 * L));
  hipMalloc(tw, t.size() * sizeof(double2));
  hipMalloc(rot, r.size() * sizeof(double2));
  hipMalloc(cs, c.size() * sizeof(double));
  hipMemcpy(*tw, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice);
  hipMemcpy(*rot, r.data(), r.size() * sizeof(double2), hipMemcpyHostToDevice);
  hipMemcpy(*cs, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice);
}

DctPlan* dct_plan_create(int nx, int ny) {
  DctPlan* p = new DctPlan();
  p->nx = nx;
  p->ny = ny;
  p->pow2_x = is_pow2(nx) && nx <= 4096;
  p->pow2_y = is_pow2(ny) && ny <= 4096;
  make_tables(nx, &p->tw_x, &p->rot_x, &p->cos_x);
  make_tables(ny, &p->tw_y, &p->rot_y, &p->cos_y);
  return p;
}

void dct_plan_destroy(DctPlan* p) {
  if (!p) return;
  hipFree(p->tw_x);
  hipFree(p->tw_y);
  hipFree(p->rot_x);
  hipFree(p->rot_y);
  hipFree(p->cos_x);
  hipFree(p->cos_y);
  delete p;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Stockham radix-2 FFT of length N held in LDS buffer a (b is scratch);
// sign -1 forward, +1 inverse (unnormalised).  Returns the buffer holding
// the result.  All threads of the workgroup take part.
__device__ double2* lds_fft(double2* a, double2* b, int N, const double2* __restrict__ tw, int sign) {
  const int half = N >> 1;
  for (int ns = 1; ns < N; ns <<= 1) {
    const int stride = N / (2 * ns);
    for (int j = threadIdx.x; j < half; j += blockDim.x) {
      const int k = j & (ns - 1);
      double2 w = tw[k * stride];
      if (sign > 0) w.y = -w.y;
      const double2 v0 = a[j];
      const double2 v1 = cmul(a[j + half], w);
      const int o = ((j - k) << 1) + k;
      b[o] = make_double2(v0.x + v1.x, v0.y + v1.y);
      b[o + ns] = make_double2(v0.x - v1.x, v0.y - v1.y);
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
  }
  return a;
}

// A line per workgroup of kRowThreads threads, kMaxPer values each (N <=
// 4096): 1024 threads give each radix-2 pass one butterfly per thread and
// 16 waves per line to hide the passes' LDS latency (the same butterflies in
// the same order as with fewer threads).
constexpr int kRowThreads = 1024, kRowWaves = kRowThreads / 64;
constexpr int kMaxPer = 4096 / kRowThreads;  // line values per thread

__device__ __forceinline__ int makhoul(int j, int N) { return (j & 1) ? (N - 1 - (j >> 1)) : (j >> 1); }

// REDFT10 of a line held in registers (r[m] = x[threadIdx.x + m*kRowThreads]).
__device__ void dct2_regs(double (&r)[kMaxPer], int N, bool pow2, double2* A, double2* B,
                          const double2* __restrict__ tw, const double2* __restrict__ rot,
                          const double* __restrict__ cs) {
  if (pow2 && N >= 2) {
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int j = threadIdx.x + m * kRowThreads;
      if (j < N) A[makhoul(j, N)] = make_double2(r[m], 0.0);
    }
    __syncthreads();
    const double2* V = lds_fft(A, B, N, tw, -1);
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int k = threadIdx.x + m * kRowThreads;
      if (k < N) {
        const double2 w = rot[k], v = V[k];
        r[m] = 2.0 * (w.x * v.x - w.y * v.y);
      }
    }
    __syncthreads();
  } else {
    // definition: Y_k = 2 sum_j x_j cos(pi (2j+1) k / (2N))
    double* xs = reinterpret_cast<double*>(A);
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int j = threadIdx.x + m * kRowThreads;
      if (j < N) xs[j] = r[m];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int k = threadIdx.x + m * kRowThreads;
      if (k < N) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) s += xs[j] * cs[((long)(2 * j + 1) * k) % (4L * N)];
        r[m] = 2.0 * s;
      }
    }
    __syncthreads();
  }
}

// REDFT01 of a line held in registers.
__device__ void dct3_regs(double (&r)[kMaxPer], int N, bool pow2, double2* A, double2* B,
                          const double2* __restrict__ tw, const double2* __restrict__ rot,
                          const double* __restrict__ cs) {
  double* ys = reinterpret_cast<double*>(B);
#pragma unroll
  for (int m = 0; m < kMaxPer; ++m) {
    const int k = threadIdx.x + m * kRowThreads;
    if (k < N) ys[k] = r[m];
  }
  __syncthreads();
  if (pow2 && N >= 2) {
    // V'_k = exp(+i pi k/(2N)) (Y_k - i Y_{N-k}), Y_N = 0; unnormalised
    // inverse FFT; x_{2n} = v_n, x_{2n+1} = v_{N-1-n}
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int k = threadIdx.x + m * kRowThreads;
      if (k < N) {
        const double a = ys[k];
        const double b = (k == 0) ? 0.0 : -ys[N - k];
        const double2 w = rot[k];
        A[k] = make_double2(w.x * a + w.y * b, -w.y * a + w.x * b);
      }
    }
    __syncthreads();
    const double2* v = lds_fft(A, B, N, tw, +1);
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int j = threadIdx.x + m * kRowThreads;
      if (j < N) r[m] = v[makhoul(j, N)].x;
    }
    __syncthreads();
  } else {
    // definition: X_k = Y_0 + 2 sum_{j>=1} Y_j cos(pi j (2k+1) / (2N))
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int k = threadIdx.x + m * kRowThreads;
      if (k < N) {
        double s = 0.0;
        for (int j = 1; j < N; ++j) s += ys[j] * cs[((long)j * (2 * k + 1)) % (4L * N)];
        r[m] = ys[0] + 2.0 * s;
      }
    }
    __syncthreads();
  }
}

// Real DFT of a line held in registers, in FFTW's halfcomplex order (R2HC):
// r_k = Re X_k for k <= N/2, r_{N-k} = Im X_k for 0 < k < N/2, X_k = sum_j
// x_j exp(-2 pi i jk/N), as the complex FFT of (x, 0).  N a power of two.
__device__ void r2hc_regs(double (&r)[kMaxPer], int N, double2* A, double2* B, const double2* __restrict__ tw) {
#pragma unroll
  for (int m = 0; m < kMaxPer; ++m) {
    const int j = threadIdx.x + m * kRowThreads;
    if (j < N) A[j] = make_double2(r[m], 0.0);
  }
  __syncthreads();
  const double2* V = lds_fft(A, B, N, tw, -1);
#pragma unroll
  for (int m = 0; m < kMaxPer; ++m) {
    const int k = threadIdx.x + m * kRowThreads;
    if (k < N) r[m] = k <= N / 2 ? V[k].x : V[N - k].y;
  }
  __syncthreads();
}

// HC2R, the unnormalised inverse of r2hc_regs: x_j = sum_k X_k exp(+2 pi i
// jk/N) of the Hermitian X the halfcomplex line holds.
__device__ void hc2r_regs(double (&r)[kMaxPer], int N, double2* A, double2* B, const double2* __restrict__ tw) {
  double* ys = reinterpret_cast<double*>(B);
#pragma unroll
  for (int m = 0; m < kMaxPer; ++m) {
    const int k = threadIdx.x + m * kRowThreads;
    if (k < N) ys[k] = r[m];
  }
  __syncthreads();
  const int h = N / 2;
#pragma unroll
  for (int m = 0; m < kMaxPer; ++m) {
    const int k = threadIdx.x + m * kRowThreads;
    if (k < N) {
      double2 v;
      if (k == 0 || k == h)
        v = make_double2(ys[k], 0.0);
      else if (k < h)
        v = make_double2(ys[k], ys[N - k]);
      else
        v = make_double2(ys[N - k], -ys[k]);
      A[k] = v;
    }
  }
  __syncthreads();
  const double2* V = lds_fft(A, B, N, tw, +1);
#pragma unroll
  for (int m = 0; m < kMaxPer; ++m) {
    const int j = threadIdx.x + m * kRowThreads;
    if (j < N) r[m] = V[j].x;
  }
  __syncthreads();
}

enum { R_DCT2 = 0, R_DCT3, R_DIV_DCT2, R_SOLVE, R_DCT3_FINISH };

struct RowArgs {
  int N;       // line length
  int rows;    // number of lines
  const double* in;
  double* out;  // may alias in
  const double2 *tw, *rot;
  const double* cs;
  bool pow2;
  int per;  // periodic x: the x passes of R_DIV_DCT2 / R_DCT3_FINISH are R2HC / HC2R
  // R_DIV_DCT2
  Grid g;
  const double *u, *v;
  double dr, dz;
  // R_SOLVE: line index = kx, eigenvalues pre[kx + (nx+1)*ky]
  const double* pre;
  // R_DCT3_FINISH
  double c4, coef;
  double *psi, *p;
  double* part;
  KState* ks;
};

// One workgroup per line; dynamic LDS = 2N complex.
template <int MODE>
__global__ __launch_bounds__(kRowThreads) void k_rows(RowArgs a, int64_t bs) {
  a.in = bofs(a.in, bs);  // the tables (tw, rot, cs, pre) are shared by the group
  a.out = bofs(a.out, bs);
  a.u = bofs(a.u, bs);
  a.v = bofs(a.v, bs);
  a.psi = bofs(a.psi, bs);
  a.p = bofs(a.p, bs);
  a.part = bofs(a.part, bs);
  a.ks = bofs(a.ks, bs);
  extern __shared__ double2 smem[];
  const int N = a.N;
  const int row = blockIdx.x;
  double2* A = smem;
  double2* B = smem + N;
  double r[kMaxPer];
#pragma unroll
  for (int m = 0; m < kMaxPer; ++m) {
    const int i = threadIdx.x + m * kRowThreads;
    r[m] = 0.0;
    if (i < N) {
      if (MODE == R_DIV_DCT2) {
        // uin(i,k) TP/PreSolve_2.f90:17-21
        const Grid& g = a.g;
        const int k = row;
        const double d =
            (a.u[g.ui(i + 1, k)] - a.u[g.ui(i, k)]) / a.dr + (a.v[g.vi(i, k + 1)] - a.v[g.vi(i, k)]) / a.dz;
        r[m] = d;
      } else {
        r[m] = a.in[(size_t)row * N + i];
      }
    }
  }
  // canonical (pairwise) sum of the row's divergence, taken before the
  // transform overwrites it; element i = m*kRowThreads + t lies in 64-chunk
  // m*kRowWaves + wave.
  double rowsum = 0.0;
  if (MODE == R_DIV_DCT2 && a.ks) {
    __shared__ double lc[kMaxPer * kRowWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const double c = wave_pairwise(r[m]);
      if (lane == 0) lc[m * kRowWaves + w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) rowsum = pairwise_fixed<kMaxPer * kRowWaves>(lc);
    __syncthreads();
  }
  if (MODE == R_DIV_DCT2 && a.per)
    r2hc_regs(r, N, A, B, a.tw);
  else if (MODE == R_DCT2 || MODE == R_DIV_DCT2 || MODE == R_SOLVE)
    dct2_regs(r, N, a.pow2, A, B, a.tw, a.rot, a.cs);
  if (MODE == R_SOLVE) {
    // rhs_wk / pre_cof TP/dct_fftw.f90:22-26 (this line is column kx = row)
    const int ldp = a.g.nx + 1;
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int ky = threadIdx.x + m * kRowThreads;
      if (ky < N) r[m] = r[m] / a.pre[row + (size_t)ldp * ky];
    }
  }
  if (MODE == R_DCT3_FINISH && a.per)
    hc2r_regs(r, N, A, B, a.tw);
  else if (MODE == R_DCT3 || MODE == R_SOLVE || MODE == R_DCT3_FINISH)
    dct3_regs(r, N, a.pow2, A, B, a.tw, a.rot, a.cs);
  if (MODE == R_DCT3_FINISH) {
    // psi = rhs_wk/coef (dct_fftw.f90:32-36), Neumann ghosts (:43-51),
    // psi *= coef (PreSolve_2.f90:125), p = psi + p (TP/mcl6.f90:502)
    const Grid& g = a.g;
    const int k = row, nx = g.nx, ny = g.ny;
    auto put = [&](int i, int kk, double val) {
      const int64_t c = g.ci(i, kk);
      const double ps = val * a.coef;
      a.psi[c] = ps;
      if (a.p) a.p[c] = ps + a.p[c];
    };
    // stage the line in LDS so the ghost columns can read their neighbour
    double* xs = reinterpret_cast<double*>(A);
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int i = threadIdx.x + m * kRowThreads;
      if (i < N) xs[i] = r[m];
    }
    __syncthreads();
    for (int i = (int)threadIdx.x - 1; i <= nx; i += blockDim.x) {
      const int ic = g.xcol(i);  // Neumann or periodic side ghosts
      const double val = xs[ic] / a.c4;
      put(i, k, val);
      if (k == 0) put(i, -1, val);
      if (k == ny - 1) put(i, ny, val);
    }
  } else {
#pragma unroll
    for (int m = 0; m < kMaxPer; ++m) {
      const int i = threadIdx.x + m * kRowThreads;
      if (i < N) a.out[(size_t)row * N + i] = r[m];
    }
  }
  if (MODE == R_DIV_DCT2 && a.ks && threadIdx.x == 0) a.part[row] = rowsum;  // rows summed by k_finalize
}

// out (cols x rows) = transpose of in (rows x cols), 32x32 LDS tiles
__global__ __launch_bounds__(kBlock) void k_transpose(int rows, int cols, const double* __restrict__ in,
                                                      double* __restrict__ out, int64_t bs) {
  in = bofs(in, bs);
  out = bofs(out, bs);
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    if (r < rows && c < cols) tile[j][tx] = in[(size_t)r * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = bx + j, c = by + tx;  // out row = in col
    if (r < cols && c < rows) out[(size_t)r * rows + c] = tile[tx][j];
  }
}

static size_t row_lds(int N) { return 2 * (size_t)N * sizeof(double2); }

template <int MODE>
static void launch_rows(hipStream_t st, const RowArgs& a) {
  // lines of 2048+ need more than 64 KiB of LDS; raised once (a function-
  // local static: concurrent first callers wait for it)
  static const bool raised = [] {
    (void)hipFuncSetAttribute((const void*)k_rows<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)row_lds(4096));
    return true;
  }();
  (void)raised;
  hipLaunchKernelGGL(k_rows<MODE>, dim3(a.rows, 1, bat_n()), dim3(kRowThreads), row_lds(a.N), st, a, bat_bs());
}

static void launch_transpose(hipStream_t st, int rows, int cols, const double* in, double* out) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32, bat_n());
  hipLaunchKernelGGL(k_transpose, grid, dim3(kBlock), 0, st, rows, cols, in, out, bat_bs());
}

static RowArgs rows_x(DctPlan* pl) {
  RowArgs a{};
  a.N = pl->nx;
  a.rows = pl->ny;
  a.tw = pl->tw_x;
  a.rot = pl->rot_x;
  a.cs = pl->cos_x;
  a.pow2 = pl->pow2_x;
  return a;
}
static RowArgs rows_y(DctPlan* pl) {
  RowArgs a{};
  a.N = pl->ny;
  a.rows = pl->nx;
  a.tw = pl->tw_y;
  a.rot = pl->rot_y;
  a.cs = pl->cos_y;
  a.pow2 = pl->pow2_y;
  return a;
}

void launch_dct2_2d(hipStream_t st, DctPlan* pl, double* a, double* work) {
  RowArgs rx = rows_x(pl);
  rx.in = a;
  rx.out = a;
  launch_rows<R_DCT2>(st, rx);
  launch_transpose(st, pl->ny, pl->nx, a, work);
  RowArgs ry = rows_y(pl);
  ry.in = work;
  ry.out = work;
  launch_rows<R_DCT2>(st, ry);
  launch_transpose(st, pl->nx, pl->ny, work, a);
}

void launch_dct3_2d(hipStream_t st, DctPlan* pl, double* a, double* work) {
  RowArgs rx = rows_x(pl);
  rx.in = a;
  rx.out = a;
  launch_rows<R_DCT3>(st, rx);
  launch_transpose(st, pl->ny, pl->nx, a, work);
  RowArgs ry = rows_y(pl);
  ry.in = work;
  ry.out = work;
  launch_rows<R_DCT3>(st, ry);
  launch_transpose(st, pl->nx, pl->ny, work, a);
}

// PreSolve_2 TP/PreSolve_2.f90:16-125 with dct_fftw TP/dct_fftw.f90:13-51
// and the caller's p = psi + p.  work holds 2*nx*ny doubles.
void launch_pressure(hipStream_t st, DctPlan* pl, const Grid& g, const Scal& s, const double* u,
                     const double* v, const double* pre_lcc, double coef, double* psi, double* p,
                     double* work, double* part, KState* ks) {
  double* w0 = work;
  double* w1 = work + (size_t)g.nx * g.ny;
  RowArgs a = rows_x(pl);
  a.g = g;
  a.u = u;
  a.v = v;
  a.dr = s.dr;
  a.dz = s.dz;
  a.out = w0;
  a.part = part;
  a.ks = ks;
  a.per = g.px;
  launch_rows<R_DIV_DCT2>(st, a);
  if (ks) launch_finalize(st, part, a.rows, 1, 1, ks->red, nullptr, PH_NONE, FIN_ALWAYS);
  launch_transpose(st, g.ny, g.nx, w0, w1);
  RowArgs b = rows_y(pl);
  b.g = g;
  b.in = w1;
  b.out = w1;
  b.pre = pre_lcc;
  launch_rows<R_SOLVE>(st, b);
  launch_transpose(st, g.nx, g.ny, w1, w0);
  RowArgs c = rows_x(pl);
  c.g = g;
  c.in = w0;
  c.per = g.px;
  c.c4 = g.px ? 2.0 * g.nx * g.ny : 4.0 * g.nx * g.ny;  // R2HC + HC2R scale by nx, DCT-II + III by 2 nx
  c.coef = coef;
  c.psi = psi;
  c.p = p;
  launch_rows<R_DCT3_FINISH>(st, c);
}

}  // namespace drd
