// kernels_prec.hip -- preconditioners of the fast path (no reference
// counterpart; the reference's own ones are in kernels_stencil.hip).
//
// 1. CH 2x2 block Jacobi: per cell, the exact 2x2 diagonal block of the
//    ph/mu operator of MATRIX_VECTOR_PH_MU_CG (TP/ph_mu_solve.f90:203-299),
//    inverted in closed form.  The reference's scalar Jacobi ignores the
//    phi-mu coupling, which makes its BiCGSTAB stall.
//
// 2. Velocity fast-Poisson preconditioner M^-1 = S^-1 A0^-1 S^-1, with A0 the
//    constant-coefficient operator a0 I - dt (2 dxx + dyy) on u and
//    a0 I - dt (dxx + 2 dyy) on v under the same ghost rules as the
//    reference's operator (u: mirror at i=0,nx -> DCT-I in x, odd at the
//    walls -> DST-II in y; v: even at the side walls -> DCT-II in x,
//    Dirichlet at k=0,ny -> DST-I in y), and S = sqrt(diag(A)/diag(A0)).
//    The transforms and the diagonal scalings run in T = float (FP32) or
//    double; input and output vectors are FP64.
//
// Lines are transformed in LDS by a complex radix-2 Stockham FFT whose length
// is a power of two (2nx, nx, ny, 2ny); other grids are rejected at set-up.
#include <cmath>
#include <vector>

#include "fft_regs.hpp"
#include "kernels.hpp"

namespace drd {

// ---- CH block Jacobi --------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_phmu_block(Grid g, Scal s, const double* __restrict__ bu,
                                                       const double* __restrict__ minner,
                                                       const double* __restrict__ m0,
                                                       const double* __restrict__ b, double* __restrict__ x,
                                                       const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (st && (st->done || st->early)) return;
  bu = bofs(bu, bs);
  minner = bofs(minner, bs);
  m0 = bofs(m0, bs);
  b = bofs(b, bs);
  x = bofs(x, bs);
  const int64_t C = g.ncell();
  const int64_t q = (int64_t)tile_id() * kBlock + threadIdx.x;  // XCD-contiguous rows
  if (q >= C) return;
  const int j = g.by_nx.div(q), i = (int)(q - (int64_t)j * g.nx);
  phmu_block_solve(g, s, bu, minner, m0, i, j, b[q], b[C + q], x[q], x[C + q]);
}

// BiCGSTAB vector updates with the block preconditioner applied on the fly
// (one thread per cell holds both the phi and the mu entry):
//   p = r (first iteration) or beta (p - omega v) + r;  phat = B^-1 p
__global__ __launch_bounds__(kBlock) void k_bicg_p_block(Grid g, Scal s, const double* __restrict__ bu,
                                                         const double* __restrict__ minner,
                                                         const double* __restrict__ m0,
                                                         const double* __restrict__ r,
                                                         const double* __restrict__ v, double* __restrict__ p,
                                                         double* __restrict__ ph, const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (st->done || st->early) return;
  bu = bofs(bu, bs);
  minner = bofs(minner, bs);
  m0 = bofs(m0, bs);
  r = bofs(r, bs);
  v = bofs(v, bs);
  p = bofs(p, bs);
  ph = bofs(ph, bs);
  const bool first = st->iter <= 1;
  const double omega = st->omega;
  const double beta = first ? 0.0 : (st->rho / st->rho1) * (st->alpha / st->omega);
  const int64_t C = g.ncell();
#pragma unroll
  for (int jj = 0; jj < kSlots; ++jj) {
    const int64_t q = tile_elem(C, 0, jj);
    if (q < 0) break;
    // p and v are read in the first iteration too (unused there), so all
    // loads of the slot issue together
    double pv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t e = q + h * C;
      const double re = r[e], pe = p[e], ve = v[e];
      double t = pe + (-omega) * ve;
      t = beta * t;
      pv[h] = first ? re : t + re;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) p[q + h * C] = pv[h];
    const int j = g.by_nx.div(q), i = (int)(q - (int64_t)j * g.nx);
    phmu_block_solve(g, s, bu, minner, m0, i, j, pv[0], pv[1], ph[q], ph[C + q]);
  }
}

//   s = r - alpha v (S aliases R), ||s||^2 (canonical partials), shat = B^-1 s
__global__ __launch_bounds__(kBlock) void k_bicg_s_block(Grid g, Scal s, const double* __restrict__ bu,
                                                         const double* __restrict__ minner,
                                                         const double* __restrict__ m0, double* __restrict__ r,
                                                         const double* __restrict__ v, double* __restrict__ sh,
                                                         double* part, const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (st->done || st->early) return;
  bu = bofs(bu, bs);
  minner = bofs(minner, bs);
  m0 = bofs(m0, bs);
  r = bofs(r, bs);
  v = bofs(v, bs);
  sh = bofs(sh, bs);
  part = bofs(part, bs);
  const double alpha = st->alpha;
  const int64_t C = g.ncell();
  double t[2][kSlots];
#pragma unroll
  for (int jj = 0; jj < kSlots; ++jj) {
    const int64_t q = tile_elem(C, 0, jj);
    double sv[2] = {0.0, 0.0};
    if (q >= 0) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t e = q + h * C;
        sv[h] = r[e] + (-alpha) * v[e];
        r[e] = sv[h];
      }
      const int j = g.by_nx.div(q), i = (int)(q - (int64_t)j * g.nx);
      phmu_block_solve(g, s, bu, minner, m0, i, j, sv[0], sv[1], sh[q], sh[C + q]);
    }
    t[0][jj] = sv[0] * sv[0];
    t[1][jj] = sv[1] * sv[1];
  }
  tile_publish<2, 1>(t, part);
}

void launch_bicg_p_block(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const double* r,
                         const double* v, double* p, double* ph, KState* ks) {
  hipLaunchKernelGGL(k_bicg_p_block, dim3(tiles_of(g.ncell()), 1, bat_n()), dim3(kBlock), 0, st, g, s, c.bu,
                     c.minner, c.m0, r, v, p, ph, ks, bat_bs());
}

void launch_bicg_s_block(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, double* r, const double* v,
                         double* sh, double* part, KState* ks) {
  hipLaunchKernelGGL(k_bicg_s_block, dim3(tiles_of(g.ncell()), 1, bat_n()), dim3(kBlock), 0, st, g, s, c.bu,
                     c.minner, c.m0, r, v, sh, part, ks, bat_bs());
  launch_finalize(st, part, tiles_of(g.ncell()), 2, 1, nullptr, ks, PH_S, FIN_ACTIVE);
}

void launch_phmu_block(hipStream_t st, const Grid& g, const Scal& s, const Coef& c, const double* b, double* x,
                       const KState* ks) {
  hipLaunchKernelGGL(k_phmu_block, dim3((unsigned)((g.ncell() + kBlock - 1) / kBlock), 1, bat_n()), dim3(kBlock),
                     0, st, g, s, c.bu, c.minner, c.m0, b, x, ks, bat_bs());
}

// ---- exact diagonal of the BDF2 velocity operator and the scaling S ---------
// (MatrixVector_uv_2_CG TP/VeloSolve_2_combined.f90:607-670 restricted to the
// coefficient of the row's own unknown)
__global__ __launch_bounds__(kBlock) void k_velo_scale(Grid g, Scal s, const double* __restrict__ rho,
                                                       const double* __restrict__ eta,
                                                       const double* __restrict__ chi, double a0,
                                                       double* __restrict__ inv_s64, float* __restrict__ inv_s32,
                                                       int64_t bs) {
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  chi = bofs(chi, bs);
  inv_s64 = bofs(inv_s64, bs);
  inv_s32 = bofs(inv_s32, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= g.nuv()) return;
  const double d = velo_diag2(g, s, rho, eta, chi, q);
  const double d0 = q < g.mn() ? a0 + s.dt * (4.0 / s.dr2 + 2.0 / s.dz2) : a0 + s.dt * (2.0 / s.dr2 + 4.0 / s.dz2);
  const double is = sqrt(d0 / d);
  inv_s64[q] = is;
  inv_s32[q] = (float)is;
}

template <class T>
struct PrecTables {
  const Cx<T>* tw2nx;  // L = 2nx (u rows, DCT-I)
  const Cx<T>* twnx;   // L = nx  (v rows, DCT-II/III)
  const Cx<T>* rotnx;
  const Cx<T>* twny;   // L = ny  (u columns, DST-II/III)
  const Cx<T>* rotny;
  const Cx<T>* tw2ny;  // L = 2ny (v columns, DST-I)
  // eigenvalues of A0 times the transform scale (2nx)(2ny):
  //   lambda = sc (a0 + ex[kx] + ey[ky]),
  // u: ex_u[kx] = 8 dt sin^2(pi kx / 2nx) / dr2, ey_u[k] = 4 dt sin^2(pi (k+1) / 2ny) / dz2
  // v: ex_v[kx] = 4 dt sin^2(pi kx / 2nx) / dr2, ey_v[k] = 8 dt sin^2(pi (k+1) / 2ny) / dz2
  const T* ex_u;  // nx+1
  const T* ey_u;  // ny
  const T* ex_v;  // nx
  const T* ey_v;  // ny-1
  T sc, a0;
};

// ---- paired real line transforms -------------------------------------------
// Two real lines share one complex FFT (line 0 in the real part, line 1 in
// the imaginary part), FFTW conventions, unnormalised.  load(i) returns the
// pair's i-th inputs; store(i, y0, y1) takes the outputs.  A is the LDS
// line (FFT length complex values, element i at A[lp(i)]); every helper
// ends with a barrier.
__device__ __forceinline__ int makhoul_ix(int j, int n) { return (j & 1) ? (n - 1 - (j >> 1)) : (j >> 1); }

// REDFT00 of n = N+1 points: FFT of the even extension (length 2N); both
// spectra are real, so line 0 is Re V and line 1 is Im V.
template <class T, int LMAX, int NT = kBlock, class Load, class Store>
__device__ void dct1_pair(Cx<T>* A, int N, const Cx<T>* __restrict__ tw, Load load, Store store) {
  const int L = 2 * N;
  for (int m = threadIdx.x; m < L; m += blockDim.x) A[lp(m)] = load(m <= N ? m : L - m);
  __syncthreads();
  fft_lds<T, LMAX, NT>(A, L, tw, T(-1));
  for (int k = threadIdx.x; k <= N; k += blockDim.x) store(k, A[lp(k)].x, A[lp(k)].y);
  __syncthreads();
}

// REDFT10 (dst = false) or RODFT10 (dst = true: inputs (-1)^j x_j, outputs
// reversed) by Makhoul's reordering; the two spectra are separated by
// V0_k = (V_k + conj V_{-k}) / 2, V1_k = (V_k - conj V_{-k}) / 2i, and
// y_k = 2 Re(exp(-i pi k / 2n) V_k).
template <class T, int LMAX, int NT = kBlock, class Load, class Store>
__device__ void dct2_pair(Cx<T>* A, int n, const Cx<T>* __restrict__ tw, const Cx<T>* __restrict__ rot, bool dst,
                          Load load, Store store) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    Cx<T> v = load(j);
    if (dst && (j & 1)) v = Cx<T>{-v.x, -v.y};
    A[lp(makhoul_ix(j, n))] = v;
  }
  __syncthreads();
  fft_lds<T, LMAX, NT>(A, n, tw, T(-1));
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const Cx<T> V = A[lp(k)], W = A[lp((n - k) & (n - 1))];
    const Cx<T> sp{V.x + W.x, V.y - W.y}, sm{V.x - W.x, V.y + W.y};  // V +- conj W
    const Cx<T> r = rot[k];
    store(dst ? n - 1 - k : k, r.x * sp.x - r.y * sp.y, r.x * sm.y + r.y * sm.x);
  }
  __syncthreads();
}

// REDFT01 (dst = false) or RODFT01 (dst = true) of the pair held where
// load(i) reads it (random access): inverse Makhoul; the two inverse FFT
// outputs are real, so they separate as Re and Im.
template <class T, int LMAX, int NT = kBlock, class Load, class Store>
__device__ void dct3_pair(Cx<T>* A, int n, const Cx<T>* __restrict__ tw, const Cx<T>* __restrict__ rot, bool dst,
                          Load load, Store store) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const Cx<T> a = load(dst ? n - 1 - k : k);
    Cx<T> b{T(0), T(0)};
    if (k > 0) {
      const Cx<T> t = load(dst ? k - 1 : n - k);
      b = Cx<T>{-t.x, -t.y};
    }
    const Cx<T> r = rot[k];
    const Cx<T> a0{r.x * a.x + r.y * b.x, -r.y * a.x + r.x * b.x};  // line 0
    const Cx<T> a1{r.x * a.y + r.y * b.y, -r.y * a.y + r.x * b.y};  // line 1
    A[lp(k)] = Cx<T>{a0.x - a1.y, a0.y + a1.x};
  }
  __syncthreads();
  fft_lds<T, LMAX, NT>(A, n, tw, T(1));
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    Cx<T> v = A[lp(makhoul_ix(j, n))];
    if (dst && (j & 1)) v = Cx<T>{-v.x, -v.y};
    store(j, v.x, v.y);
  }
  __syncthreads();
}

// RODFT00 of n points: FFT of the odd extension (length 2(n+1)); both
// spectra are imaginary: y0_k = -Im V_{k+1}, y1_k = Re V_{k+1}.
template <class T, int LMAX, int NT = kBlock, class Load, class Store>
__device__ void dst1_pair(Cx<T>* A, int n, const Cx<T>* __restrict__ tw, Load load, Store store) {
  const int L = 2 * (n + 1);
  for (int m = threadIdx.x; m < L; m += blockDim.x) {
    Cx<T> v{T(0), T(0)};
    if (m >= 1 && m <= n) v = load(m - 1);
    if (m >= n + 2) {
      const Cx<T> t = load(L - m - 1);
      v = Cx<T>{-t.x, -t.y};
    }
    A[lp(m)] = v;
  }
  __syncthreads();
  fft_lds<T, LMAX, NT>(A, L, tw, T(-1));
  for (int k = threadIdx.x; k < n; k += blockDim.x) store(k, -A[lp(k + 1)].y, A[lp(k + 1)].x);
  __syncthreads();
}

// pass 1 (fwd) / 3 (inv) over rows, two rows per workgroup: u rows (pairs
// [0, nu)), v rows (pairs [nu, nu + nv)).
// fwd: w = DCT_x(inv_s * r) in T;  inv: z = inv_s * DCT_x^-1(w) in FP64
template <class T, bool FWD, int LMAX>
__global__ __launch_bounds__(kBlock) void k_prec_rows(Grid g, PrecTables<T> tb, const double* __restrict__ r,
                                                      const T* __restrict__ inv_s, T* __restrict__ wu,
                                                      T* __restrict__ wv, double* __restrict__ z,
                                                      const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (st && (st->done || st->early)) return;
  r = bofs(r, bs);  // the tables tb are shared by the group
  inv_s = bofs(inv_s, bs);
  wu = bofs(wu, bs);
  wv = bofs(wv, bs);
  z = bofs(z, bs);
  extern __shared__ unsigned char smem_raw[];
  Cx<T>* A = reinterpret_cast<Cx<T>*>(smem_raw);
  const int nx = g.nx, ny = g.ny;
  const int nu = (ny + 1) / 2;
  if ((int)blockIdx.x < nu) {  // u rows 2b, 2b+1: n = nx+1 points, DCT-I both ways
    const int r0 = 2 * blockIdx.x, n = nx + 1;
    const bool has1 = r0 + 1 < ny;
    const int64_t b0 = (int64_t)r0 * n, b1 = b0 + n;
    auto load = [&](int i) -> Cx<T> {
      if (FWD) return Cx<T>{T(r[b0 + i]) * inv_s[b0 + i], has1 ? T(r[b1 + i]) * inv_s[b1 + i] : T(0)};
      return Cx<T>{wu[b0 + i], has1 ? wu[b1 + i] : T(0)};
    };
    auto store = [&](int k, T y0, T y1) {
      if (FWD) {
        wu[b0 + k] = y0;
        if (has1) wu[b1 + k] = y1;
      } else {
        z[b0 + k] = (double)(y0 * inv_s[b0 + k]);
        if (has1) z[b1 + k] = (double)(y1 * inv_s[b1 + k]);
      }
    };
    dct1_pair<T, LMAX>(A, nx, tb.tw2nx, load, store);
  } else {  // v rows 2b, 2b+1 of the ny-1: n = nx points, DCT-II forward, DCT-III back
    const int r0 = 2 * (blockIdx.x - nu), n = nx;
    const bool has1 = r0 + 1 < ny - 1;
    const int64_t w0 = (int64_t)r0 * n, w1 = w0 + n;           // in wv
    const int64_t b0 = g.mn() + w0, b1 = g.mn() + w1;           // in r, z, inv_s
    if (FWD) {
      auto load = [&](int i) -> Cx<T> {
        return Cx<T>{T(r[b0 + i]) * inv_s[b0 + i], has1 ? T(r[b1 + i]) * inv_s[b1 + i] : T(0)};
      };
      auto store = [&](int k, T y0, T y1) {
        wv[w0 + k] = y0;
        if (has1) wv[w1 + k] = y1;
      };
      dct2_pair<T, LMAX>(A, n, tb.twnx, tb.rotnx, false, load, store);
    } else {
      auto load = [&](int i) -> Cx<T> { return Cx<T>{wv[w0 + i], has1 ? wv[w1 + i] : T(0)}; };
      auto store = [&](int j, T x0, T x1) {
        z[b0 + j] = (double)(x0 * inv_s[b0 + j]);
        if (has1) z[b1 + j] = (double)(x1 * inv_s[b1 + j]);
      };
      dct3_pair<T, LMAX>(A, n, tb.twnx, tb.rotnx, false, load, store);
    }
  }
}

// pass 2 over columns of the transposed blocks, two columns per workgroup:
// u columns (pairs [0, nu)), v columns ([nu, nu + nv)); forward transform,
// divide by the eigenvalue, inverse.  The divided spectra wait in LDS
// behind the FFT line.  A 256-thread workgroup per pair leaves threads idle
// in the radix-16 passes of a 1024-point line, but one wave per pair was
// slower still (measured: 128 against 106 us per application at 2048 x
// 1024): the LDS a pair needs then allows only six waves per CU.
template <class T, int LMAX>
constexpr int cols_threads() {
  return kBlock;
}
// STRIDED: the columns are read and written in place in the row-major
// blocks of pass 1 (element i of u column c at tu[i (nx + 1) + c], of v
// column c at tv[i nx + c]) instead of in transposed blocks: for small grids,
// whose blocks stay in cache, this saves the two transpose launches.  The
// same transforms in the same order either way.
template <class T, int LMAX, bool STRIDED = false>
__global__ __launch_bounds__((cols_threads<T, LMAX>())) void k_prec_cols(Grid g, PrecTables<T> tb, T* __restrict__ tu,
                                                                       T* __restrict__ tv, const KState* st,
                                                                       int64_t bs) {
  constexpr int NT = cols_threads<T, LMAX>();
  st = bofs(st, bs);
  if (st && (st->done || st->early)) return;
  tu = bofs(tu, bs);
  tv = bofs(tv, bs);
  extern __shared__ unsigned char smem_raw[];
  const int nx = g.nx, ny = g.ny;
  Cx<T>* A = reinterpret_cast<Cx<T>*>(smem_raw);
  T* l0 = reinterpret_cast<T*>(A + lds_len(2 * ny));
  T* l1 = l0 + ny;
  const int nu = (nx + 2) / 2;
  if ((int)blockIdx.x < nu) {  // u columns: DST-II, / lambda, DST-III (n = ny)
    const int c0 = 2 * blockIdx.x, n = ny;
    const bool has1 = c0 + 1 < nx + 1;
    // element i of the pair's columns: col0[i * cs], col1[i * cs]
    const int64_t cs = STRIDED ? nx + 1 : 1;
    T* col0 = STRIDED ? tu + c0 : tu + (int64_t)c0 * n;
    T* col1 = STRIDED ? col0 + 1 : col0 + n;
    const T e0 = tb.a0 + tb.ex_u[c0], e1 = has1 ? tb.a0 + tb.ex_u[c0 + 1] : T(1);
    dct2_pair<T, LMAX, NT>(A, n, tb.twny, tb.rotny, true,
                 [&](int i) { return Cx<T>{col0[i * cs], has1 ? col1[i * cs] : T(0)}; },
                 [&](int k, T y0, T y1) {
                   const T ey = tb.ey_u[k];
                   l0[k] = y0 / (tb.sc * (e0 + ey));
                   l1[k] = has1 ? y1 / (tb.sc * (e1 + ey)) : T(0);
                 });
    dct3_pair<T, LMAX, NT>(A, n, tb.twny, tb.rotny, true, [&](int i) { return Cx<T>{l0[i], l1[i]}; },
                 [&](int j, T x0, T x1) {
                   col0[j * cs] = x0;
                   if (has1) col1[j * cs] = x1;
                 });
  } else {  // v columns: DST-I, / lambda, DST-I (n = ny-1)
    const int c0 = 2 * (blockIdx.x - nu), n = ny - 1;
    const bool has1 = c0 + 1 < nx;
    const int64_t cs = STRIDED ? nx : 1;
    T* col0 = STRIDED ? tv + c0 : tv + (int64_t)c0 * n;
    T* col1 = STRIDED ? col0 + 1 : col0 + n;
    const T e0 = tb.a0 + tb.ex_v[c0], e1 = has1 ? tb.a0 + tb.ex_v[c0 + 1] : T(1);
    dst1_pair<T, LMAX, NT>(A, n, tb.tw2ny, [&](int i) { return Cx<T>{col0[i * cs], has1 ? col1[i * cs] : T(0)}; },
                 [&](int k, T y0, T y1) {
                   const T ey = tb.ey_v[k];
                   l0[k] = y0 / (tb.sc * (e0 + ey));
                   l1[k] = has1 ? y1 / (tb.sc * (e1 + ey)) : T(0);
                 });
    dst1_pair<T, LMAX, NT>(A, n, tb.tw2ny, [&](int i) { return Cx<T>{l0[i], l1[i]}; },
                 [&](int k, T y0, T y1) {
                   col0[k * cs] = y0;
                   if (has1) col1[k * cs] = y1;
                 });
  }
}

// One block of a transpose: out (cols x rows) = in^T, in row-major rows x cols.
template <class T>
struct TBlock {
  int rows, cols;
  const T* in;
  T* out;
};

// The u and the v block of a pass transposed by one launch: grid rows
// [0, ya) tile block a, the others block b, and the grid is as wide as the
// wider block.  Each tile goes through a TR x TC LDS tile (TR rows of in,
// TC columns); the tile shape is the autotuner's choice.
template <class T, int TR, int TC>
__global__ __launch_bounds__(kBlock) void k_transpose2T(TBlock<T> a, TBlock<T> b, int ya, const KState* st,
                                                        int64_t bs) {
  st = bofs(st, bs);
  if (st && (st->done || st->early)) return;
  const bool in_a = (int)blockIdx.y < ya;
  const int rows = in_a ? a.rows : b.rows, cols = in_a ? a.cols : b.cols;
  const T* __restrict__ in = bofs(in_a ? a.in : b.in, bs);
  T* __restrict__ out = bofs(in_a ? a.out : b.out, bs);
  const int bx = blockIdx.x * TC, by = (in_a ? (int)blockIdx.y : (int)blockIdx.y - ya) * TR;
  if (bx >= cols) return;  // the narrower block's surplus tiles (whole workgroup)
  __shared__ T tile[TR][TC + 1];
  {
    const int tx = threadIdx.x % TC;
    for (int j = threadIdx.x / TC; j < TR; j += kBlock / TC) {
      const int rr = by + j, c = bx + tx;
      if (rr < rows && c < cols) tile[j][tx] = in[(size_t)rr * cols + c];
    }
  }
  __syncthreads();
  {
    const int tx = threadIdx.x % TR;
    for (int j = threadIdx.x / TR; j < TC; j += kBlock / TR) {
      const int rr = bx + j, c = by + tx;  // out row = in column
      if (rr < cols && c < rows) out[(size_t)rr * rows + c] = tile[tx][j];
    }
  }
}

// tile shapes the autotuner chooses from: TR x TC
constexpr int kTransposeShapes = 3;
constexpr int kShapeTR[kTransposeShapes] = {32, 16, 64};
constexpr int kShapeTC[kTransposeShapes] = {32, 64, 16};

template <class T>
static void launch_transpose2T(hipStream_t st, int shape, TBlock<T> a, TBlock<T> b, const KState* ks, unsigned B,
                               int64_t bs) {
  const int TR = kShapeTR[shape], TC = kShapeTC[shape];
  const int ya = (a.rows + TR - 1) / TR, yb = (b.rows + TR - 1) / TR;
  const dim3 grid((unsigned)((std::max(a.cols, b.cols) + TC - 1) / TC), (unsigned)(ya + yb), B);
#define DRD_T2(R, C) hipLaunchKernelGGL((k_transpose2T<T, R, C>), grid, dim3(kBlock), 0, st, a, b, ya, ks, bs)
  switch (shape) {
    case 1: DRD_T2(16, 64); break;
    case 2: DRD_T2(64, 16); break;
    default: DRD_T2(32, 32); break;
  }
#undef DRD_T2
}

// ---- host side -------------------------------------------------------------------
struct VeloPrec {
  int nx, ny;
  bool fp32;
  void* tabs_mem = nullptr;  // twiddles and rotations (depend on nx, ny only)
  void* eig_mem = nullptr;   // ex_u, ey_u, ex_v, ey_v (depend on dt, dx, dz)
  PrecTables<float> tf{};
  PrecTables<double> td{};
  double* inv_s64 = nullptr;
  float* inv_s32 = nullptr;
  void* work = nullptr;  // wu, wv, tu, tv in T
  bool own_bufs = true;  // false: inv_s64, inv_s32, work belong to the caller
  int tshape = 0;        // transpose tile shape (velo_prec_tune)
  double a0 = 0;
};

template <class T>
static void host_tw(std::vector<Cx<T>>& v, int L) {  // exp(-2 pi i t / L), t < L
  for (int t = 0; t < L; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)L;
    v.push_back(Cx<T>{(T)std::cos(a), (T)std::sin(a)});
  }
}
template <class T>
static void host_rot(std::vector<Cx<T>>& v, int L) {
  for (int k = 0; k < L; ++k) {
    const double a = -M_PI * (double)k / (2.0 * L);
    v.push_back(Cx<T>{(T)std::cos(a), (T)std::sin(a)});
  }
}

template <class T>
static int build_twiddles(VeloPrec* pc, PrecTables<T>& tb) {
  const int nx = pc->nx, ny = pc->ny;
  std::vector<Cx<T>> cx;
  std::vector<size_t> off;
  off.push_back(cx.size());
  host_tw<T>(cx, 2 * nx);
  off.push_back(cx.size());
  host_tw<T>(cx, nx);
  off.push_back(cx.size());
  host_rot<T>(cx, nx);
  off.push_back(cx.size());
  host_tw<T>(cx, ny);
  off.push_back(cx.size());
  host_rot<T>(cx, ny);
  off.push_back(cx.size());
  host_tw<T>(cx, 2 * ny);
  const size_t bytes = cx.size() * sizeof(Cx<T>);
  if (hipMalloc(&pc->tabs_mem, bytes) != hipSuccess) return DRD_E_HIP;
  if (hipMemcpy(pc->tabs_mem, cx.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return DRD_E_HIP;
  const Cx<T>* base = reinterpret_cast<const Cx<T>*>(pc->tabs_mem);
  tb.tw2nx = base + off[0];
  tb.twnx = base + off[1];
  tb.rotnx = base + off[2];
  tb.twny = base + off[3];
  tb.rotny = base + off[4];
  tb.tw2ny = base + off[5];
  return 0;
}

// 1-D parts of the eigenvalues of A0 = a0 I - dt (2 dxx + dyy) (u) and
// a0 I - dt (dxx + 2 dyy) (v) under the reference's ghost rules, and the
// scale (2nx)(2ny) of the unnormalised transform pair.
template <class T>
static int build_eigen(VeloPrec* pc, const Scal& s, PrecTables<T>& tb) {
  const int nx = pc->nx, ny = pc->ny;
  std::vector<T> e;
  for (int kx = 0; kx <= nx; ++kx) {
    const double sx = std::sin(M_PI * kx / (2.0 * nx));
    e.push_back((T)(s.dt * 2.0 * 4.0 * sx * sx / s.dr2));
  }
  for (int ky = 0; ky < ny; ++ky) {
    const double sy = std::sin(M_PI * (ky + 1) / (2.0 * ny));
    e.push_back((T)(s.dt * 4.0 * sy * sy / s.dz2));
  }
  for (int kx = 0; kx < nx; ++kx) {
    const double sx = std::sin(M_PI * kx / (2.0 * nx));
    e.push_back((T)(s.dt * 4.0 * sx * sx / s.dr2));
  }
  for (int ky = 1; ky < ny; ++ky) {
    const double sy = std::sin(M_PI * ky / (2.0 * ny));
    e.push_back((T)(s.dt * 2.0 * 4.0 * sy * sy / s.dz2));
  }
  const size_t bytes = e.size() * sizeof(T);
  if (!pc->eig_mem && hipMalloc(&pc->eig_mem, bytes) != hipSuccess) return DRD_E_HIP;
  if (hipMemcpy(pc->eig_mem, e.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return DRD_E_HIP;
  const T* b = reinterpret_cast<const T*>(pc->eig_mem);
  tb.ex_u = b;
  tb.ey_u = b + (nx + 1);
  tb.ex_v = b + (nx + 1) + ny;
  tb.ey_v = b + (nx + 1) + ny + nx;
  tb.sc = (T)((2.0 * nx) * (2.0 * ny));
  tb.a0 = (T)pc->a0;
  return 0;
}

int velo_prec_update(VeloPrec* pc, const Scal& s) {
  pc->a0 = 1.5 * s.re;  // unit density, unit viscosity, no penalty
  return pc->fp32 ? build_eigen<float>(pc, s, pc->tf) : build_eigen<double>(pc, s, pc->td);
}

static bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

VeloPrec* velo_prec_create(const Grid& g, const Scal& s, bool fp32, int* err, const VeloPrecBufs* ext) {
  *err = 0;
  if (!pow2(g.nx) || !pow2(g.ny) || g.nx > 4096 || g.ny > 4096) {
    *err = DRD_E_UNSUPPORTED;
    return nullptr;
  }
  {  // the row and column passes hold a padded FFT line (and the column spectra) in LDS
    const size_t cx = fp32 ? 2 * sizeof(float) : 2 * sizeof(double), re = cx / 2;
    const size_t rows = (size_t)lds_len(2 * g.nx) * cx, cols = (size_t)lds_len(2 * g.ny) * cx + 2 * (size_t)g.ny * re;
    if (std::max(rows, cols) > 160 * 1024) {
      *err = DRD_E_UNSUPPORTED;
      return nullptr;
    }
  }
  VeloPrec* pc = new VeloPrec();
  pc->nx = g.nx;
  pc->ny = g.ny;
  pc->fp32 = fp32;
  int rc = fp32 ? build_twiddles<float>(pc, pc->tf) : build_twiddles<double>(pc, pc->td);
  if (!rc) rc = velo_prec_update(pc, s);
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  if (ext) {
    pc->inv_s64 = ext->inv_s64;
    pc->inv_s32 = ext->inv_s32;
    pc->work = ext->work;
    pc->own_bufs = false;
  } else {
    if (!rc && hipMalloc(&pc->inv_s64, g.nuv() * sizeof(double)) != hipSuccess) rc = DRD_E_HIP;
    if (!rc && hipMalloc(&pc->inv_s32, g.nuv() * sizeof(float)) != hipSuccess) rc = DRD_E_HIP;
    if (!rc && hipMalloc(&pc->work, 2 * g.nuv() * esz) != hipSuccess) rc = DRD_E_HIP;
  }
  if (rc) {
    velo_prec_destroy(pc);
    *err = rc;
    return nullptr;
  }
  return pc;
}

void velo_prec_destroy(VeloPrec* pc) {
  if (!pc) return;
  (void)hipFree(pc->tabs_mem);
  (void)hipFree(pc->eig_mem);
  if (pc->own_bufs) {
    (void)hipFree(pc->inv_s64);
    (void)hipFree(pc->inv_s32);
    (void)hipFree(pc->work);
  }
  delete pc;
}

void velo_prec_setup(hipStream_t st, VeloPrec* pc, const Grid& g, const Scal& s, const Coef& c) {
  hipLaunchKernelGGL(k_velo_scale, dim3((unsigned)((g.nuv() + kBlock - 1) / kBlock), 1, bat_n()), dim3(kBlock), 0,
                     st, g, s, c.rho, c.eta, c.chi, pc->a0, pc->inv_s64, pc->inv_s32, bat_bs());
}

// Grids up to this nx take the column pass in place (STRIDED); the coarsest
// multigrid level, 64 x 32 at c1, is one of them.
constexpr int kPrecStridedMaxNx = 256;
// LMAX: longest row line (2nx); LMAXC: longest column line (2ny)
template <class T, int LMAX, int LMAXC>
static void launch_prec_passes(hipStream_t st, const Grid& g, const PrecTables<T>& tb, const T* inv_s,
                               const double* r, double* z, T* wu, T* wv, T* tu, T* tv, unsigned rows_blocks,
                               unsigned cols_blocks, size_t lds_rows, size_t lds_cols, const KState* ks, int shape) {
  const int nx = g.nx, ny = g.ny;
  // lines of 2nx or 2ny complex may exceed 64 KiB; raised once (a function-
  // local static: concurrent first callers wait for it)
  static const bool raised = [] {
    const int lim = 160 * 1024;
    (void)hipFuncSetAttribute((const void*)k_prec_rows<T, true, LMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    (void)hipFuncSetAttribute((const void*)k_prec_rows<T, false, LMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    (void)hipFuncSetAttribute((const void*)k_prec_cols<T, LMAXC>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void*)k_prec_cols<T, LMAXC, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    return true;
  }();
  (void)raised;
  const unsigned B = bat_n();
  const int64_t bs = bat_bs();
  hipLaunchKernelGGL((k_prec_rows<T, true, LMAX>), dim3(rows_blocks, 1, B), dim3(kBlock), lds_rows, st, g, tb, r,
                     inv_s, wu, wv, (double*)nullptr, ks, bs);
  if (nx <= kPrecStridedMaxNx) {  // small grids: columns in place, no transposes
    hipLaunchKernelGGL((k_prec_cols<T, LMAXC, true>), dim3(cols_blocks, 1, B), dim3(cols_threads<T, LMAXC>()),
                       lds_cols, st, g, tb, wu, wv, ks, bs);
    hipLaunchKernelGGL((k_prec_rows<T, false, LMAX>), dim3(rows_blocks, 1, B), dim3(kBlock), lds_rows, st, g, tb,
                       (const double*)nullptr, inv_s, wu, wv, z, ks, bs);
    return;
  }
  launch_transpose2T<T>(st, shape, TBlock<T>{ny, nx + 1, wu, tu}, TBlock<T>{ny - 1, nx, wv, tv}, ks, B, bs);
  hipLaunchKernelGGL((k_prec_cols<T, LMAXC>), dim3(cols_blocks, 1, B), dim3(cols_threads<T, LMAXC>()), lds_cols, st,
                     g, tb, tu, tv, ks, bs);
  launch_transpose2T<T>(st, shape, TBlock<T>{nx + 1, ny, tu, wu}, TBlock<T>{nx, ny - 1, tv, wv}, ks, B, bs);
  hipLaunchKernelGGL((k_prec_rows<T, false, LMAX>), dim3(rows_blocks, 1, B), dim3(kBlock), lds_rows, st, g, tb,
                     (const double*)nullptr, inv_s, wu, wv, z, ks, bs);
}

template <class T>
static void apply_T(hipStream_t st, VeloPrec* pc, const Grid& g, const PrecTables<T>& tb, const T* inv_s,
                    const double* r, double* z, const KState* ks) {
  const int nx = g.nx, ny = g.ny;
  T* wu = reinterpret_cast<T*>(pc->work);
  T* wv = wu + (size_t)(nx + 1) * ny;
  T* tu = wv + (size_t)nx * (ny - 1);
  T* tv = tu + (size_t)(nx + 1) * ny;
  // two lines per workgroup; LDS holds the longest FFT line (2nx for the
  // u rows' DCT-I, 2ny for the v columns' DST-I) and, for columns, the
  // divided spectra of the pair
  const unsigned rows_blocks = (unsigned)((ny + 1) / 2 + ny / 2);
  const unsigned cols_blocks = (unsigned)((nx + 2) / 2 + (nx + 1) / 2);
  const size_t lds_rows = (size_t)lds_len(2 * nx) * sizeof(Cx<T>);
  const size_t lds_cols = (size_t)lds_len(2 * ny) * sizeof(Cx<T>) + 2 * (size_t)ny * sizeof(T);
#define DRD_PREC_PASSES(LR, LC)                                                                          \
  launch_prec_passes<T, LR, LC>(st, g, tb, inv_s, r, z, wu, wv, tu, tv, rows_blocks, cols_blocks, lds_rows, \
                                lds_cols, ks, pc->tshape)
  if (nx <= 2048) {
    if (ny <= 1024)
      DRD_PREC_PASSES(4096, 2048);
    else
      DRD_PREC_PASSES(4096, 8192);
  } else {
    if (ny <= 1024)
      DRD_PREC_PASSES(8192, 2048);
    else
      DRD_PREC_PASSES(8192, 8192);
  }
#undef DRD_PREC_PASSES
}

void velo_prec_apply(hipStream_t st, VeloPrec* pc, const Grid& g, const double* r, double* z, const KState* ks) {
  if (pc->fp32)
    apply_T<float>(st, pc, g, pc->tf, pc->inv_s32, r, z, ks);
  else
    apply_T<double>(st, pc, g, pc->td, pc->inv_s64, r, z, ks);
}

// Autotuner: times `reps` applications with each transpose tile shape (after
// one warm-up each) and keeps the fastest.  ms[kTransposeShapes] gets the
// mean times; returns the chosen shape (negative on a HIP error).
int velo_prec_tune(hipStream_t st, VeloPrec* pc, const Grid& g, const double* r, double* z, int reps, double* ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
  int best = 0;
  for (int k = 0; k < kTransposeShapes; ++k) {
    pc->tshape = k;
    velo_prec_apply(st, pc, g, r, z, nullptr);
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < reps; ++i) velo_prec_apply(st, pc, g, r, z, nullptr);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    ms[k] = (double)t / reps;
    if (ms[k] < ms[best]) best = k;
  }
  pc->tshape = best;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return hipGetLastError() == hipSuccess ? best : -1;
}

}  // namespace drd
