// kernels_krylov.hip -- fused BiCGSTAB vector kernels (FP64).
//
// Each iteration of LIB/BiCGSTABREVCOM.f:257-388 runs as
//   p-update | M^-1 | A + (rtld,v) | s-update + ||s|| | [early x] | M^-1 |
//   A + (t,s),(t,t) | x,r-update + ||r||,(rtld,r)
// where each reducing kernel publishes per-tile partial sums and a
// one-workgroup k_finalize adds them in the canonical order and takes the
// branch decisions (kfinalize).  The kernels test KState::done/early first,
// so the host may queue whole iterations ahead and only polls `done` now
// and then.
//
// The vector updates keep DAXPY's y + a*x rounding (no contraction), so
// they reproduce the CPU oracle bit for bit; only the reduction order differs.
#include "kernels.hpp"

namespace drd {

__device__ __forceinline__ bool kact(const KState* st) { return !st->done && !st->early; }

// ||x||^2, ||b||^2                                 REVCOM:229-253
template <int NS>
__global__ __launch_bounds__(kBlock) void k_init_norms(int64_t seg, const double* __restrict__ x,
                                                       const double* __restrict__ b, double* part,
                                                       KState* st, int64_t bs) {
  x = bofs(x, bs);
  b = bofs(b, bs);
  part = bofs(part, bs);
  double t[NS * 2][kSlots];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t q = tile_elem(seg, s, j);
      const double a = q >= 0 ? x[q] : 0.0, c = q >= 0 ? b[q] : 0.0;
      t[s * 2][j] = a * a;
      t[s * 2 + 1][j] = c * c;
    }
  tile_publish<NS, 2>(t, part);
}

// p = r (first iteration) or p = beta*(p - omega*v) + r   REVCOM:266-275
__global__ __launch_bounds__(kBlock) void k_bicg_p(int64_t n, const double* __restrict__ r,
                                                   const double* __restrict__ v, double* __restrict__ p,
                                                   const KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (!kact(st)) return;
  r = bofs(r, bs);
  v = bofs(v, bs);
  p = bofs(p, bs);
  const bool first = st->iter <= 1;
  const double omega = st->omega;
  const double beta = first ? 0.0 : (st->rho / st->rho1) * (st->alpha / st->omega);
  const int64_t base = (int64_t)blockIdx.x * kTile;
  // every slot's loads first (indices past n clamped to n-1, their values
  // unused), so they are all in flight together; then the updates
  double rr[kSlots], pp[kSlots], vv[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t q = base + j * kBlock + threadIdx.x, qc = q < n ? q : n - 1;
    rr[j] = r[qc];
    pp[j] = p[qc];
    vv[j] = v[qc];
  }
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t q = base + j * kBlock + threadIdx.x;
    if (q < n) {
      if (first) {
        p[q] = rr[j];
      } else {
        double t = pp[j] + (-omega) * vv[j];
        t = beta * t;
        p[q] = t + rr[j];
      }
    }
  }
}

// s = r - alpha*v (S aliases R), ||s||^2           REVCOM:311-313
template <int NS>
__global__ __launch_bounds__(kBlock) void k_bicg_s(int64_t seg, double* __restrict__ r,
                                                   const double* __restrict__ v, double* part, KState* st,
                                                   int64_t bs) {
  st = bofs(st, bs);
  if (!kact(st)) return;
  r = bofs(r, bs);
  v = bofs(v, bs);
  part = bofs(part, bs);
  const double alpha = st->alpha;
  double t[NS][kSlots], rr[NS][kSlots], vv[NS][kSlots];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {  // loads first (clamped past the segment's end)
      const int64_t q = tile_elem(seg, s, j), qc = q >= 0 ? q : (int64_t)s * seg + seg - 1;
      rr[s][j] = r[qc];
      vv[s][j] = v[qc];
    }
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t q = tile_elem(seg, s, j);
      double sv = 0.0;
      if (q >= 0) {
        sv = rr[s][j] + (-alpha) * vv[s][j];
        r[q] = sv;
      }
      t[s][j] = sv * sv;
    }
  tile_publish<NS, 1>(t, part);
}

// x = x + alpha*phat + omega*shat; r = s - omega*t; ||r||^2, (rtld,r)
// REVCOM:354-361 and the next iteration's rho (:263).  With the identity
// preconditioner shat aliases r (and phat aliases p), so r and sh carry no
// __restrict__; each element is read before it is written by its own thread.
// After an early exit (||s|| <= tol, REVCOM:311-316; every kernel since the
// s-update has skipped) the sweep makes the exit's update x = x + alpha*phat
// (:314) instead, and its last workgroup to finish (relaxed ticket) marks
// the solve done: every workgroup read the flags before taking its ticket,
// and nothing here reads x or the flags afterwards, so no fence is needed.
template <int NS>
__global__ __launch_bounds__(kBlock) void k_bicg_xr(int64_t seg, double* __restrict__ x, double* r,
                                                    const double* __restrict__ rt, const double* __restrict__ ph,
                                                    const double* sh, const double* __restrict__ t,
                                                    double* part, KState* st, int64_t bs) {
  st = bofs(st, bs);
  if (st->done) return;
  x = bofs(x, bs);
  r = bofs(r, bs);
  rt = bofs(rt, bs);
  ph = bofs(ph, bs);
  sh = bofs(sh, bs);
  t = bofs(t, bs);
  part = bofs(part, bs);
  const double alpha = st->alpha, omega = st->omega;
  if (st->early) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int64_t q = tile_elem(seg, s, j);
        if (q >= 0) x[q] = x[q] + alpha * ph[q];
      }
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&st->eticket, 1u) == gridDim.x - 1) {
      st->eticket = 0;
      st->done = 1;
    }
    return;
  }
  // every slot's six loads first (clamped past the segment's end, values
  // unused there), all in flight together; each element is still read
  // before its own thread writes it, which the aliasing above needs
  double lx[NS][kSlots], lph[NS][kSlots], lsh[NS][kSlots], lr[NS][kSlots], lt[NS][kSlots], lrt[NS][kSlots];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t q = tile_elem(seg, s, j), qc = q >= 0 ? q : (int64_t)s * seg + seg - 1;
      lx[s][j] = x[qc];
      lph[s][j] = ph[qc];
      lsh[s][j] = sh[qc];
      lr[s][j] = r[qc];
      lt[s][j] = t[qc];
      lrt[s][j] = rt[qc];
    }
  double tm[NS * 2][kSlots];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t q = tile_elem(seg, s, j);
      double rv = 0.0, rtv = 0.0;
      if (q >= 0) {
        double xv = lx[s][j] + alpha * lph[s][j];
        xv = xv + omega * lsh[s][j];
        x[q] = xv;
        rv = lr[s][j] + (-omega) * lt[s][j];
        r[q] = rv;
        rtv = lrt[s][j];
      }
      tm[s * 2][j] = rv * rv;
      tm[s * 2 + 1][j] = rtv * rv;
    }
  tile_publish<NS, 2>(tm, part);
}

__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b,
                                                 int64_t bs) {
  a = bofs(a, bs);
  b = bofs(b, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q < n) b[q] = a[q];
}

__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double alpha, const double* __restrict__ x,
                                                  double beta, double* __restrict__ y, int64_t bs) {
  x = bofs(x, bs);
  y = bofs(y, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q < n) y[q] = alpha * x[q] + beta * y[q];
}

__global__ __launch_bounds__(kBlock) void k_cell_axpy(int64_t n, const double* __restrict__ a,
                                                      double* __restrict__ b, int64_t bs) {
  a = bofs(a, bs);
  b = bofs(b, bs);
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q < n) b[q] = a[q] + b[q];
}

// ||b - ax||^2 and ||b||^2 into ks->red[0..1] (true-residual check)
template <int NS>
__global__ __launch_bounds__(kBlock) void k_resid_norms(int64_t seg, const double* __restrict__ b,
                                                        const double* __restrict__ ax, double* part,
                                                        KState* st, int64_t bs) {
  (void)st;
  b = bofs(b, bs);
  ax = bofs(ax, bs);
  part = bofs(part, bs);
  double t[NS * 2][kSlots];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t q = tile_elem(seg, s, j);
      const double bv = q >= 0 ? b[q] : 0.0;
      const double d = q >= 0 ? bv - ax[q] : 0.0;
      t[s * 2][j] = d * d;
      t[s * 2 + 1][j] = bv * bv;
    }
  tile_publish<NS, 2>(t, part);
}

// ---- FGMRES orthogonalisation ------------------------------------------------
// res[i] = (w, V_i) for i < nv, V_i = V + i*ldv.  One sweep reads w once and
// each V_i once; partials part[(tile*NS + s)*nv + i], finished by k_finalize.
template <int NS>
__global__ __launch_bounds__(kBlock) void k_multidot(int64_t seg, const double* __restrict__ w,
                                                     const double* __restrict__ V, int64_t ldv, int nv,
                                                     double* part) {
  double wl[NS][kSlots];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t q = tile_elem(seg, s, j);
      wl[s][j] = q >= 0 ? w[q] : 0.0;
    }
  for (int i = 0; i < nv; ++i) {
    double t[NS][kSlots], p[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int64_t q = tile_elem(seg, s, j);
        t[s][j] = q >= 0 ? wl[s][j] * V[i * ldv + q] : 0.0;
      }
    tile_pairwise<NS>(t, p);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int s = 0; s < NS; ++s) part[((size_t)tile_id() * NS + s) * nv + i] = p[s];
    }
  }
}

// w = w + sign * sum_i h[i] V_i ; if want_norm, res[0] = ||w||^2
template <int NS>
__global__ __launch_bounds__(kBlock) void k_multiaxpy(int64_t seg, double* __restrict__ w,
                                                      const double* __restrict__ V, int64_t ldv, int nv,
                                                      const double* __restrict__ h, double sign, double* part,
                                                      int want_norm) {
  double t[NS][kSlots];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int64_t q = tile_elem(seg, s, j);
      double sv = 0.0;
      if (q >= 0) {
        sv = w[q];
        for (int i = 0; i < nv; ++i) sv = sv + (sign * h[i]) * V[i * ldv + q];
        w[q] = sv;
      }
      t[s][j] = sv * sv;
    }
  if (want_norm) tile_publish<NS, 1>(t, part);
}

// Canonical total of the preceding kernel's tile partials; see kernels.hpp.
// kFinBlock threads: thread t adds the aligned run [t P, (t+1) P) of the
// tiles padded to a power of two ntp, P = max(1, ntp / kFinBlock), in runs of
// 8 (then the run sums), the waves add their threads' sums (wave_pairwise)
// and thread 0 the 16 wave sums, all pairwise: the same tree as any other
// aligned split of the padded range.  Two values are summed per pass (their
// runs and run sums fill the 128 registers a thread of 1024 may hold), every
// load of a pass issued before any is used (indices clamped, values masked).
// kFinMaxK bounds ns * nv: FGMRES orthogonalises against at most restart + 2
// <= 202 vectors (drd_fgmres accepts restart <= 200) of at most 2 segments.
constexpr int kFinBlock = 1024, kFinWaves = kFinBlock / 64, kFinMaxK = 512, kFinPer = 2;
__device__ __forceinline__ double pairwise8(const double (&r)[8]) {
  return ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
}
__global__ __launch_bounds__(kFinBlock) void k_finalize(const double* __restrict__ part, uint32_t nt, int ns,
                                                        int nv, double* out, KState* st, int phase, int guard,
                                                        int64_t bs) {
  st = bofs(st, bs);
  if (guard == FIN_ACTIVE && (st->done || st->early)) return;
  part = bofs(part, bs);
  out = bofs(out, bs);
  __shared__ double lw[kFinPer][kFinWaves];
  __shared__ double res[kFinMaxK];
  const int K = ns * nv;
  uint32_t ntp = 1;
  while (ntp < nt) ntp <<= 1;
  const uint32_t P = ntp > (uint32_t)kFinBlock ? ntp / kFinBlock : 1u;  // 1 .. 64
  const uint32_t runs = P / 8u, len = P;  // runs: P > 8; len: P <= 8
  const uint32_t base = threadIdx.x * P;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k0 = 0; k0 < K; k0 += kFinPer) {
    double tv[kFinPer];
    if (P <= 8u) {  // one run per thread (up to 8192 tiles): one round of loads for both values
      double r[kFinPer][8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t q = base + (uint32_t)i;
        const bool ok = (uint32_t)i < len && q < nt;
        const uint32_t qc = ok ? q : 0u;
#pragma unroll
        for (int v = 0; v < kFinPer; ++v) {
          const int k = k0 + v < K ? k0 + v : k0;
          const double x = part[(size_t)qc * K + k];
          r[v][i] = ok ? x : 0.0;
        }
      }
#pragma unroll
      for (int v = 0; v < kFinPer; ++v) tv[v] = pairwise8(r[v]);
    } else {  // larger grids: runs of 8, one value at a time
#pragma unroll
      for (int v = 0; v < kFinPer; ++v) {
        const int k = k0 + v < K ? k0 + v : k0;
        double g[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (uint32_t gi = 0; gi < runs; ++gi) {
          double r[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const uint32_t q = base + gi * 8u + (uint32_t)i;
            const double x = part[(size_t)(q < nt ? q : 0u) * K + k];
            r[i] = q < nt ? x : 0.0;
          }
          const double sv = pairwise8(r);
#pragma unroll
          for (int gj = 0; gj < 8; ++gj)
            if ((uint32_t)gj == gi) g[gj] = sv;
        }
        tv[v] = pairwise8(g);
      }
    }
#pragma unroll
    for (int v = 0; v < kFinPer; ++v) {
      const double sv = wave_pairwise(tv[v]);
      if (lane == 0) lw[v][w] = sv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int v = 0; v < kFinPer; ++v)
        if (k0 + v < K) res[k0 + v] = pairwise_fixed<kFinWaves>(lw[v]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double tot[2] = {0.0, 0.0};
    for (int v = 0; v < nv; ++v) {
      double s = res[v];
      for (int sg = 1; sg < ns; ++sg) {
        const double x = res[sg * nv + v];
        s = s + x;
      }
      if (phase == PH_NONE)
        out[v] = s;
      else if (v < 2)
        tot[v] = s;
    }
    if (phase != PH_NONE) kfinalize(st, phase, tot);
  }
}

__global__ __launch_bounds__(kBlock) void k_scale_into(int64_t n, const double* a, double s, double* b) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q < n) b[q] = a[q] / s;
}

static inline unsigned nb(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// Reductions see a vector of n = nseg * seg elements (nseg 1 or 2) and run
// one workgroup per kTile elements of a segment (times the batch group's
// instances; the FGMRES kernels run single-instance).
#define DRD_SEG_LAUNCH(kern, n, nseg, ...)                                                                    \
  do {                                                                                                        \
    const int64_t seg_ = (n) / (nseg);                                                                        \
    if ((nseg) == 2)                                                                                          \
      hipLaunchKernelGGL(kern<2>, dim3(nb(seg_, kTile), 1, bat_n()), dim3(kBlock), 0, s, seg_, __VA_ARGS__, \
                         bat_bs());                                                                           \
    else                                                                                                      \
      hipLaunchKernelGGL(kern<1>, dim3(nb(seg_, kTile), 1, bat_n()), dim3(kBlock), 0, s, seg_, __VA_ARGS__, \
                         bat_bs());                                                                           \
  } while (0)
#define DRD_SEG_LAUNCH1(kern, n, nseg, ...)                                                          \
  do {                                                                                               \
    const int64_t seg_ = (n) / (nseg);                                                               \
    if ((nseg) == 2)                                                                                 \
      hipLaunchKernelGGL(kern<2>, dim3(nb(seg_, kTile)), dim3(kBlock), 0, s, seg_, __VA_ARGS__);    \
    else                                                                                             \
      hipLaunchKernelGGL(kern<1>, dim3(nb(seg_, kTile)), dim3(kBlock), 0, s, seg_, __VA_ARGS__);    \
  } while (0)

void launch_finalize(hipStream_t s, const double* part, int64_t ntiles, int nseg, int nv, double* out, KState* ks,
                     int phase, int guard) {
  hipLaunchKernelGGL(k_finalize, dim3(1, 1, bat_n()), dim3(kFinBlock), 0, s, part, (uint32_t)ntiles, nseg, nv,
                     out, ks, phase, guard, bat_bs());
}
static void launch_finalize1(hipStream_t s, const double* part, int64_t ntiles, int nseg, int nv, double* out) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, part, (uint32_t)ntiles, nseg, nv, out,
                     (KState*)nullptr, (int)PH_NONE, (int)FIN_ALWAYS, (int64_t)0);
}

void launch_multidot(hipStream_t s, int64_t n, int nseg, const double* w, const double* V, int64_t ldv, int nv,
                     double* part, double* res) {
  DRD_SEG_LAUNCH1(k_multidot, n, nseg, w, V, ldv, nv, part);
  launch_finalize1(s, part, tiles_of(n / nseg), nseg, nv, res);
}
void launch_multiaxpy(hipStream_t s, int64_t n, int nseg, double* w, const double* V, int64_t ldv, int nv,
                      const double* h, double sign, double* part, double* res, int want_norm) {
  DRD_SEG_LAUNCH1(k_multiaxpy, n, nseg, w, V, ldv, nv, h, sign, part, want_norm);
  if (want_norm) launch_finalize1(s, part, tiles_of(n / nseg), nseg, 1, res);
}
void launch_scale_into(hipStream_t s, int64_t n, const double* a, double sc, double* b) {
  hipLaunchKernelGGL(k_scale_into, dim3(nb(n, kBlock)), dim3(kBlock), 0, s, n, a, sc, b);
}

void launch_init_norms(hipStream_t s, int64_t n, int nseg, const double* x, const double* b, double* part,
                       KState* ks) {
  DRD_SEG_LAUNCH(k_init_norms, n, nseg, x, b, part, ks);
  launch_finalize(s, part, tiles_of(n / nseg), nseg, 2, nullptr, ks, PH_INIT_NORMS, FIN_ALWAYS);
}
void launch_bicg_p(hipStream_t s, int64_t n, const double* r, const double* v, double* p, KState* ks) {
  hipLaunchKernelGGL(k_bicg_p, dim3(nb(n, kTile), 1, bat_n()), dim3(kBlock), 0, s, n, r, v, p, ks, bat_bs());
}
void launch_bicg_s(hipStream_t s, int64_t n, int nseg, double* r, const double* v, double* part, KState* ks) {
  DRD_SEG_LAUNCH(k_bicg_s, n, nseg, r, v, part, ks);
  launch_finalize(s, part, tiles_of(n / nseg), nseg, 1, nullptr, ks, PH_S, FIN_ACTIVE);
}
void launch_bicg_xr(hipStream_t s, int64_t n, int nseg, double* x, double* r, const double* rt, const double* ph,
                    const double* sh, const double* t, double* part, KState* ks, bool finish) {
  DRD_SEG_LAUNCH(k_bicg_xr, n, nseg, x, r, rt, ph, sh, t, part, ks);
  if (finish) launch_finalize(s, part, tiles_of(n / nseg), nseg, 2, nullptr, ks, PH_XR, FIN_ACTIVE);
}
// The reductions a fused matvec sweep makes, for a product y = A x formed
// elsewhere (a caller's MATVEC on the host, copied in): MV_INIT a0 := y and
// (y, y); MV_DOT1 (a0, y); MV_DOT2 (y, a0), (y, y).  Same tile partials and
// the same products as the fused sweeps, so the finish is theirs.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_ext_mv(int64_t n, const double* __restrict__ y, double* __restrict__ a0,
                                                   double* part, KState* st) {
  if (!kact(st)) return;
  constexpr int K = MODE == MV_DOT2 ? 2 : 1;
  double t[K][kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int64_t q = tile_elem(n, 0, j);
    double z = 0.0, a = 0.0;
    if (q >= 0) {
      z = y[q];
      if (MODE == MV_INIT)
        a0[q] = z;
      else
        a = a0[q];
    }
    if (MODE == MV_INIT) t[0][j] = z * z;
    if (MODE == MV_DOT1) t[0][j] = a * z;
    if (MODE == MV_DOT2) {
      t[0][j] = z * a;
      t[1][j] = z * z;
    }
  }
  tile_publish<1, K>(t, part);
}

void launch_ext_mv(hipStream_t s, int mode, int64_t n, const double* y, double* a0, double* part, KState* ks) {
  const dim3 grid(tiles_of(n));
  if (mode == MV_INIT) hipLaunchKernelGGL(k_ext_mv<MV_INIT>, grid, dim3(kBlock), 0, s, n, y, a0, part, ks);
  if (mode == MV_DOT1) hipLaunchKernelGGL(k_ext_mv<MV_DOT1>, grid, dim3(kBlock), 0, s, n, y, a0, part, ks);
  if (mode == MV_DOT2) hipLaunchKernelGGL(k_ext_mv<MV_DOT2>, grid, dim3(kBlock), 0, s, n, y, a0, part, ks);
  launch_mv_finalize(s, mode, part, tiles_of(n), 1, ks);
}

void launch_copy(hipStream_t s, int64_t n, const double* a, double* b) {
  hipLaunchKernelGGL(k_copy, dim3(nb(n, kBlock), 1, bat_n()), dim3(kBlock), 0, s, n, a, b, bat_bs());
}
void launch_axpby(hipStream_t s, int64_t n, double alpha, const double* x, double beta, double* y) {
  hipLaunchKernelGGL(k_axpby, dim3(nb(n, kBlock), 1, bat_n()), dim3(kBlock), 0, s, n, alpha, x, beta, y, bat_bs());
}
void launch_cell_axpy(hipStream_t s, int64_t n, const double* a, double* b) {
  hipLaunchKernelGGL(k_cell_axpy, dim3(nb(n, kBlock), 1, bat_n()), dim3(kBlock), 0, s, n, a, b, bat_bs());
}
void launch_resid_norms(hipStream_t s, int64_t n, int nseg, const double* b, const double* ax, double* part,
                        KState* ks) {
  DRD_SEG_LAUNCH(k_resid_norms, n, nseg, b, ax, part, ks);
  launch_finalize(s, part, tiles_of(n / nseg), nseg, 2, ks->red, nullptr, PH_NONE, FIN_ALWAYS);
}

}  // namespace drd
