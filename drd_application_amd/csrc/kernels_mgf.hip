// kernels_mgf.hip -- LDS-tiled multigrid legs: one launch per level and leg.
//
// The V-cycle of kernels_mg.hip takes five launches per level (two sweeps,
// residual, restriction on the way down; prolongation and two sweeps on the
// way up).  Below 2^18 unknowns each of them costs about 5 us whatever its
// size, and on the finest level the five passes re-read the level's
// vectors and coefficients from HBM.  Here a workgroup owns a TW x TH block
// of the level's cells and runs a whole leg on it in LDS:
//   down: x1 = omega b / d, x2 = x1 + (b - A x1) omega / d, r = b - A x2 and
//         the restriction of r to the coarse unknowns it owns;
//   up:   x = x2 + P xc, then two weighted Jacobi sweeps
// (omega / d is stored per unknown: 0.7 / d on the finest level, 1.2 / d below).
// Each stage is evaluated on the owned unknowns plus a ring of the width the
// later stages read (3, 2, 1 going down; 2, 1, 0 going up), so the stages
// need no data from other workgroups; neighbouring blocks recompute their
// rings.  Every value is formed by the same expression in the same type as
// in the per-pass kernels (the stencil and transfers of velo_rows.hpp), so
// the cycle's result is the same bit for bit.
#include "kernels.hpp"
#include "velo_rows.hpp"

namespace drd {

constexpr int kMgfThreads = 512;

// Block geometry.  The workgroup owns cells [ci0, ci0 + TW) x [ck0, ck0 + TH),
// the v faces of those cells' rows and the u faces [ci0, ci0 + TW) (plus
// face nx on the last block of a row).  A u or v position at ring h lies
// within h positions of the owned ones (u: columns up to ci0 + TW + h).
// LDS holds positions up to ring H: u and v arrays of ROWS rows of LD,
// element 0 at (ci0 - H, ck0 - H); and the cells the stencils at ring H - 1
// read, element 0 at (ci0 - H, ck0 - H).  Thread t handles positions
// j * kMgfThreads + t (slot j) in every stage.
template <int TW, int TH, int H>
struct MgBlock {
  static constexpr int LD = TW + 2 * H + 1;
  static constexpr int ROWS = TH + 2 * H;
  static constexpr int NV = LD * ROWS;  // per vector component
  static constexpr int NP = 2 * NV;     // u positions, then v positions
  static constexpr int NS = (NP + kMgfThreads - 1) / kMgfThreads;
  static constexpr int CLD = TW + 2 * H;
  static constexpr int NC = CLD * (TH + 2 * H);
  static constexpr int NCS = (NC + kMgfThreads - 1) / kMgfThreads;
};

// Ring of position p of the block (u for p < NV); 99 when the position is
// not an unknown of the level.
struct MgPos {
  bool u;
  int i, k, ring;
  bool own;  // written by this block
};
template <int TW, int TH, int H>
__device__ __forceinline__ MgPos mg_pos(int p, int ci0, int ck0, int nx, int ny) {
  using B = MgBlock<TW, TH, H>;
  MgPos r;
  r.u = p < B::NV;
  const int l = r.u ? p : p - B::NV;
  const int lk = l / B::LD, li = l - lk * B::LD;
  r.i = ci0 - H + li;
  r.k = ck0 - H + lk;
  const int dx = r.u ? max(ci0 - r.i, r.i - (ci0 + TW)) : max(ci0 - r.i, r.i - (ci0 + TW - 1));
  const int dy = max(ck0 - r.k, r.k - (ck0 + TH - 1));
  const bool valid = p < B::NP && (r.u ? (r.i >= 0 && r.i <= nx && r.k >= 0 && r.k <= ny - 1)
                                       : (r.i >= 0 && r.i <= nx - 1 && r.k >= 1 && r.k <= ny - 1));
  r.ring = valid ? max(0, max(dx, dy)) : 99;
  r.own = r.ring == 0 && (!r.u || r.i < ci0 + TW || r.i == nx);
  return r;
}
__device__ __forceinline__ uint32_t mg_packed(const Grid& g, bool u, int i, int k) {
  return u ? (uint32_t)(i + (g.nx + 1) * k) : (uint32_t)(g.mn() + i + g.nx * (k - 1));
}

// The block's cell coefficients: loads issued together into registers,
// then stored to LDS (cells outside (-1:nx, -1:ny) hold the value of cell
// (-1, -1); no stencil of the block reads them).
template <class T, int TW, int TH, int H>
struct MgCoefStage {
  using B = MgBlock<TW, TH, H>;
  T r[B::NCS], e[B::NCS], x[B::NCS];
  __device__ __forceinline__ void load(const Grid& g, const T* __restrict__ rho, const T* __restrict__ eta,
                                       const T* __restrict__ chi, int ci0, int ck0) {
#pragma unroll
    for (int it = 0; it < B::NCS; ++it) {
      const int p = it * kMgfThreads + (int)threadIdx.x;
      const int lk = p / B::CLD, li = p - lk * B::CLD;
      const int i = ci0 - H + li, k = ck0 - H + lk;
      const bool in = p < B::NC && i >= -1 && i <= g.nx && k >= -1 && k <= g.ny;
      const uint32_t c = in ? g.cio(i, k) : 0u;
      r[it] = ld32(rho, c);
      e[it] = ld32(eta, c);
      x[it] = ld32(chi, c);
    }
  }
  __device__ __forceinline__ void store(T* sc) const {
#pragma unroll
    for (int it = 0; it < B::NCS; ++it) {
      const int p = it * kMgfThreads + (int)threadIdx.x;
      if (p < B::NC) {
        sc[p] = r[it];
        sc[B::NC + p] = e[it];
        sc[2 * B::NC + p] = x[it];
      }
    }
  }
};

// Positions of the block: slot j of thread t is position j * kMgfThreads + t.
template <int TW, int TH, int H>
struct MgSlots {
  using B = MgBlock<TW, TH, H>;
  MgPos m[B::NS];
  __device__ __forceinline__ MgSlots(int ci0, int ck0, int nx, int ny) {
#pragma unroll
    for (int j = 0; j < B::NS; ++j) m[j] = mg_pos<TW, TH, H>(j * kMgfThreads + (int)threadIdx.x, ci0, ck0, nx, ny);
  }
};

// One stencil evaluation at slot j of the block.  FP64 levels: velo_u_at /
// velo_v_at.  FP32 levels: the regrouped form of velo_rows.hpp, with the
// corner viscosities of the staged cells in LDS and each slot's centre
// coefficient D in a register, both formed once per leg.
template <class T>
struct MgEvalRef {
  const ScalT<T>& s;
  const CoefL<T>& C;
  template <class XV>
  __device__ __forceinline__ double operator()(const XV& X, const MgPos& m, int) const {
    return m.u ? (double)velo_u_at<2, T>(s, C, X, m.i, m.k) : (double)velo_v_at<2, T>(s, C, X, m.i, m.k);
  }
};
template <class T, int NS>
struct MgEvalPc {
  PcK<T> K;
  const CoefL<T>& C;
  PcQArr<T> Q;
  T d[NS];
  template <class XV>
  __device__ __forceinline__ double operator()(const XV& X, const MgPos& m, int j) const {
    return m.u ? (double)velo_u_pc(K, C, Q, X, m.i, m.k, d[j]) : (double)velo_v_pc(K, C, Q, X, m.i, m.k, d[j]);
  }
};
// Forms the corners (into sq, C's index space: the corner north-east of each
// staged cell whose three neighbours there are staged; 0 elsewhere, where no
// stencil of the leg reads it) and D at the slots of rings <= RMAX.  Ends
// with a barrier; the cells must be staged and visible.
template <class T, int TW, int TH, int H, int RMAX>
__device__ __forceinline__ MgEvalPc<T, MgBlock<TW, TH, H>::NS> mg_eval_pc(const ScalT<T>& s, const CoefL<T>& C,
                                                                          T* sq, const MgSlots<TW, TH, H>& P) {
  using B = MgBlock<TW, TH, H>;
  constexpr int CROWS = B::NC / B::CLD;
#pragma unroll
  for (int it = 0; it < B::NCS; ++it) {
    const int p = it * kMgfThreads + (int)threadIdx.x;
    if (p >= B::NC) continue;
    const int lk = p / B::CLD, li = p - lk * B::CLD;
    sq[p] = (li < B::CLD - 1 && lk < CROWS - 1) ? pc_corner<T>(C, (uint32_t)p) : T(0);
  }
  __syncthreads();
  MgEvalPc<T, B::NS> A{pc_k<T>(s), C, PcQArr<T>{sq}, {}};
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    A.d[j] = m.ring > RMAX ? T(0) : (m.u ? pc_du(A.K, C, A.Q, m.i, m.k) : pc_dv(A.K, C, A.Q, m.i, m.k));
  }
  return A;
}

// Stages of the down leg after the first (x1 in sa): x2 into sb and xd, r
// over sa, restriction into bc.  XV is the LDS vector view (ghost rules, or
// none in an interior block).
template <class XV, class EV, class T, class TBC, int TW, int TH>
__device__ __forceinline__ void mg_down_stages(const Grid& g, const Grid& gc, const EV& A, const MgSlots<TW, TH, 3>& P,
                                               const XV& XA, const XV& XB, T* sa, T* sb,
                                               const T (&bv)[MgBlock<TW, TH, 3>::NS],
                                               const T (&wv)[MgBlock<TW, TH, 3>::NS], T* __restrict__ xd,
                                               TBC* __restrict__ bc, int ci0, int ck0) {
  using B = MgBlock<TW, TH, 3>;
  // x2 = x1 + (b - A x1) omega / d at rings <= 2
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    if (m.ring > 2) continue;
    const int p = j * kMgfThreads + (int)threadIdx.x;
    const double z = A(XA, m, j);
    const T x2 = (T)((double)sa[p] + ((double)bv[j] - z) * (double)wv[j]);
    sb[p] = x2;
    if (m.own) xd[mg_packed(g, m.u, m.i, m.k)] = x2;
  }
  __syncthreads();
  // r = b - A x2 at rings <= 1, over x1 (no longer read)
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    if (m.ring > 1) continue;
    const int p = j * kMgfThreads + (int)threadIdx.x;
    const double z = A(XB, m, j);
    sa[p] = (T)((double)bv[j] - z);
  }
  __syncthreads();
  // restriction to the coarse unknowns of the block: u faces I in
  // [ci0/2, ci0/2 + TW/2) (and gc.nx on the last block), v faces I in
  // [ci0/2, ci0/2 + TW/2); rows K in [ck0/2, ck0/2 + TH/2)
  const LdsUV<T> R{sa, sa + B::NV, ci0 - 3, ck0 - 3, B::LD};
  constexpr int CU = (TW / 2 + 1) * (TH / 2), CV = (TW / 2) * (TH / 2);
  for (int p = threadIdx.x; p < CU + CV; p += kMgfThreads) {
    const bool u = p < CU;
    const int l = u ? p : p - CU;
    const int w = u ? TW / 2 + 1 : TW / 2;
    const int lk = l / w, li = l - lk * w;
    const int I = ci0 / 2 + li, K = ck0 / 2 + lk;
    if (u) {
      if (K > gc.ny - 1 || I > gc.nx || (li == TW / 2 && I != gc.nx)) continue;
      bc[mg_uidx(gc, I, K)] = (TBC)(0.25 * mg_restrict_u(g, gc, R, I, K));
    } else {
      if (K < 1 || K > gc.ny - 1 || I > gc.nx - 1) continue;
      bc[mg_vidx(gc, I, K)] = (TBC)(0.25 * mg_restrict_v(g, gc, R, I, K));
    }
  }
}

// A block is interior when every position up to ring H is an unknown of the
// grid at which the ghost rules map each read to itself.
template <int TW, int TH, int H>
__device__ __forceinline__ bool mg_interior(const Grid& g, int ci0, int ck0) {
  return ci0 - H >= 0 && ci0 + TW + H <= g.nx - 1 && ck0 - H >= 1 && ck0 + TH - 1 + H <= g.ny - 1;
}

// Down leg of level g (coarse level gc): x2 to xd, bc = R (b - A x2); bcopy
// (if given) receives b as FP32 at the owned unknowns.
template <class T, class TB, class TBC, int TW, int TH>
__global__ __launch_bounds__(kMgfThreads) void k_mg_down(Grid g, Grid gc, Scal s, const T* __restrict__ rho,
                                                         const T* __restrict__ eta, const T* __restrict__ chi,
                                                         const TB* __restrict__ b, const T* __restrict__ dw,
                                                         T* __restrict__ xd, float* __restrict__ bcopy,
                                                         TBC* __restrict__ bc, const KState* st, int64_t bs) {
  constexpr int H = 3;
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  chi = bofs(chi, bs);
  b = bofs(b, bs);
  dw = bofs(dw, bs);
  xd = bofs(xd, bs);
  bcopy = bofs(bcopy, bs);
  bc = bofs(bc, bs);
  st = bofs(st, bs);
  using B = MgBlock<TW, TH, H>;
  constexpr bool PC = sizeof(T) == sizeof(float);
  __shared__ T sc[3 * B::NC];
  __shared__ T sq[PC ? B::NC : 1];
  __shared__ T sa[B::NP];
  __shared__ T sb[B::NP];
  const int ntx = (g.nx + TW - 1) / TW;
  const int tile = (int)tile_id();
  const int ty = tile / ntx, tx = tile - ty * ntx;
  const int ci0 = tx * TW, ck0 = ty * TH;
  const MgSlots<TW, TH, H> P(ci0, ck0, g.nx, g.ny);
  // every load of the first stage issued before the first wait
  MgCoefStage<T, TW, TH, H> cs;
  cs.load(g, rho, eta, chi, ci0, ck0);
  TB braw[B::NS];
  T wv[B::NS];
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    const uint32_t q = m.ring <= 3 ? mg_packed(g, m.u, m.i, m.k) : 0u;
    braw[j] = ld32(b, q);
    wv[j] = ld32(dw, q);
  }
  if (st && (st->done || st->early)) return;
  cs.store(sc);
  // x1 = omega b / d at rings <= 3
  T bv[B::NS];
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    bv[j] = (T)braw[j];
    if (m.ring > 3) continue;
    sa[j * kMgfThreads + (int)threadIdx.x] = (T)((double)braw[j] * (double)wv[j]);
    if (bcopy && m.own) bcopy[mg_packed(g, m.u, m.i, m.k)] = (float)braw[j];
  }
  __syncthreads();
  const CoefL<T> C{sc, sc + B::NC, sc + 2 * B::NC, ci0 - H, ck0 - H, (uint32_t)B::CLD};
  const ScalT<T> sT = ScalT<T>::make(s);  // converted once, not at every unknown
  auto legs = [&](const auto& A) {
    if (mg_interior<TW, TH, H>(g, ci0, ck0)) {
      const VeloLdsIn<T> XA{sa, sa + B::NV, ci0 - H, ck0 - H, B::LD};
      const VeloLdsIn<T> XB{sb, sb + B::NV, ci0 - H, ck0 - H, B::LD};
      mg_down_stages(g, gc, A, P, XA, XB, sa, sb, bv, wv, xd, bc, ci0, ck0);
    } else {
      const VeloLds<T> XA{sa, sa + B::NV, g.nx, g.ny, s.lr_d, ci0 - H, ck0 - H, B::LD};
      const VeloLds<T> XB{sb, sb + B::NV, g.nx, g.ny, s.lr_d, ci0 - H, ck0 - H, B::LD};
      mg_down_stages(g, gc, A, P, XA, XB, sa, sb, bv, wv, xd, bc, ci0, ck0);
    }
  };
  if constexpr (PC)
    legs(mg_eval_pc<T, TW, TH, H, 2>(sT, C, sq, P));  // stencils at rings <= 2
  else
    legs(MgEvalRef<T>{sT, C});
}

// Stages of the up leg after the first (x2 + P xc in sa): a sweep into sb,
// a sweep into out.
template <class XV, class EV, class T, class TO, int TW, int TH>
__device__ __forceinline__ void mg_up_stages(const Grid& g, const EV& A, const MgSlots<TW, TH, 2>& P,
                                             const XV& XA, const XV& XB, T* sa, T* sb,
                                             const T (&bv)[MgBlock<TW, TH, 2>::NS],
                                             const T (&wv)[MgBlock<TW, TH, 2>::NS], TO* __restrict__ out) {
  using B = MgBlock<TW, TH, 2>;
  // first sweep at rings <= 1
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    if (m.ring > 1) continue;
    const int p = j * kMgfThreads + (int)threadIdx.x;
    const double z = A(XA, m, j);
    sb[p] = (T)((double)sa[p] + ((double)bv[j] - z) * (double)wv[j]);
  }
  __syncthreads();
  // second sweep at the owned unknowns
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    if (!m.own) continue;
    const int p = j * kMgfThreads + (int)threadIdx.x;
    const double z = A(XB, m, j);
    out[mg_packed(g, m.u, m.i, m.k)] = (TO)((double)sb[p] + ((double)bv[j] - z) * (double)wv[j]);
  }
}

// Up leg of level g (coarse level gc): x = xd + P xc, two weighted Jacobi
// sweeps, the second written to out.
template <class T, class TB, class TC, class TO, int TW, int TH>
__global__ __launch_bounds__(kMgfThreads) void k_mg_up(Grid g, Grid gc, Scal s, const T* __restrict__ rho,
                                                       const T* __restrict__ eta, const T* __restrict__ chi,
                                                       const TB* __restrict__ b, const T* __restrict__ dw,
                                                       const T* __restrict__ xd, const TC* __restrict__ xc,
                                                       TO* __restrict__ out, const KState* st, int64_t bs) {
  constexpr int H = 2;
  rho = bofs(rho, bs);
  eta = bofs(eta, bs);
  chi = bofs(chi, bs);
  b = bofs(b, bs);
  dw = bofs(dw, bs);
  xd = bofs(xd, bs);
  xc = bofs(xc, bs);
  out = bofs(out, bs);
  st = bofs(st, bs);
  using B = MgBlock<TW, TH, H>;
  constexpr bool PC = sizeof(T) == sizeof(float);
  __shared__ T sc[3 * B::NC];
  __shared__ T sq[PC ? B::NC : 1];
  __shared__ T sa[B::NP];
  __shared__ T sb[B::NP];
  const int ntx = (g.nx + TW - 1) / TW;
  const int tile = (int)tile_id();
  const int ty = tile / ntx, tx = tile - ty * ntx;
  const int ci0 = tx * TW, ck0 = ty * TH;
  const MgSlots<TW, TH, H> P(ci0, ck0, g.nx, g.ny);
  MgCoefStage<T, TW, TH, H> cs;
  cs.load(g, rho, eta, chi, ci0, ck0);
  T xdv[B::NS], bv[B::NS], wv[B::NS];
  TC xcv[B::NS][4];
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    const bool in = m.ring <= 2;
    const uint32_t q = in ? mg_packed(g, m.u, m.i, m.k) : 0u;
    xdv[j] = ld32(xd, q);
    bv[j] = (T)ld32(b, q);
    wv[j] = ld32(dw, q);
    int I[4], K[4];
    if (m.u)
      mg_prolong_u_at(gc, m.i, m.k, I, K);
    else
      mg_prolong_v_at(gc, m.i, m.k, I, K);
#pragma unroll
    for (int t = 0; t < 4; ++t) xcv[j][t] = xc[in ? (m.u ? mg_uidx(gc, I[t], K[t]) : mg_vidx(gc, I[t], K[t])) : 0];
  }
  if (st && (st->done || st->early)) return;
  cs.store(sc);
  // x = x2 + P xc at rings <= 2
#pragma unroll
  for (int j = 0; j < B::NS; ++j) {
    const MgPos& m = P.m[j];
    if (m.ring > 2) continue;
    const double xv[4] = {(double)xcv[j][0], (double)xcv[j][1], (double)xcv[j][2], (double)xcv[j][3]};
    const double acc = m.u ? mg_prolong_u_sum(gc, m.i, m.k, xv) : mg_prolong_v_sum(gc, m.i, m.k, xv);
    sa[j * kMgfThreads + (int)threadIdx.x] = (T)((double)xdv[j] + acc);
  }
  __syncthreads();
  const CoefL<T> C{sc, sc + B::NC, sc + 2 * B::NC, ci0 - H, ck0 - H, (uint32_t)B::CLD};
  const ScalT<T> sT = ScalT<T>::make(s);  // converted once, not at every unknown
  auto legs = [&](const auto& A) {
    if (mg_interior<TW, TH, H>(g, ci0, ck0)) {
      const VeloLdsIn<T> XA{sa, sa + B::NV, ci0 - H, ck0 - H, B::LD};
      const VeloLdsIn<T> XB{sb, sb + B::NV, ci0 - H, ck0 - H, B::LD};
      mg_up_stages(g, A, P, XA, XB, sa, sb, bv, wv, out);
    } else {
      const VeloLds<T> XA{sa, sa + B::NV, g.nx, g.ny, s.lr_d, ci0 - H, ck0 - H, B::LD};
      const VeloLds<T> XB{sb, sb + B::NV, g.nx, g.ny, s.lr_d, ci0 - H, ck0 - H, B::LD};
      mg_up_stages(g, A, P, XA, XB, sa, sb, bv, wv, out);
    }
  };
  if constexpr (PC)
    legs(mg_eval_pc<T, TW, TH, H, 1>(sT, C, sq, P));  // stencils at rings <= 1
  else
    legs(MgEvalRef<T>{sT, C});
}

// Block shapes: 32 x 16 cells for levels of nx >= 1024 (FP64: 48 KB of LDS
// going down), 16 x 8 below (more blocks where the level is small).  For
// the FP32 levels 32 x 16 beat 32 x 32, 64 x 16, 32 x 24, 48 x 16 and 32 x 8
// (V-cycle at 2048 x 1024: 199 us against 209-216 us); on the small levels
// 16 x 8 beat 16 x 16, 8 x 8 and 32 x 8 (V-cycle at 256 x 128: 47.9 us
// against 52.5 us with 16 x 16).
template <class T, bool SMALL>
struct MgfShape;
template <>
struct MgfShape<float, false> {
  static constexpr int TW = 32, TH = 16;
};
template <>
struct MgfShape<double, false> {
  static constexpr int TW = 32, TH = 16;
};
template <class T>
struct MgfShape<T, true> {
  static constexpr int TW = 16, TH = 8;
};
static bool mgf_small(const Grid& g) { return g.nx < 1024; }

template <class S>
static unsigned mgf_blocks(const Grid& g) {
  return (unsigned)(((g.nx + S::TW - 1) / S::TW) * ((g.ny + S::TH - 1) / S::TH));
}

template <class T, class TB, class TBC, bool SMALL>
static void mg_down_s(hipStream_t st, const Grid& g, const Grid& gc, const Scal& s, const MgCoef& c, const void* b,
                      const void* dw, void* xd, float* bcopy, void* bc, const KState* ks) {
  using S = MgfShape<T, SMALL>;
  hipLaunchKernelGGL((k_mg_down<T, TB, TBC, S::TW, S::TH>), dim3(mgf_blocks<S>(g), 1, bat_n()), dim3(kMgfThreads), 0,
                     st, g, gc, s, (const T*)c.rho, (const T*)c.eta, (const T*)c.chi, (const TB*)b, (const T*)dw,
                     (T*)xd, bcopy, (TBC*)bc, ks, bat_bs());
}
template <class T, class TB, class TBC>
static void mg_down_t(hipStream_t st, const Grid& g, const Grid& gc, const Scal& s, const MgCoef& c, const void* b,
                      const void* dw, void* xd, float* bcopy, void* bc, const KState* ks) {
  if (mgf_small(g))
    mg_down_s<T, TB, TBC, true>(st, g, gc, s, c, b, dw, xd, bcopy, bc, ks);
  else
    mg_down_s<T, TB, TBC, false>(st, g, gc, s, c, b, dw, xd, bcopy, bc, ks);
}

void launch_mg_down(hipStream_t st, bool f32, bool b32, bool bc32, const Grid& g, const Grid& gc, const Scal& s,
                    const MgCoef& c, const void* b, const void* dw, void* xd, float* bcopy, void* bc,
                    const KState* ks) {
  if (!f32)
    mg_down_t<double, double, double>(st, g, gc, s, c, b, dw, xd, bcopy, bc, ks);
  else if (!b32)
    bc32 ? mg_down_t<float, double, float>(st, g, gc, s, c, b, dw, xd, bcopy, bc, ks)
         : mg_down_t<float, double, double>(st, g, gc, s, c, b, dw, xd, bcopy, bc, ks);
  else
    bc32 ? mg_down_t<float, float, float>(st, g, gc, s, c, b, dw, xd, bcopy, bc, ks)
         : mg_down_t<float, float, double>(st, g, gc, s, c, b, dw, xd, bcopy, bc, ks);
}

template <class T, class TB, class TC, class TO, bool SMALL>
static void mg_up_s(hipStream_t st, const Grid& g, const Grid& gc, const Scal& s, const MgCoef& c, const void* b,
                    const void* dw, const void* xd, const void* xc, void* out, const KState* ks) {
  using S = MgfShape<T, SMALL>;
  hipLaunchKernelGGL((k_mg_up<T, TB, TC, TO, S::TW, S::TH>), dim3(mgf_blocks<S>(g), 1, bat_n()), dim3(kMgfThreads), 0,
                     st, g, gc, s, (const T*)c.rho, (const T*)c.eta, (const T*)c.chi, (const TB*)b, (const T*)dw,
                     (const T*)xd, (const TC*)xc, (TO*)out, ks, bat_bs());
}
template <class T, class TB, class TC, class TO>
static void mg_up_t(hipStream_t st, const Grid& g, const Grid& gc, const Scal& s, const MgCoef& c, const void* b,
                    const void* dw, const void* xd, const void* xc, void* out, const KState* ks) {
  if (mgf_small(g))
    mg_up_s<T, TB, TC, TO, true>(st, g, gc, s, c, b, dw, xd, xc, out, ks);
  else
    mg_up_s<T, TB, TC, TO, false>(st, g, gc, s, c, b, dw, xd, xc, out, ks);
}

void launch_mg_up(hipStream_t st, bool f32, bool c32, bool o32, const Grid& g, const Grid& gc, const Scal& s,
                  const MgCoef& c, const void* b, const void* dw, const void* xd, const void* xc, void* out,
                  const KState* ks) {
  if (!f32) {
    mg_up_t<double, double, double, double>(st, g, gc, s, c, b, dw, xd, xc, out, ks);
  } else if (c32) {
    o32 ? mg_up_t<float, float, float, float>(st, g, gc, s, c, b, dw, xd, xc, out, ks)
        : mg_up_t<float, float, float, double>(st, g, gc, s, c, b, dw, xd, xc, out, ks);
  } else {
    o32 ? mg_up_t<float, float, double, float>(st, g, gc, s, c, b, dw, xd, xc, out, ks)
        : mg_up_t<float, float, double, double>(st, g, gc, s, c, b, dw, xd, xc, out, ks);
  }
}

}  // namespace drd
