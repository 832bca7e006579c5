// fft_regs.hpp -- complex line FFTs with the butterflies in registers.
//
// One workgroup of NT threads (kBlock unless a kernel says otherwise)
// transforms a complex line of power-of-two length L <= 8192 held in LDS, in
// place.  The transform is a mixed-radix
// Stockham FFT: one pass of radix 2, 4 or 8 when log2 L is not a multiple
// of 4, then radix-16 passes.  In a pass each group j of R inputs
//   v[t] = A[j + t L/R],  t < R,
// is multiplied by the twiddles w^(t k), w = exp(+-2 pi i / (Ns R)),
// k = j mod Ns (Ns = product of the earlier radices; w^k is read from the
// table, its powers formed by multiplication), goes through an
// R-point DFT in registers and is stored at A[(j - k) R + k + t Ns].  A
// thread holds all its groups of a pass (one for lines up to 4096), so the
// whole pass is read before the barrier and written after it: one LDS
// buffer and two barriers per pass (3 passes for L = 4096 instead of 12
// radix-2 stages).
#pragma once

#include "fft_lines.hpp"

namespace drd {

// LDS position of line element i: one pad element after every 16, so the
// strided stores of a pass (element stride R or 16 Ns) fall into different
// banks while unit-stride accesses stay contiguous.  A line of L elements
// occupies lds_len(L) slots.
__host__ __device__ __forceinline__ constexpr int lp(int i) { return i + (i >> 4); }
__host__ __device__ __forceinline__ constexpr int lds_len(int L) { return lp(L); }

// cos(j pi / 8) and sin(j pi / 8), j = 0..7: the twiddles inside the
// radix-2..16 butterflies (2 pi k / m with m | 16).
__device__ __forceinline__ constexpr double cos8(int j) {
  return j == 0 ? 1.0
       : j == 1 ? 0.92387953251128675613
       : j == 2 ? 0.70710678118654752440
       : j == 3 ? 0.38268343236508977173
       : j == 4 ? 0.0
       : j == 5 ? -0.38268343236508977173
       : j == 6 ? -0.70710678118654752440
                : -0.92387953251128675613;
}
__device__ __forceinline__ constexpr double sin8(int j) { return j <= 4 ? cos8(4 - j) : cos8(j - 4); }

__device__ __forceinline__ constexpr int bitrev(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
  return r;
}

// In-register R-point DFT (R = 2, 4, 8, 16), natural order in and out:
// radix-2 decimation in time after a bit-reversal permutation.
// sign = -1 forward, +1 inverse (unnormalised).
template <class T, int R>
__device__ __forceinline__ void dft_regs(Cx<T> (&v)[R], T sign) {
  constexpr int bits = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = bitrev(i, bits);
    if (j > i) {
      const Cx<T> t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int m = 2; m <= R; m <<= 1) {
#pragma unroll
    for (int k = 0; k < m / 2; ++k) {
      const int e = k * (16 / m);  // 2 pi k / m in units of pi / 8
      const T c = T(cos8(e)), s = sign * T(sin8(e));
#pragma unroll
      for (int b = 0; b < R; b += m) {
        const Cx<T> u = v[b + k], t = v[b + k + m / 2];
        const Cx<T> tw = k == 0 ? t : Cx<T>{t.x * c - t.y * s, t.x * s + t.y * c};
        v[b + k] = Cx<T>{u.x + tw.x, u.y + tw.y};
        v[b + k + m / 2] = Cx<T>{u.x - tw.x, u.y - tw.y};
      }
    }
  }
}

// Groups of a radix-R pass one of NT threads holds for lines up to LMAX:
// the radix-R pass (R < 16) occurs only when log2 L = log2 R (mod 4).
__host__ __device__ constexpr int pass_groups(int R, int LMAX, int NT) {
  int L = 0;
  if (R == 16) {
    L = LMAX;
  } else {
    const int r = R == 2 ? 1 : R == 4 ? 2 : 3;
    for (int p = 0; (1 << p) <= LMAX; ++p)
      if ((p & 3) == r) L = 1 << p;
  }
  const int g = L / R / NT;
  return g > 1 ? g : 1;
}

// One Stockham pass of radix R over the LDS line A of length L.
// tw[m] = exp(-2 pi i m / L), m < L.
template <class T, int R, int G, int NT>
__device__ __forceinline__ void stockham_pass(Cx<T>* A, int L, int Ns, const Cx<T>* __restrict__ tw, T sign) {
  const int groups = L / R;
  Cx<T> v[G][R];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int j = (int)threadIdx.x + g * NT;
    if (j < groups) {
      const int k = j & (Ns - 1);
#pragma unroll
      for (int t = 0; t < R; ++t) v[g][t] = A[lp(j + t * groups)];
      if (Ns > 1) {  // w^t by repeated multiplication from one table entry
        Cx<T> w1 = tw[k * (L / (Ns * R))];
        if (sign > T(0)) w1.y = -w1.y;
        Cx<T> w = w1;
#pragma unroll
        for (int t = 1; t < R; ++t) {
          v[g][t] = cmulT(v[g][t], w);
          if (t + 1 < R) w = cmulT(w, w1);
        }
      }
      dft_regs<T, R>(v[g], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int j = (int)threadIdx.x + g * NT;
    if (j < groups) {
      const int k = j & (Ns - 1);
      const int o = (j - k) * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) A[lp(o + t * Ns)] = v[g][t];
    }
  }
  __syncthreads();
}

// Complex FFT of the LDS line A (length L, power of two, 1 <= L <= LMAX <=
// 8192, element i at A[lp(i)]) in place; the caller has synchronised after
// filling A.  LMAX fixes
// how many groups a thread holds per pass (1 for LMAX = 4096 and NT = 256).
template <class T, int LMAX, int NT = kBlock>
__device__ void fft_lds(Cx<T>* A, int L, const Cx<T>* __restrict__ tw, T sign) {
  int p = 0;
  while ((1 << p) < L) ++p;
  int Ns = 1;
  switch (p & 3) {
    case 1:
      stockham_pass<T, 2, pass_groups(2, LMAX, NT), NT>(A, L, Ns, tw, sign);
      Ns = 2;
      break;
    case 2:
      stockham_pass<T, 4, pass_groups(4, LMAX, NT), NT>(A, L, Ns, tw, sign);
      Ns = 4;
      break;
    case 3:
      stockham_pass<T, 8, pass_groups(8, LMAX, NT), NT>(A, L, Ns, tw, sign);
      Ns = 8;
      break;
    default:
      break;
  }
  for (int q = 0; q < (p >> 2); ++q) {
    stockham_pass<T, 16, pass_groups(16, LMAX, NT), NT>(A, L, Ns, tw, sign);
    Ns *= 16;
  }
}

}  // namespace drd
