// fft_lines.hpp -- real line transforms of the solve path in LDS.
//
// A complex radix-2 Stockham FFT of power-of-two length runs in LDS with one
// workgroup per line; the FFTW real-to-real kinds used by the solve path are
// built on it (unnormalised, FFTW conventions):
//   REDFT00 (DCT-I, n = N+1 points, FFT of 2N after even extension),
//   REDFT10 / REDFT01 (DCT-II / III, Makhoul reordering, FFT of N),
//   RODFT10 / RODFT01 (DST-II / III, via the DCT-II/III with sign flips),
//   RODFT00 (DST-I, FFT of 2(N+1) after odd extension).
#pragma once

#include "drd_common.hpp"

namespace drd {

template <class T>
struct Cx {
  T x, y;
};

template <class T>
__device__ __forceinline__ Cx<T> cmulT(Cx<T> a, Cx<T> b) {
  return Cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// radix-2 Stockham, twiddles tw[t] = exp(-2 pi i t / L) for t < L/2
template <class T>
__device__ Cx<T>* fftT(Cx<T>* a, Cx<T>* b, int L, const Cx<T>* __restrict__ tw, int sign) {
  const int half = L >> 1;
  for (int ns = 1; ns < L; ns <<= 1) {
    const int stride = L / (2 * ns);
    for (int j = threadIdx.x; j < half; j += blockDim.x) {
      const int k = j & (ns - 1);
      Cx<T> w = tw[k * stride];
      if (sign > 0) w.y = -w.y;
      const Cx<T> v0 = a[j];
      const Cx<T> v1 = cmulT(a[j + half], w);
      const int o = ((j - k) << 1) + k;
      b[o] = Cx<T>{v0.x + v1.x, v0.y + v1.y};
      b[o + ns] = Cx<T>{v0.x - v1.x, v0.y - v1.y};
    }
    __syncthreads();
    Cx<T>* t = a;
    a = b;
    b = t;
  }
  return a;
}

enum { TR_DCT1 = 0, TR_DCT2, TR_DCT3, TR_DST2, TR_DST3, TR_DST1 };

// Transforms the real line held in `line` (LDS, length n) in place.
// Unnormalised FFTW conventions: REDFT00 (n = N+1), REDFT10/01, RODFT10/01,
// RODFT00.  A and B are complex scratch of length >= FFT length.
template <class T>
__device__ void line_transform(int kind, T* line, int n, Cx<T>* A, Cx<T>* B, const Cx<T>* __restrict__ tw,
                               const Cx<T>* __restrict__ rot) {
  switch (kind) {
    case TR_DCT1: {  // even extension, FFT of length 2(n-1): Y_k = Re V_k
      const int N = n - 1, L = 2 * N;
      for (int m = threadIdx.x; m < L; m += blockDim.x) A[m] = Cx<T>{line[m <= N ? m : L - m], T(0)};
      __syncthreads();
      const Cx<T>* V = fftT(A, B, L, tw, -1);
      for (int k = threadIdx.x; k < n; k += blockDim.x) line[k] = V[k].x;
      __syncthreads();
      break;
    }
    case TR_DCT2:
    case TR_DST2: {  // Makhoul; DST-II(x)_k = DCT-II((-1)^j x_j)_{n-1-k}
      const bool sn = (kind == TR_DST2);
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int m = (j & 1) ? (n - 1 - (j >> 1)) : (j >> 1);
        const T v = line[j];
        A[m] = Cx<T>{(sn && (j & 1)) ? -v : v, T(0)};
      }
      __syncthreads();
      const Cx<T>* V = fftT(A, B, n, tw, -1);
      for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const Cx<T> r = rot[k], v = V[k];
        const T y = T(2) * (r.x * v.x - r.y * v.y);
        line[sn ? (n - 1 - k) : k] = y;
      }
      __syncthreads();
      break;
    }
    case TR_DCT3:
    case TR_DST3: {  // inverse pair of the above; DST-III(y) = (-1)^j DCT-III(reverse y)_j
      const bool sn = (kind == TR_DST3);
      for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const T a = sn ? line[n - 1 - k] : line[k];
        const T b = (k == 0) ? T(0) : -(sn ? line[k - 1] : line[n - k]);
        const Cx<T> r = rot[k];
        A[k] = Cx<T>{r.x * a + r.y * b, -r.y * a + r.x * b};
      }
      __syncthreads();
      const Cx<T>* v = fftT(A, B, n, tw, +1);
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int m = (j & 1) ? (n - 1 - (j >> 1)) : (j >> 1);
        const T x = v[m].x;
        line[j] = (sn && (j & 1)) ? -x : x;
      }
      __syncthreads();
      break;
    }
    case TR_DST1: {  // odd extension, FFT of length 2(n+1): Y_k = -Im V_{k+1}
      const int L = 2 * (n + 1);
      for (int m = threadIdx.x; m < L; m += blockDim.x) {
        T v = T(0);
        if (m >= 1 && m <= n) v = line[m - 1];
        if (m >= n + 2) v = -line[L - m - 1];
        A[m] = Cx<T>{v, T(0)};
      }
      __syncthreads();
      const Cx<T>* V = fftT(A, B, L, tw, -1);
      for (int k = threadIdx.x; k < n; k += blockDim.x) line[k] = -V[k + 1].y;
      __syncthreads();
      break;
    }
  }
}

}  // namespace drd
