// kernels_unet.hip -- forward, backward and Adam of the learned correction
// (layer table in unet.hpp) with hand-written kernels.
//
// 3x3 convolutions run as implicit GEMMs on the gfx950 MFMA units
// (v_mfma_f32_16x16x32_bf16): a wave computes 16 consecutive pixels of a row
// times all output channels; the GEMM depth k = tap*cin + ci (tap = 3*(dy+1)
// + (dx+1)) is padded to a multiple of 32, so the 8 k of a lane are 8
// consecutive channels of one tap: one 16-byte NHWC load.  Lane layout of
// the 16x16x32 MFMA: lane l holds A[l%16][8(l/16)+j], B[8(l/16)+j][l%16],
// and D[4(l/16)+i][l%16].
//
// The data gradient of a convolution is a convolution of the masked output
// gradient with the flipped, transposed weights (same kernel); the weight
// gradient is an MFMA GEMM over pixel chunks with FP32 atomics into the
// gradient vector (training is not required to be bit-reproducible).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "drd_common.hpp"
#include "unet.hpp"

namespace drd {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ bf16x8 zero8() {
  const uint4 z = {0u, 0u, 0u, 0u};
  return __builtin_bit_cast(bf16x8, z);
}
__device__ __forceinline__ bf16x8 load16B(const bf16* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// ---- 3x3 convolution ---------------------------------------------------------
struct ConvIO {
  const bf16* x1;  // first input: (H >> up1) x (W >> up1) x c1
  int c1, up1;
  const bf16* x2;  // second input: H x W x c2 (channels c1 .. c1+c2-1)
  int c2;
  const bf16* w;  // [cout][kp]
  int kp;
  const float* bias;  // or null
  int H, W, cout, relu;
  bf16* y;  // H x W x cout
};

__device__ __forceinline__ bf16x8 conv_load8(const ConvIO& a, int yy, int xx, int ci0) {
  if (yy < 0 || yy >= a.H || xx < 0 || xx >= a.W) return zero8();
  if (ci0 < a.c1) {
    const int w1 = a.W >> a.up1;
    return load16B(a.x1 + ((size_t)(yy >> a.up1) * w1 + (xx >> a.up1)) * a.c1 + ci0);
  }
  return load16B(a.x2 + ((size_t)yy * a.W + xx) * a.c2 + (ci0 - a.c1));
}

__device__ __forceinline__ float conv_load1(const ConvIO& a, int yy, int xx, int ci) {
  if (yy < 0 || yy >= a.H || xx < 0 || xx >= a.W) return 0.f;
  if (ci < a.c1) {
    const int w1 = a.W >> a.up1;
    return (float)a.x1[((size_t)(yy >> a.up1) * w1 + (xx >> a.up1)) * a.c1 + ci];
  }
  return (float)a.x2[((size_t)yy * a.W + xx) * a.c2 + (ci - a.c1)];
}

// grid (ceil(W / 64), H), 256 threads: wave w of block (bx, y) covers pixels
// x0 = (4 bx + w) * 16 .. x0 + 15 of row y.  NB = cout / 16.
template <int NB>
__global__ __launch_bounds__(256) void k_conv3x3(ConvIO a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int y = blockIdx.y;
  const int x0 = (blockIdx.x * 4 + wave) * 16;
  if (x0 >= a.W) return;  // whole wave
  const int cin = a.c1 + a.c2, k9 = 9 * cin;
  const int m = lane & 15, g = lane >> 4;
  f32x4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < a.kp; kc += 32) {
    const int kg = kc + 8 * g;
    bf16x8 av = zero8();
    if (kg < k9) {
      const int tap = kg / cin, ci0 = kg - tap * cin;
      av = conv_load8(a, y + tap / 3 - 1, x0 + m + tap % 3 - 1, ci0);
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const bf16x8 bv = load16B(a.w + (size_t)(nb * 16 + m) * a.kp + kg);
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[nb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = nb * 16 + m;
    const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int px = x0 + 4 * g + i;
      if (px < a.W) {
        float v = acc[nb][i] + b;
        if (a.relu) v = v > 0.f ? v : 0.f;
        a.y[((size_t)y * a.W + px) * a.cout + co] = (bf16)v;
      }
    }
  }
}

static void launch_conv(hipStream_t st, const ConvIO& a) {
  dim3 grid((unsigned)((a.W + 63) / 64), (unsigned)a.H);
  switch (a.cout / 16) {
    case 1: hipLaunchKernelGGL(k_conv3x3<1>, grid, dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_conv3x3<2>, grid, dim3(256), 0, st, a); break;
    case 3: hipLaunchKernelGGL(k_conv3x3<3>, grid, dim3(256), 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_conv3x3<4>, grid, dim3(256), 0, st, a); break;
  }
}

// ---- weight gradient: gW[co][k] += sum_p dz[p][co] * in_k(p) -------------------
// One wave per (16 output channels, 16 k, chunk of pixels); the GEMM reduction
// runs over pixels, 32 per MFMA.  Results added to the FP32 gradient vector
// (row length k9; padded k are skipped).
constexpr int kDwChunk = 2048;  // pixels per wave

__global__ __launch_bounds__(64) void k_conv_dw(ConvIO a, const bf16* __restrict__ dz, float* __restrict__ gw) {
  const int lane = threadIdx.x, m = lane & 15, g = lane >> 4;
  const int kb = blockIdx.x, mb = blockIdx.y;
  const int npix = a.H * a.W;
  const int p_begin = blockIdx.z * kDwChunk;
  const int cin = a.c1 + a.c2, k9 = 9 * cin;
  const int k = kb * 16 + m;  // this lane's B column
  const bool kval = k < k9;
  const int tap = kval ? k / cin : 0, ci = kval ? k - tap * cin : 0;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int p0 = p_begin; p0 < p_begin + kDwChunk && p0 < npix; p0 += 32) {
    bf16x8 av, bv;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = p0 + 8 * g + j;
      float va = 0.f, vb = 0.f;
      if (p < npix) {
        va = (float)dz[(size_t)p * a.cout + mb * 16 + m];
        if (kval) {
          const int yy = p / a.W, xx = p - yy * a.W;
          vb = conv_load1(a, yy + dy, xx + dx, ci);
        }
      }
      av[j] = (bf16)va;
      bv[j] = (bf16)vb;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  }
  // D[4g+i][m]: output channel mb*16 + 4g + i, k = kb*16 + m
  if (kval) {
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&gw[(size_t)(mb * 16 + 4 * g + i) * k9 + k], acc[i]);
  }
}

// gb[co] += sum_p dz[p][co]
__global__ __launch_bounds__(256) void k_bias_grad(const bf16* __restrict__ dz, int npix, int cout,
                                                   float* __restrict__ gb) {
  __shared__ float red[256];
  const int co = blockIdx.y;
  float s = 0.f;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < npix; p += gridDim.x * 256) s += (float)dz[(size_t)p * cout + co];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&gb[co], red[0]);
}

// ---- elementwise pieces --------------------------------------------------------
// dz = bf16(g * (y > 0))  (ReLU derivative from the stored output)
__global__ __launch_bounds__(256) void k_relu_mask(int64_t n, const float* __restrict__ g, const bf16* __restrict__ y,
                                                   bf16* __restrict__ dz) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dz[i] = (bf16)((float)y[i] > 0.f ? g[i] : 0.f);
}

// out = 2x2 average of in (H x W x c -> H/2 x W/2 x c)
__global__ __launch_bounds__(256) void k_pool2(int H, int W, int c, const bf16* __restrict__ in, bf16* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int h = H / 2, w = W / 2;
  if (i >= (int64_t)h * w * c) return;
  const int ch = (int)(i % c);
  const int64_t px = i / c;
  const int yy = (int)(px / w), xx = (int)(px - (int64_t)yy * w);
  const size_t b = ((size_t)(2 * yy) * W + 2 * xx) * c + ch;
  const float s = (float)in[b] + (float)in[b + c] + (float)in[b + (size_t)W * c] + (float)in[b + (size_t)W * c + c];
  out[i] = (bf16)(0.25f * s);
}

// gradient routing into an FP32 accumulator gd (Hd x Wd x cd) from a bf16
// data gradient src (channels [coff, coff+cd) of a cs-channel tensor):
//   mode 0: same size; mode 1: src is 2x larger, sum of the 2x2 children
//   (nearest upsampling); mode 2: src is 2x smaller, 1/4 of the parent (pooling)
__global__ __launch_bounds__(256) void k_route(int mode, int Hd, int Wd, int cd, float* __restrict__ gd,
                                               const bf16* __restrict__ src, int cs, int coff) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)Hd * Wd * cd) return;
  const int ch = (int)(i % cd);
  const int64_t px = i / cd;
  const int yy = (int)(px / Wd), xx = (int)(px - (int64_t)yy * Wd);
  float v;
  if (mode == 0) {
    v = (float)src[((size_t)yy * Wd + xx) * cs + coff + ch];
  } else if (mode == 1) {
    const int Ws = 2 * Wd;
    const size_t b = ((size_t)(2 * yy) * Ws + 2 * xx) * cs + coff + ch;
    v = (float)src[b] + (float)src[b + cs] + (float)src[b + (size_t)Ws * cs] + (float)src[b + (size_t)Ws * cs + cs];
  } else {
    const int Ws = Wd / 2;
    v = 0.25f * (float)src[((size_t)(yy / 2) * Ws + xx / 2) * cs + coff + ch];
  }
  gd[i] += v;
}

// ---- 1x1 output layer (FP32 weights) -------------------------------------------
__global__ __launch_bounds__(256) void k_out_fwd(int npix, const bf16* __restrict__ x, const float* __restrict__ wo,
                                                 const float* __restrict__ bo, float* __restrict__ out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  float xv[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) xv[c] = (float)x[(size_t)p * 16 + c];
#pragma unroll
  for (int o = 0; o < kUnetOut; ++o) {
    float s = bo[o];
#pragma unroll
    for (int c = 0; c < 16; ++c) s += wo[o * 16 + c] * xv[c];
    out[(size_t)p * kUnetOut + o] = s;
  }
}

// loss = 0.5 sum (out - t)^2 / n,  go = (out - t) / n,  then the 1x1 layer's
// gradients: gd1[p][c] += sum_o wo[o][c] go[p][o]; gwo, gbo by block sums
__global__ __launch_bounds__(256) void k_out_bwd(int npix, const bf16* __restrict__ x, const float* __restrict__ wo,
                                                 const float* __restrict__ out, const float* __restrict__ t,
                                                 float* __restrict__ gd1, float* __restrict__ gwo,
                                                 float* __restrict__ gbo, double* __restrict__ loss) {
  __shared__ float red[256];
  const int p = blockIdx.x * 256 + threadIdx.x;
  const float inv = 1.f / (float)(npix * kUnetOut);
  float go[kUnetOut] = {0.f, 0.f}, xv[16], l = 0.f;
  if (p < npix) {
#pragma unroll
    for (int o = 0; o < kUnetOut; ++o) {
      const float d = out[(size_t)p * kUnetOut + o] - t[(size_t)p * kUnetOut + o];
      l += 0.5f * d * d * inv;
      go[o] = d * inv;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      xv[c] = (float)x[(size_t)p * 16 + c];
      float s = 0.f;
#pragma unroll
      for (int o = 0; o < kUnetOut; ++o) s += wo[o * 16 + c] * go[o];
      gd1[(size_t)p * 16 + c] += s;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 16; ++c) xv[c] = 0.f;
  }
  // block sums of go[o] * x[c] (32 values), go[o] (2) and the loss (1)
  for (int q = 0; q < kUnetOut * 16 + kUnetOut + 1; ++q) {
    float v;
    if (q < kUnetOut * 16)
      v = go[q / 16] * xv[q % 16];
    else if (q < kUnetOut * 16 + kUnetOut)
      v = go[q - kUnetOut * 16];
    else
      v = l;
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (q < kUnetOut * 16)
        atomicAdd(&gwo[q], red[0]);
      else if (q < kUnetOut * 16 + kUnetOut)
        atomicAdd(&gbo[q - kUnetOut * 16], red[0]);
      else
        atomicAdd(loss, (double)red[0]);
    }
    __syncthreads();
  }
}

// ---- parameters ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 forward weights wf[co][kp] (zero padded) and backward weights
// wb[ci][kpb], kpb = conv_kp(cout), wb[ci][tap' cout + co] = W[co][(8-tap') cin + ci]
__global__ __launch_bounds__(256) void k_pack_w(const float* __restrict__ w, int cout, int cin, bf16* __restrict__ wf,
                                                int kp, bf16* __restrict__ wb, int kpb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int k9 = 9 * cin;
  if (i < cout * kp) {
    const int co = i / kp, k = i - co * kp;
    wf[i] = (bf16)(k < k9 ? w[(size_t)co * k9 + k] : 0.f);
  }
  if (wb && i < cin * kpb) {
    const int ci = i / kpb, kk = i - ci * kpb;
    float v = 0.f;
    if (kk < 9 * cout) {
      const int tp = kk / cout, co = kk - tp * cout;
      v = w[(size_t)co * k9 + (8 - tp) * cin + ci];
    }
    wb[i] = (bf16)v;
  }
}

// ---- the network object ------------------------------------------------------------
struct UNet {
  int H = 0, W = 0;
  int np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  size_t offW[kUnetConvs], offB[kUnetConvs], offWo = 0, offBo = 0;
  bf16 *wf[kUnetConvs] = {}, *wb[kUnetConvs] = {};
  bf16* act[kUnetConvs] = {};  // outputs of the 3x3 layers
  bf16 *p1 = nullptr, *p2 = nullptr;
  float* out = nullptr;
  float* gact[kUnetConvs] = {};  // FP32 gradients of the layer outputs
  bf16 *dz = nullptr, *dx = nullptr;
  double* loss = nullptr;
  int step = 0;
  std::vector<void*> allocs;
};

static int hl(int H, int lev) { return H >> lev; }

static void* ualloc(UNet* u, size_t bytes, int* err) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    *err = DRD_E_HIP;
    return nullptr;
  }
  (void)hipMemset(p, 0, bytes);
  u->allocs.push_back(p);
  return p;
}

static void pack_weights(hipStream_t st, UNet* u) {
  for (int l = 0; l < kUnetConvs; ++l) {
    const ConvSpec& s = kUnetSpec[l];
    const int cin = conv_cin(s), kp = conv_kp(cin), kpb = conv_kp(s.cout);
    const int n = std::max(s.cout * kp, cin * kpb);
    hipLaunchKernelGGL(k_pack_w, dim3((n + 255) / 256), dim3(256), 0, st, u->p + u->offW[l], s.cout, cin, u->wf[l],
                       kp, l > 0 ? u->wb[l] : (bf16*)nullptr, kpb);
  }
}

UNet* unet_create(int H, int W, uint64_t seed, int* err) {
  *err = 0;
  if (H < 4 || W < 4 || H % 4 || W % 4) {
    *err = DRD_E_ARG;
    return nullptr;
  }
  UNet* u = new UNet();
  u->H = H;
  u->W = W;
  size_t off = 0;
  for (int l = 0; l < kUnetConvs; ++l) {
    const ConvSpec& s = kUnetSpec[l];
    u->offW[l] = off;
    off += (size_t)s.cout * 9 * conv_cin(s);
    u->offB[l] = off;
    off += s.cout;
  }
  u->offWo = off;
  off += kUnetOut * 16;
  u->offBo = off;
  off += kUnetOut;
  u->np = (int)off;
  // He-normal initial weights (SplitMix64 + Box-Muller), zero biases; the 1x1
  // layer starts small so the initial correction is close to zero
  std::vector<float> h(off, 0.f);
  uint64_t st = seed;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto gauss = [&]() { return std::sqrt(-2.0 * std::log(next())) * std::cos(2.0 * M_PI * next()); };
  for (int l = 0; l < kUnetConvs; ++l) {
    const ConvSpec& s = kUnetSpec[l];
    const int k9 = 9 * conv_cin(s);
    const double sd = std::sqrt(2.0 / k9);
    for (int i = 0; i < s.cout * k9; ++i) h[u->offW[l] + i] = (float)(sd * gauss());
  }
  for (int i = 0; i < kUnetOut * 16; ++i) h[u->offWo + i] = (float)(1e-3 * gauss());
  int e = 0;
  u->p = (float*)ualloc(u, off * 4, &e);
  u->g = (float*)ualloc(u, off * 4, &e);
  u->m = (float*)ualloc(u, off * 4, &e);
  u->v = (float*)ualloc(u, off * 4, &e);
  for (int l = 0; !e && l < kUnetConvs; ++l) {
    const ConvSpec& s = kUnetSpec[l];
    const int cin = conv_cin(s);
    const size_t px = (size_t)hl(H, s.level) * hl(W, s.level);
    u->wf[l] = (bf16*)ualloc(u, (size_t)s.cout * conv_kp(cin) * 2, &e);
    if (l > 0) u->wb[l] = (bf16*)ualloc(u, (size_t)cin * conv_kp(s.cout) * 2, &e);
    u->act[l] = (bf16*)ualloc(u, px * s.cout * 2, &e);
    u->gact[l] = (float*)ualloc(u, px * s.cout * 4, &e);
  }
  const size_t full = (size_t)H * W;
  if (!e) u->p1 = (bf16*)ualloc(u, full / 4 * 16 * 2, &e);
  if (!e) u->p2 = (bf16*)ualloc(u, full / 16 * 32 * 2, &e);
  if (!e) u->out = (float*)ualloc(u, full * kUnetOut * 4, &e);
  if (!e) u->dz = (bf16*)ualloc(u, full * 32 * 2, &e);
  if (!e) u->dx = (bf16*)ualloc(u, full * 64 * 2, &e);
  if (!e) u->loss = (double*)ualloc(u, 8, &e);
  if (!e && hipMemcpy(u->p, h.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess) e = DRD_E_HIP;
  if (e) {
    unet_destroy(u);
    *err = e;
    return nullptr;
  }
  pack_weights(0, u);
  if (hipDeviceSynchronize() != hipSuccess) {
    unet_destroy(u);
    *err = DRD_E_HIP;
    return nullptr;
  }
  return u;
}

void unet_destroy(UNet* u) {
  if (!u) return;
  for (void* p : u->allocs) (void)hipFree(p);
  delete u;
}

int unet_set_params(UNet* u, const float* host, int n) {
  if (n != u->np) return DRD_E_ARG;
  if (hipMemcpy(u->p, host, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess) return DRD_E_HIP;
  pack_weights(0, u);
  return hipDeviceSynchronize() == hipSuccess ? 0 : DRD_E_HIP;
}

int unet_get_params(UNet* u, float* host, int n) {
  if (n != u->np) return DRD_E_ARG;
  return hipMemcpy(host, u->p, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : DRD_E_HIP;
}

int unet_get_grads(UNet* u, float* host, int n) {
  if (n != u->np) return DRD_E_ARG;
  return hipMemcpy(host, u->g, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : DRD_E_HIP;
}

// the ConvIO of layer l reading the forward tensors (input image `in` for l = 0)
static ConvIO conv_io(UNet* u, int l, const bf16* in) {
  const ConvSpec& s = kUnetSpec[l];
  ConvIO a{};
  a.H = hl(u->H, s.level);
  a.W = hl(u->W, s.level);
  a.c1 = s.c1;
  a.up1 = s.up1;
  a.c2 = s.c2;
  a.cout = s.cout;
  a.kp = conv_kp(conv_cin(s));
  a.w = u->wf[l];
  a.bias = u->p + u->offB[l];
  a.relu = 1;
  a.y = u->act[l];
  switch (l) {
    case 0: a.x1 = in; break;
    case 2: a.x1 = u->p1; break;
    case 4: a.x1 = u->p2; break;
    case 6: a.x1 = u->act[5]; a.x2 = u->act[3]; break;  // up(e3) || e2
    case 8: a.x1 = u->act[7]; a.x2 = u->act[1]; break;  // up(d2) || e1
    default: a.x1 = u->act[l - 1]; break;
  }
  return a;
}

static unsigned nblk256(int64_t n) { return (unsigned)((n + 255) / 256); }

int unet_forward(hipStream_t st, UNet* u, const void* in, float* out) {
  const bf16* x = reinterpret_cast<const bf16*>(in);
  for (int l = 0; l < kUnetConvs; ++l) {
    if (l == 2)
      hipLaunchKernelGGL(k_pool2, dim3(nblk256((int64_t)u->H * u->W / 4 * 16)), dim3(256), 0, st, u->H, u->W, 16,
                         (const bf16*)u->act[1], u->p1);
    if (l == 4)
      hipLaunchKernelGGL(k_pool2, dim3(nblk256((int64_t)u->H * u->W / 16 * 32)), dim3(256), 0, st, u->H / 2,
                         u->W / 2, 32, (const bf16*)u->act[3], u->p2);
    launch_conv(st, conv_io(u, l, x));
  }
  const int npix = u->H * u->W;
  hipLaunchKernelGGL(k_out_fwd, dim3(nblk256(npix)), dim3(256), 0, st, npix, (const bf16*)u->act[9],
                     (const float*)(u->p + u->offWo), (const float*)(u->p + u->offBo), out ? out : u->out);
  return hipGetLastError() == hipSuccess ? 0 : DRD_E_HIP;
}

int unet_zero_grad(hipStream_t st, UNet* u) {
  if (hipMemsetAsync(u->loss, 0, 8, st) != hipSuccess) return DRD_E_HIP;
  return hipMemsetAsync(u->g, 0, (size_t)u->np * 4, st) == hipSuccess ? 0 : DRD_E_HIP;
}

int unet_read_loss(hipStream_t st, UNet* u, double* loss) {
  if (hipMemcpyAsync(loss, u->loss, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return DRD_E_HIP;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : DRD_E_HIP;
}

int unet_backward(hipStream_t st, UNet* u, const void* in, const float* target) {
  const bf16* x = reinterpret_cast<const bf16*>(in);
  const int npix = u->H * u->W;
  for (int l = 0; l < kUnetConvs; ++l) {
    const ConvSpec& s = kUnetSpec[l];
    const size_t n = (size_t)hl(u->H, s.level) * hl(u->W, s.level) * s.cout;
    if (hipMemsetAsync(u->gact[l], 0, n * 4, st) != hipSuccess) return DRD_E_HIP;
  }
  hipLaunchKernelGGL(k_out_bwd, dim3(nblk256(npix)), dim3(256), 0, st, npix, (const bf16*)u->act[9],
                     (const float*)(u->p + u->offWo), (const float*)u->out, target, u->gact[9], u->g + u->offWo,
                     u->g + u->offBo, u->loss);
  for (int l = kUnetConvs - 1; l >= 0; --l) {
    const ConvSpec& s = kUnetSpec[l];
    const int Hl = hl(u->H, s.level), Wl = hl(u->W, s.level), cin = conv_cin(s);
    const int64_t n = (int64_t)Hl * Wl * s.cout;
    hipLaunchKernelGGL(k_relu_mask, dim3(nblk256(n)), dim3(256), 0, st, n, (const float*)u->gact[l],
                       (const bf16*)u->act[l], u->dz);
    ConvIO a = conv_io(u, l, x);
    const unsigned chunks = (unsigned)((Hl * Wl + kDwChunk - 1) / kDwChunk);
    hipLaunchKernelGGL(k_conv_dw, dim3((unsigned)((9 * cin + 15) / 16), (unsigned)(s.cout / 16), chunks), dim3(64), 0,
                       st, a, (const bf16*)u->dz, u->g + u->offW[l]);
    hipLaunchKernelGGL(k_bias_grad, dim3(std::min<unsigned>(nblk256(Hl * Wl), 64u), (unsigned)s.cout), dim3(256), 0,
                       st, (const bf16*)u->dz, Hl * Wl, s.cout, u->g + u->offB[l]);
    if (l == 0) break;
    // data gradient: conv of dz with the flipped, transposed weights
    ConvIO b{};
    b.H = Hl;
    b.W = Wl;
    b.x1 = u->dz;
    b.c1 = s.cout;
    b.up1 = 0;
    b.c2 = 0;
    b.w = u->wb[l];
    b.kp = conv_kp(s.cout);
    b.bias = nullptr;
    b.cout = cin;
    b.relu = 0;
    b.y = u->dx;
    launch_conv(st, b);
    // route dx (Hl x Wl x cin) to the gradients of the layer's inputs
    auto route = [&](int mode, int Hd, int Wd, int cd, float* gd, int coff) {
      hipLaunchKernelGGL(k_route, dim3(nblk256((int64_t)Hd * Wd * cd)), dim3(256), 0, st, mode, Hd, Wd, cd, gd,
                         (const bf16*)u->dx, cin, coff);
    };
    switch (l) {
      case 2: route(2, u->H, u->W, 16, u->gact[1], 0); break;          // pool(e1)
      case 4: route(2, u->H / 2, u->W / 2, 32, u->gact[3], 0); break;  // pool(e2)
      case 6:
        route(1, u->H / 4, u->W / 4, 32, u->gact[5], 0);   // up(e3)
        route(0, u->H / 2, u->W / 2, 32, u->gact[3], 32);  // e2
        break;
      case 8:
        route(1, u->H / 2, u->W / 2, 32, u->gact[7], 0);  // up(d2)
        route(0, u->H, u->W, 16, u->gact[1], 32);         // e1
        break;
      default: route(0, Hl, Wl, cin, u->gact[l - 1], 0); break;
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : DRD_E_HIP;
}

int unet_adam(hipStream_t st, UNet* u, float lr, float b1, float b2, float eps) {
  u->step += 1;
  const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
  hipLaunchKernelGGL(k_adam, dim3(nblk256(u->np)), dim3(256), 0, st, u->np, u->p, (const float*)u->g, u->m, u->v, lr,
                     b1, b2, eps, c1, c2);
  pack_weights(st, u);
  return hipGetLastError() == hipSuccess ? 0 : DRD_E_HIP;
}

// ---- velocity vectors <-> network images ---------------------------------------
// Pixel (k, i) of the ny x nx image carries the u face i of row k (i < nx)
// and the v face k of column i (k >= 1; 0 on k = 0).  The u face i = nx is
// left to the DCT solve.
constexpr int kSsBlocks = 256;

// part[v * kSsBlocks + b] = sum over block b's grid-stride share of x_v^2
__global__ __launch_bounds__(256) void k_sumsq3(int64_t n, const double* __restrict__ x0,
                                                const double* __restrict__ x1, const double* __restrict__ x2,
                                                double* __restrict__ part) {
  __shared__ double red[3][256];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kSsBlocks * 256) {
    s0 += x0[i] * x0[i];
    s1 += x1[i] * x1[i];
    s2 += x2[i] * x2[i];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int v = 0; v < 3; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[threadIdx.x * kSsBlocks + blockIdx.x] = red[threadIdx.x][0];
}

// ss[v] = sum of part[v * kSsBlocks ...] (fixed order: deterministic)
__global__ __launch_bounds__(256) void k_sumsq3_fin(const double* __restrict__ part, double* __restrict__ ss) {
  __shared__ double red[256];
  for (int v = 0; v < 3; ++v) {
    red[threadIdx.x] = part[v * kSsBlocks + threadIdx.x];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) ss[v] = red[0];
    __syncthreads();
  }
}

__device__ __forceinline__ double rms_scale(double ss, int64_t n) {
  const double s = sqrt(ss / (double)n);
  return s > 0.0 ? s : 1.0;
}

// img = [r_u/s_r, r_v/s_r, d_u/s_d, d_v/s_d, z_u/s_z, z_v/s_z, log eta, chi]
// (bf16 NHWC) with ss = (|r|^2, |d|^2, |z|^2)
__global__ __launch_bounds__(256) void k_pack_features(Grid g, const double* __restrict__ r,
                                                       const double* __restrict__ d, const double* __restrict__ z,
                                                       const double* __restrict__ eta,
                                                       const double* __restrict__ chi, const double* __restrict__ ss,
                                                       bf16* __restrict__ img) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (p >= (int64_t)nx * ny) return;
  const int k = (int)(p / nx), i = (int)(p - (int64_t)k * nx);
  const int64_t n = g.nuv();
  const double ir = 1.0 / rms_scale(ss[0], n), id = 1.0 / rms_scale(ss[1], n), iz = 1.0 / rms_scale(ss[2], n);
  const int64_t qu = i + (int64_t)(nx + 1) * k, qv = g.mn() + i + (int64_t)nx * (k - 1);
  const bool hv = k >= 1;
  float v[8];
  v[0] = (float)(r[qu] * ir);
  v[1] = hv ? (float)(r[qv] * ir) : 0.f;
  v[2] = (float)(d[qu] * id);
  v[3] = hv ? (float)(d[qv] * id) : 0.f;
  v[4] = (float)(z[qu] * iz);
  v[5] = hv ? (float)(z[qv] * iz) : 0.f;
  v[6] = (float)log(eta[g.ci(i, k)]);
  v[7] = (float)chi[g.ci(i, k)];
  bf16x8 o;
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = (bf16)v[c];
  *reinterpret_cast<uint4*>(img + p * 8) = __builtin_bit_cast(uint4, o);
}

// w = s_d * out as a velocity Krylov vector (the u faces i = nx get 0)
__global__ __launch_bounds__(256) void k_unpack_correction(Grid g, const float* __restrict__ out,
                                                           const double* __restrict__ ss, double* __restrict__ w) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (p >= (int64_t)nx * ny) return;
  const int k = (int)(p / nx), i = (int)(p - (int64_t)k * nx);
  const double sd = rms_scale(ss[1], g.nuv());
  const int64_t qu = i + (int64_t)(nx + 1) * k;
  w[qu] = sd * (double)out[p * 2];
  if (i == nx - 1) w[qu + 1] = 0.0;
  if (k >= 1) w[g.mn() + i + (int64_t)nx * (k - 1)] = sd * (double)out[p * 2 + 1];
}

// training target v / s_d on the image
__global__ __launch_bounds__(256) void k_make_target(Grid g, const double* __restrict__ v,
                                                     const double* __restrict__ ss, float* __restrict__ t) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (p >= (int64_t)nx * ny) return;
  const int k = (int)(p / nx), i = (int)(p - (int64_t)k * nx);
  const double id = 1.0 / rms_scale(ss[1], g.nuv());
  const int64_t qu = i + (int64_t)(nx + 1) * k, qv = g.mn() + i + (int64_t)nx * (k - 1);
  t[p * 2] = (float)(v[qu] * id);
  t[p * 2 + 1] = k >= 1 ? (float)(v[qv] * id) : 0.f;
}

void launch_norms3(hipStream_t st, int64_t n, const double* r, const double* d, const double* z, double* part,
                   double* ss) {
  hipLaunchKernelGGL(k_sumsq3, dim3(kSsBlocks), dim3(256), 0, st, n, r, d, z, part);
  hipLaunchKernelGGL(k_sumsq3_fin, dim3(1), dim3(256), 0, st, (const double*)part, ss);
}
void launch_pack_features(hipStream_t st, const Grid& g, const double* r, const double* d, const double* z,
                          const double* eta, const double* chi, const double* ss, void* img) {
  hipLaunchKernelGGL(k_pack_features, dim3(nblk256(g.ncell())), dim3(256), 0, st, g, r, d, z, eta, chi, ss,
                     reinterpret_cast<bf16*>(img));
}
// z += alpha c, alpha = dots[0] / dots[1] (0 when dots[1] is not positive)
__global__ __launch_bounds__(256) void k_axpy_ratio(int64_t n, const double* __restrict__ dots,
                                                    const double* __restrict__ c, double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double den = dots[1];
  const double a = den > 0.0 ? dots[0] / den : 0.0;
  z[i] += a * c[i];
}

void launch_axpy_ratio(hipStream_t st, int64_t n, const double* dots, const double* c, double* z) {
  hipLaunchKernelGGL(k_axpy_ratio, dim3(nblk256(n)), dim3(256), 0, st, n, dots, c, z);
}

void launch_unpack_correction(hipStream_t st, const Grid& g, const float* out, const double* ss, double* w) {
  hipLaunchKernelGGL(k_unpack_correction, dim3(nblk256(g.ncell())), dim3(256), 0, st, g, out, ss, w);
}
void launch_make_target(hipStream_t st, const Grid& g, const double* v, const double* ss, float* t) {
  hipLaunchKernelGGL(k_make_target, dim3(nblk256(g.ncell())), dim3(256), 0, st, g, v, ss, t);
}

// crop (y0, x0, h x w) of an H x W image with c channels of esz bytes each
__global__ __launch_bounds__(256) void k_crop(int W, int c, int esz, const unsigned char* __restrict__ src, int y0,
                                              int x0, int h, int w, unsigned char* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;  // one 4-byte word
  const int row_words = w * c * esz / 4;
  if (i >= (int64_t)h * row_words) return;
  const int yy = (int)(i / row_words), wd = (int)(i - (int64_t)yy * row_words);
  const size_t so = ((size_t)(y0 + yy) * W + x0) * c * esz;
  reinterpret_cast<uint32_t*>(dst + (size_t)yy * w * c * esz)[wd] = reinterpret_cast<const uint32_t*>(src + so)[wd];
}

void launch_crop(hipStream_t st, int W, int c, int esz, const void* src, int y0, int x0, int h, int w, void* dst) {
  hipLaunchKernelGGL(k_crop, dim3(nblk256((int64_t)h * w * c * esz / 4)), dim3(256), 0, st, W, c, esz,
                     (const unsigned char*)src, y0, x0, h, w, (unsigned char*)dst);
}

int unet_copy_params(hipStream_t st, UNet* dst, const UNet* src) {
  if (dst->np != src->np) return DRD_E_ARG;
  if (hipMemcpyAsync(dst->p, src->p, (size_t)src->np * 4, hipMemcpyDeviceToDevice, st) != hipSuccess) return DRD_E_HIP;
  pack_weights(st, dst);
  return hipGetLastError() == hipSuccess ? 0 : DRD_E_HIP;
}

int unet_size(const UNet* u, int* H, int* W) {
  *H = u->H;
  *W = u->W;
  return 0;
}

double unet_forward_flops(const UNet* u) {
  double f = 0;
  for (int l = 0; l < kUnetConvs; ++l) {
    const ConvSpec& s = kUnetSpec[l];
    f += 2.0 * hl(u->H, s.level) * hl(u->W, s.level) * 9.0 * conv_cin(s) * s.cout;
  }
  return f + 2.0 * u->H * u->W * 16.0 * kUnetOut;
}

}  // namespace drd
