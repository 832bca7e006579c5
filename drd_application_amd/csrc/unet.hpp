// unet.hpp -- the learned correction of the velocity preconditioner (no
// reference counterpart): a 3-level U-Net with 16/32/32 channels applied
// after the fast-Poisson (DCT) preconditioner,
//   z' = z + alpha s_d N(r / s_r, d / s_d, z / s_z, log eta, chi),
//   z = M r,  d = r - A z,
// with M the fast-Poisson (DCT) solve; N approximates the error the DCT
// solve leaves, (A^-1 r - z) / s_d.  The correction c = s_d N enters with
// the step alpha = (d, A c) / (A c, A c) that minimises |r - A (z + alpha c)|,
// so it never enlarges the residual of the plain DCT solve.  Images are ny x nx
// (u faces i < nx and v faces k >= 1 mapped to cell (i, k)).  BF16 weights
// and activations (NHWC), FP32 accumulation on the MFMA units, FP32 master
// weights and Adam moments.
//
// Layers (3x3 convolutions, zero padding, ReLU; "||" is channel concat,
// "up" nearest-neighbour upsampling by 2, "pool" 2x2 averaging):
//   e1a = conv(in[8] -> 16)         e1 = conv(e1a -> 16)          H x W
//   e2a = conv(pool(e1) -> 32)      e2 = conv(e2a -> 32)          H/2 x W/2
//   e3a = conv(pool(e2) -> 32)      e3 = conv(e3a -> 32)          H/4 x W/4
//   d2a = conv(up(e3) || e2 -> 32)  d2 = conv(d2a -> 32)          H/2 x W/2
//   d1a = conv(up(d2) || e1 -> 16)  d1 = conv(d1a -> 16)          H x W
//   out = conv1x1(d1 -> 2), no activation, FP32
#pragma once

#include <stdint.h>

#include <vector>

#include <hip/hip_runtime.h>

namespace drd {

constexpr int kUnetIn = 8;   // r_u, r_v, d_u, d_v, z_u, z_v, log eta, chi
constexpr int kUnetOut = 2;  // corrections of u and v
constexpr int kUnetConvs = 10;

struct ConvSpec {
  int c1, up1;  // channels read from the first input (upsampled by 2^up1)
  int c2;       // channels read from the second input (skip connection)
  int cout;
  int level;    // image size H >> level, W >> level
};

// the ten 3x3 convolutions in forward order
constexpr ConvSpec kUnetSpec[kUnetConvs] = {
    {kUnetIn, 0, 0, 16, 0}, {16, 0, 0, 16, 0},  // e1a, e1
    {16, 0, 0, 32, 1},      {32, 0, 0, 32, 1},  // e2a, e2 (input pooled)
    {32, 0, 0, 32, 2},      {32, 0, 0, 32, 2},  // e3a, e3 (input pooled)
    {32, 1, 32, 32, 1},     {32, 0, 0, 32, 1},  // d2a (up(e3) || e2), d2
    {32, 1, 16, 16, 0},     {16, 0, 0, 16, 0},  // d1a (up(d2) || e1), d1
};

__host__ __device__ constexpr int conv_cin(const ConvSpec& s) { return s.c1 + s.c2; }
// GEMM depth 9*cin rounded up to the MFMA step of 32
__host__ __device__ constexpr int conv_kp(int cin) { return (9 * cin + 31) / 32 * 32; }

// Flat FP32 parameter vector: for each conv l, W_l[cout][9*cin] (k = tap*cin
// + ci, tap = 3*(dy+1) + (dx+1)) then b_l[cout]; then the 1x1 layer
// Wo[2][16], bo[2].
inline int unet_param_count() {
  int n = 0;
  for (int l = 0; l < kUnetConvs; ++l) n += kUnetSpec[l].cout * 9 * conv_cin(kUnetSpec[l]) + kUnetSpec[l].cout;
  return n + kUnetOut * 16 + kUnetOut;
}

struct UNet;
UNet* unet_create(int H, int W, uint64_t seed, int* err);  // H, W multiples of 4
void unet_destroy(UNet* u);
int unet_set_params(UNet* u, const float* host, int n);
int unet_get_params(UNet* u, float* host, int n);
// forward on a device NHWC bf16 input (H x W x 8); out: device FP32 H x W x 2
// (null: the network's own output buffer, which unet_backward reads)
int unet_forward(hipStream_t st, UNet* u, const void* in_bf16, float* out);
// Backward of loss = 0.5 * sum((out - target)^2) / (2 H W) after unet_forward
// on the same input.  Gradients and the loss accumulate from unet_zero_grad on
// (a batch is several forward/backward pairs); unet_read_loss waits for the
// stream and returns the accumulated loss.
int unet_zero_grad(hipStream_t st, UNet* u);
int unet_backward(hipStream_t st, UNet* u, const void* in_bf16, const float* target);
int unet_read_loss(hipStream_t st, UNet* u, double* loss);
int unet_get_grads(UNet* u, float* host, int n);
int unet_adam(hipStream_t st, UNet* u, float lr, float beta1, float beta2, float eps);
// per-layer MFMA FLOPs of one forward pass (for the roofline)
double unet_forward_flops(const UNet* u);
int unet_copy_params(hipStream_t st, UNet* dst, const UNet* src);  // same parameters, any image size
int unet_size(const UNet* u, int* H, int* W);
// (h x w) window at (y0, x0) of an H x W image with c channels of esz bytes (c * esz * w % 4 == 0)
void launch_crop(hipStream_t st, int W, int c, int esz, const void* src, int y0, int x0, int h, int w, void* dst);

// velocity vectors <-> ny x nx network images (kernels_unet.hip)
struct Grid;
// ss[0..2] = |r|^2, |d|^2, |z|^2 (device; part: 768 doubles of scratch)
void launch_norms3(hipStream_t st, int64_t n, const double* r, const double* d, const double* z, double* part,
                   double* ss);
// 8-channel bf16 input image of r (vector the preconditioner is applied to),
// d = r - A z, z = DCT solve, and the coefficient fields log eta and chi
void launch_pack_features(hipStream_t st, const Grid& g, const double* r, const double* d, const double* z,
                          const double* eta, const double* chi, const double* ss, void* img);
// w = s_d * out as a velocity Krylov vector, s_d = rms of d (0 on the u faces i = nx)
void launch_unpack_correction(hipStream_t st, const Grid& g, const float* out, const double* ss, double* w);
// z += (dots[0] / dots[1]) c (device scalars; no step when dots[1] <= 0)
void launch_axpy_ratio(hipStream_t st, int64_t n, const double* dots, const double* c, double* z);
// t = v / s_d on the image (training target, v = A^-1 r - z)
void launch_make_target(hipStream_t st, const Grid& g, const double* v, const double* ss, float* t);

}  // namespace drd
