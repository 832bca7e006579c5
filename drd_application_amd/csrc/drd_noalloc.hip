// drd_noalloc.hip -- the particle velocity averages without a context and
// without any device allocation (drd_particle_velocity_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) pass device arrays
// for every input and output; the work buffers (bodies in the kernels'
// layout, per-chunk partial sums, per-body finishes) live in one file-scope
// device array of kScratchBytes per GPU.  Its address on each device is
// taken once, with that device current, by hipGetSymbolAddress; host code
// never hands the array's own name to a kernel.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "kernels.hpp"

using namespace drd;

namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB per device
constexpr int64_t kScratchDoubles = (int64_t)(kScratchBytes / sizeof(double));
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // per-device address of `scratch`

// The calling thread's current device's copy of `scratch` (looked up once
// per device).
int scratch_ptr(double** out, std::string* err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    *err = "hipGetDevice failed";
    return DRD_E_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    *err = "device index beyond the per-device table";
    return DRD_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return DRD_E_HIP;
    }
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

// bodies[5b ..] = (x, y, r, 0, 0): the layout the body kernels read
__global__ __launch_bounds__(kBlock) void k_pack_bodies(int n, const double* __restrict__ bx,
                                                        const double* __restrict__ by,
                                                        const double* __restrict__ br, double* __restrict__ bodies) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= n) return;
  double* e = bodies + 5 * (int64_t)b;
  e[0] = bx[b];
  e[1] = by[b];
  e[2] = br[b];
  e[3] = 0.0;
  e[4] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_unpack_aver(int n, const double* __restrict__ fin,
                                                        double* __restrict__ ua, double* __restrict__ va) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= n) return;
  ua[b] = fin[3 * (int64_t)b + 1];
  va[b] = fin[3 * (int64_t)b + 2];
}

int fail(int code, const std::string& msg) {
  api_set_error(msg);
  return code;
}

}  // namespace

extern "C" int64_t drd_noalloc_scratch_bytes(void) { return (int64_t)kScratchBytes; }

extern "C" int drd_particle_velocity_noalloc(const drd_params* prm, const double* u, const double* v,
                                             const double* bx, const double* by, const double* br, int32_t nbody,
                                             double lx, double ly, double* u_aver, double* v_aver,
                                             void* stream) {
  if (!prm || nbody < 0 || !(lx > 0.0) || !(ly > 0.0)) return fail(DRD_E_ARG, "bad argument");
  if (prm->nx < 4 || prm->ny < 4 || prm->nx > 4096 || prm->ny > 4096)
    return fail(DRD_E_ARG, "nx, ny must be in [4, 4096]");
  if (nbody == 0) return 0;
  if (!u || !v || !bx || !by || !br || !u_aver || !v_aver) return fail(DRD_E_ARG, "null array");
  const Grid g = Grid::make(prm->nx, prm->ny);
  // the chunking of drd_particle_velocity (its upload slot holds max(nuf, nc)
  // doubles), so both forms add the same partial sums in the same order
  const int64_t slot = std::max<int64_t>(g.nuf(), g.nc());
  if (5 * (int64_t)nbody > slot || nbody > 65535) return fail(DRD_E_ARG, "too many bodies for the grid");
  const int chunks = body_chunks(g, nbody, slot);
  const int64_t need = 8 * (int64_t)nbody + 3 * (int64_t)chunks * nbody;
  if (need > kScratchDoubles)
    return fail(DRD_E_ARG, "bodies x row chunks exceed the " + std::to_string(kScratchBytes >> 20) +
                               " MiB work array: " + std::to_string(need * 8) + " bytes needed");
  double* base = nullptr;
  std::string err;
  if (int rc = scratch_ptr(&base, &err)) return fail(rc, err);
  double* bodies = base;                        // 5 nbody
  double* fin = base + 5 * (int64_t)nbody;      // 3 nbody
  double* part = base + 8 * (int64_t)nbody;     // 3 chunks nbody
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned nbk = (unsigned)((nbody + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_pack_bodies, dim3(nbk), dim3(kBlock), 0, st, nbody, bx, by, br, bodies);
  launch_body_sums(st, g, prm->dx, prm->dz, prm->I_thickness, lx, ly, bodies, nbody, chunks, u, v, part, fin);
  hipLaunchKernelGGL(k_unpack_aver, dim3(nbk), dim3(kBlock), 0, st, nbody, fin, u_aver, v_aver);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(DRD_E_HIP, std::string("launch: ") + hipGetErrorString(e));
  return 0;
}
