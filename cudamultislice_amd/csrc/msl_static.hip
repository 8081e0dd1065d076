// msl_static.hip -- the batched 2D FFT of msl_fft2_device for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops): the twiddle tables msl_fft2_device uploads into buffers it
// allocates per call live here in one file-scope device array instead.
//
// scratch holds the row twiddles exp(-2 pi i k / nx) at [0, nx) and the
// column twiddles exp(-2 pi i k / ny) at [kTwMax, kTwMax + ny), nx, ny <= 2048:
// 2 x 2048 float2 = 32 KiB per GPU, kept until the process ends.  Every call
// on a GPU rewrites it: calls that overlap on one GPU (concurrent streams or
// host threads) race on it and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/msl.h"
#include "msl_internal.hpp"

namespace msl {
namespace {

constexpr int kTwMax = 2048;
constexpr int kMaxDevices = 64;

__device__ __align__(256) float2 scratch[2 * kTwMax];

// Device address of scratch per device, looked up once per device with that
// device current (each GPU holds its own copy of the array).
std::mutex g_mu;
float2* g_ptr[kMaxDevices] = {};

hipError_t scratch_ptr(float2** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = static_cast<float2*>(p);
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

// tw[k] = exp(-2 pi i k / n), k < n, in double then rounded to float (the
// host tables of msl_fft2_device are computed the same way)
__global__ void k_twiddles(float2* tw, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double a = -2.0 * M_PI * (double)k / (double)n;
  tw[k] = make_float2((float)cos(a), (float)sin(a));
}

}  // namespace
}  // namespace msl

using namespace msl;

extern "C" msl_status msl_fft2_device_noalloc(int nx, int ny, int nbatch, void* data, int inverse, void* stream) {
  if (nbatch < 0 || !data) return MSL_ERR_ARG;
  // the static twiddle array holds lengths up to kTwMax; the column pass
  // puts the batch on grid.y
  if (!fft_size_supported(nx) || !fft_size_supported(ny) || nx > kTwMax || ny > kTwMax || nbatch > 65535)
    return MSL_ERR_UNSUPPORTED;
  if (nbatch == 0) return MSL_OK;
  float2* sp = nullptr;
  if (scratch_ptr(&sp) != hipSuccess) return MSL_ERR_HIP;
  const hipStream_t st = (hipStream_t)stream;
  float2* twx = sp;
  float2* twy = sp + kTwMax;
  hipLaunchKernelGGL(k_twiddles, dim3((nx + 255) / 256), dim3(256), 0, st, twx, nx);
  hipLaunchKernelGGL(k_twiddles, dim3((ny + 255) / 256), dim3(256), 0, st, twy, ny);
  if (hipGetLastError() != hipSuccess) return MSL_ERR_HIP;
  float2* w = (float2*)data;
  const long long bs = (long long)nx * ny;
  RowArgs r{};
  r.src = w; r.dst = w; r.src_bstride = r.dst_bstride = bs; r.nx = nx; r.ny = ny; r.nbatch = nbatch;
  r.t_group = 1; r.scale = 1.f; r.tw = twx;
  ColArgs c{};
  c.src = w; c.dst = w; c.src_bstride = c.dst_bstride = bs; c.nx = nx; c.ny = ny; c.nbatch = nbatch;
  c.scale = 1.f; c.tw = twy;
  hipError_t e;
  if (!inverse) {
    e = launch_row(ROW_FWD, RM_NONE, r, st);
    if (e == hipSuccess) e = launch_col(COL_FWD, CM_SCALE, c, st);
  } else {
    e = launch_col(COL_INV, CM_SCALE, c, st);
    if (e == hipSuccess) e = launch_row(ROW_INV, RM_NONE, r, st);
  }
  return e == hipSuccess ? MSL_OK : MSL_ERR_HIP;
}
