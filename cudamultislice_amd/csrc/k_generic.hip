// k_generic.hip -- 2D FFT and pointwise hooks for grids whose sides are not
// powers of two (the reference's cuFFT plan takes any resX x resY,
// U.cu:1961; its GUI offers 4008x2672, 2004x1336 and 1002x668, MW.cs:147-149).
//
// Lines of any length n are transformed by a mixed-radix Stockham FFT in
// global memory: n = r1 r2 ... rk (prime factors, 2s paired into 4s), one
// launch per factor, ping-ponging between the data and a scratch array of
// the same size.  A thread computes one output of one stage: with L the
// length already combined, output o = (i / L) L r + j + p L (i = (o / (L r)) L
// + j, j = o mod L, p = (o mod L r) / L) is
//   sum_q x[i + q n / r] exp(-+2 pi i q (j + p L) / (L r)),
// the exponents read from the table exp(-2 pi i k / n), k < n.  A radix r
// costs r complex multiply-adds per output: the large prime factors of the
// GUI sizes (167) make these passes compute-bound, but they are the
// compatibility path; power-of-two grids keep the fused LDS passes.
#include <hip/hip_runtime.h>

#include <vector>

#include "msl_internal.hpp"

namespace msl {

namespace {

__host__ __device__ inline long long gnblk(long long n, int b) { return (n + b - 1) / b; }

// line l of a batch of waves: (l / lpw) * wave_stride + (l % lpw) * line_stride
struct GenLines {
  long long nlines, lpw, wave_stride, line_stride, elem_stride;
  int n;
  int lines_fast;  // 1: consecutive threads take consecutive lines (columns)
};

__global__ void k_gen_stage(const float2* __restrict__ src, float2* __restrict__ dst, GenLines g, int r, int L,
                            int dir, const float2* __restrict__ tw) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = g.nlines * g.n;
  if (tid >= total) return;
  long long line;
  int o;
  if (g.lines_fast) {
    line = tid % g.nlines;
    o = (int)(tid / g.nlines);
  } else {
    o = (int)(tid % g.n);
    line = tid / g.n;
  }
  const long long base = (line / g.lpw) * g.wave_stride + (line % g.lpw) * g.line_stride;
  const int Lr = L * r, n = g.n, m = n / r;
  const int blk = o / Lr, rem = o - blk * Lr, p = rem / L, j = rem - p * L;
  const int i = blk * L + j;
  // exponent step of q, in units of 2 pi / n: (j + p L) n / (L r)
  const int step = (int)(((long long)(j + p * L) * (n / Lr)) % n);
  float2 acc = make_float2(0.f, 0.f);
  int e = 0;
  for (int q = 0; q < r; ++q) {
    const float2 x = src[base + (long long)(i + q * m) * g.elem_stride];
    float2 w = tw[e];
    if (dir > 0) w.y = -w.y;
    acc.x += x.x * w.x - x.y * w.y;
    acc.y += x.x * w.y + x.y * w.x;
    e += step;
    if (e >= n) e -= n;
  }
  dst[base + (long long)o * g.elem_stride] = acc;
}

// radices of n: factors 4 first, then 2, then odd primes ascending
std::vector<int> radices(int n) {
  std::vector<int> r;
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  for (int p = 3; (long long)p * p <= n; p += 2)
    while (n % p == 0) { r.push_back(p); n /= p; }
  if (n > 1) r.push_back(n);
  return r;
}

}  // namespace

hipError_t launch_gen_fft2(float2* data, float2* scratch, int nx, int ny, int nbatch, int inverse,
                           const float2* twx, const float2* twy, hipStream_t s) {
  if (nx < 1 || ny < 1 || nbatch < 0 || !data || !scratch || !twx || !twy) return hipErrorInvalidValue;
  if (nbatch == 0) return hipSuccess;
  const long long np = (long long)nx * ny;
  const int dir = inverse ? +1 : -1;
  // rows: nbatch * ny lines of nx contiguous elements
  GenLines gr{(long long)nbatch * ny, (long long)ny, np, (long long)nx, 1, nx, 0};
  // columns: nbatch * nx lines of ny elements strided by nx
  GenLines gc{(long long)nbatch * nx, (long long)nx, np, 1, (long long)nx, ny, 1};
  const size_t bytes = (size_t)np * nbatch * sizeof(float2);
  for (int pass = 0; pass < 2; ++pass) {
    const GenLines& g = pass == 0 ? gr : gc;
    const float2* tw = pass == 0 ? twx : twy;
    const std::vector<int> rs = radices(g.n);
    float2* a = data;
    float2* b = scratch;
    int L = 1;
    const long long total = g.nlines * g.n;
    for (int r : rs) {
      hipLaunchKernelGGL(k_gen_stage, dim3((unsigned)gnblk(total, 256)), dim3(256), 0, s, a, b, g, r, L, dir, tw);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      L *= r;
      float2* t = a;
      a = b;
      b = t;
    }
    if (a != data) {
      const hipError_t e = hipMemcpyAsync(data, a, bytes, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// ---------------------------------------------------------------- pointwise hooks
// w[b][y][x] *= m(x, y) * scale for the k-space multipliers of the passes:
// CM_PROP / CM_PROPCONJ the Fresnel propagator px[x] py[y] inside |k| < kmax
// (CreatePropsKernel, U.cu:419), CM_BAND the band limit |k| <= kmax
// (BandLimitKernel, U.cu:1037-1040), CM_GRAD -4 pi^2 k^2 (gradKernel,
// U.cu:959-960, pi = 3.14159f), CM_BANDGRAD both, CM_SCALE the scale only.
__global__ void k_gen_kmul(float2* w, int nx, int ny, long long np, int nbatch, int mul, const float* kx,
                           const float* ky, const float2* px, const float2* py, float kmax, float scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np * nbatch) return;
  const long long e = i % np;
  const int y = (int)(e / nx), x = (int)(e - (long long)y * nx);
  float2 m = make_float2(scale, 0.f);
  if (mul != CM_SCALE) {
    const float kxv = kx[x], kyv = ky[y];
    const float k = sqrtf(kxv * kxv + kyv * kyv);
    if (mul == CM_PROP || mul == CM_PROPCONJ) {
      float2 p = make_float2(px[x].x * py[y].x - px[x].y * py[y].y, px[x].x * py[y].y + px[x].y * py[y].x);
      if (mul == CM_PROPCONJ) p.y = -p.y;
      m = k < kmax ? make_float2(p.x * scale, p.y * scale) : make_float2(0.f, 0.f);
    } else if (mul == CM_BAND) {
      m = make_float2(k <= kmax ? scale : 0.f, 0.f);
    } else if (mul == CM_GRAD) {
      m = make_float2(scale * -4.f * 3.14159f * 3.14159f * (kxv * kxv + kyv * kyv), 0.f);
    } else {  // CM_BANDGRAD
      const float gr = -4.f * 3.14159f * 3.14159f * (kxv * kxv + kyv * kyv);
      m = make_float2(k <= kmax ? scale * gr : 0.f, 0.f);
    }
  }
  const float2 v = w[i];
  w[i] = make_float2(v.x * m.x - v.y * m.y, v.x * m.y + v.y * m.x);
}

hipError_t launch_gen_kmul(float2* w, int nx, int ny, int nbatch, int mul, const float* kx, const float* ky,
                           const float2* px, const float2* py, float kmax, float scale, hipStream_t s) {
  const long long np = (long long)nx * ny;
  hipLaunchKernelGGL(k_gen_kmul, dim3((unsigned)gnblk(np * nbatch, 256)), dim3(256), 0, s, w, nx, ny, np, nbatch,
                     mul, kx, ky, px, py, kmax, scale);
  return hipGetLastError();
}

// dst = src * t (elementwise, one wave)
__global__ void k_gen_cmul(const float2* src, const float2* t, float2* dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 a = src[i], b = t[i];
  dst[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

hipError_t launch_gen_cmul(const float2* src, const float2* t, float2* dst, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_cmul, dim3((unsigned)gnblk(n, 256)), dim3(256), 0, s, src, t, dst, n);
  return hipGetLastError();
}

// per-block sums of scale |F|^2 over kin <= |k| < kout (one wave), then one
// block adds the partial sums in a fixed order into out[0]
__global__ void k_gen_annulus(const float2* F, int nx, long long np, const float* kx, const float* ky, float kin,
                              float kout, float scale, float* partial) {
  __shared__ float sh[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float v = 0.f;
  if (i < np) {
    const int y = (int)(i / nx), x = (int)(i - (long long)y * nx);
    const float k = sqrtf(kx[x] * kx[x] + ky[y] * ky[y]);
    if (k >= kin && k < kout) v = scale * (F[i].x * F[i].x + F[i].y * F[i].y);
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_gen_sum(const float* partial, int n, float* out) {
  __shared__ float sh[256];
  float v = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) v += partial[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

int gen_annulus_blocks(long long np) { return (int)gnblk(np, 256); }

hipError_t launch_gen_annulus(const float2* F, int nx, int ny, const float* kx, const float* ky, float kin,
                              float kout, float scale, float* partial, float* out, hipStream_t s) {
  const long long np = (long long)nx * ny;
  const int nb = gen_annulus_blocks(np);
  hipLaunchKernelGGL(k_gen_annulus, dim3(nb), dim3(256), 0, s, F, nx, np, kx, ky, kin, kout, scale, partial);
  hipLaunchKernelGGL(k_gen_sum, dim3(1), dim3(256), 0, s, partial, nb, out);
  return hipGetLastError();
}

}  // namespace msl
