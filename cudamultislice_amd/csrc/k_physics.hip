// k_physics.hip -- atom binning, potential synthesis, probe, CTF, detector
// and finite-difference kernels.  Citations are to the reference's
// "GUI Simulation/UnmanagedGUISimulation.cu" (U.cu).
#include <algorithm>
#include <cmath>

#include "msl_fft.hpp"
#include "msl_internal.hpp"

namespace msl {

static inline __host__ __device__ unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// ============================================================== binning
// atombin2 (U.cu:277-303): bin ids from FP32 sub, div, mul(50), floor --
// evaluated with IEEE division and no contraction so ids are bit-exact.
__global__ void k_bin_keys(BinArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float t1 = a.x[i] - a.MinX, t2 = a.MaxX - a.MinX;
  const float t3 = t1 / t2;
  const float t4 = t3 * 50.0f;
  const int bidx = (int)floorf(t4);
  const float u1 = a.y[i] - a.MinY, u2 = a.MaxY - a.MinY;
  const float u3 = u1 / u2;
  const float u4 = u3 * 50.0f;
  const int bidy = (int)floorf(u4);
  const float w1 = a.MaxZ - a.z[i];
  const float w2 = w1 / 1.0f;
  const int zid = (int)floorf(w2);
  if (bidx < 0 || bidx >= 50 || bidy < 0 || bidy >= 50 || zid < 0 || zid >= a.nslab) {
    a.key[i] = -1;
    atomicOr(a.bad, 1);
    return;
  }
  const int key = zid * 2500 + bidy * 50 + bidx;
  a.key[i] = key;
  atomicAdd(&a.count[key], 1);
}

// exclusive scan of count[0..nb] into start[0..nb], three phases
constexpr int SCAN_B = 256, SCAN_PER = 4, SCAN_CHUNK = SCAN_B * SCAN_PER;

__device__ int block_excl_scan(int v, int* sh, int& total) {
  // Hillis-Steele over 256 threads
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_B; off <<= 1) {
    int x = tid >= off ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += x;
    __syncthreads();
  }
  total = sh[SCAN_B - 1];
  const int incl = sh[tid];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(SCAN_B) k_scan_local(const int* in, int* out, int n, int* sums) {
  __shared__ int sh[SCAN_B];
  const long long base = (long long)blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
  int v[SCAN_PER], s = 0;
  for (int k = 0; k < SCAN_PER; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = block_excl_scan(s, sh, total);
  for (int k = 0; k < SCAN_PER; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_B) k_scan_sums(int* sums, int nsums) {
  __shared__ int sh[SCAN_B];
  int carry = 0;
  for (int b0 = 0; b0 < nsums; b0 += SCAN_B) {
    const int i = b0 + threadIdx.x;
    const int v = i < nsums ? sums[i] : 0;
    int total;
    const int ex = block_excl_scan(v, sh, total);
    if (i < nsums) sums[i] = ex + carry;
    carry += total;
  }
}

__global__ void k_scan_add(int* out, int n, const int* sums) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / SCAN_CHUNK];
}

__global__ void k_bin_scatter(BinArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int key = a.key[i];
  if (key < 0) return;
  a.perm[a.start[key] + atomicAdd(&a.fill[key], 1)] = i;
}

// restore input order inside each bin (the reference's push_back order,
// U.cu:1479-1482), so the CSR is identical on every run
__global__ void k_bin_sort(BinArgs a, int nb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int s = a.start[b], e = a.start[b + 1];
  for (int i = s + 1; i < e; ++i) {
    const int v = a.perm[i];
    int j = i - 1;
    while (j >= s && a.perm[j] > v) {
      a.perm[j + 1] = a.perm[j];
      --j;
    }
    a.perm[j + 1] = v;
  }
}

// positions relative to Minimum*, U.cu:1479-1481
__global__ void k_bin_gather(BinArgs a, int nb) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.start[nb]) return;  // only slots filled by k_bin_scatter
  const int i = a.perm[s];
  if (i < 0 || i >= a.n) return;
  a.bx[s] = a.x[i] - a.MinX;
  a.by[s] = a.y[i] - a.MinY;
  a.bz[s] = a.z[i] - a.MinZ;
  a.bZ[s] = a.Z[i];
}

hipError_t launch_bin(const BinArgs& a, hipStream_t s) {
  const int nb = a.nslab * 2500;
  hipLaunchKernelGGL(k_bin_keys, dim3(nblk(a.n, 256)), dim3(256), 0, s, a);
  const int nscan = nb + 1;  // count[nb] == 0, so start[nb] = total
  const int nchunks = (int)nblk(nscan, SCAN_CHUNK);
  hipLaunchKernelGGL(k_scan_local, dim3(nchunks), dim3(SCAN_B), 0, s, a.count, a.start, nscan, a.scan_tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_B), 0, s, a.scan_tmp, nchunks);
  hipLaunchKernelGGL(k_scan_add, dim3(nblk(nscan, 256)), dim3(256), 0, s, a.start, nscan, a.scan_tmp);
  hipLaunchKernelGGL(k_bin_scatter, dim3(nblk(a.n, 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_bin_sort, dim3(nblk(nb, 256)), dim3(256), 0, s, a, nb);
  hipLaunchKernelGGL(k_bin_gather, dim3(nblk(a.n, 256)), dim3(256), 0, s, a, nb);
  return hipGetLastError();
}

// ============================================================== potential
// Per-Z constants, host side.  vatom (U.cu:160-193):
//   V(r) = sum_i A_i exp2(-B_i r) / r + sum_i C_i exp2(-D_i r^2)
//   A = 150.4121417 a, B = 2 pi sqrt(b) log2 e, C = 266.5157269 c d^-1.5,
//   D = pi^2 / d log2 e
// vzatom2 (U.cu:119-152):
//   Vz(r) = sum_i A'_i K0(B'_i r) + sum_i C'_i exp2(-D_i r^2)
//   A' = 300.8242834 a, B' = 2 pi sqrt(b), C' = 150.4121417 c / d
void kirkland_transform(const float* fp, float* zt) {
  const double pi = 3.141592654, l2e = 1.4426950408889634;
  for (int z = 0; z < 103; ++z) {
    const float* p = fp + z * 12;
    float* o = zt + z * 24;
    for (int i = 0; i < 3; ++i) {
      double a = p[2 * i], b = p[2 * i + 1], c = p[6 + 2 * i], d = p[6 + 2 * i + 1];
      o[2 * i] = (float)(150.4121417 * a);
      o[2 * i + 1] = (float)(2.0 * pi * std::sqrt(b) * l2e);
      o[6 + 2 * i] = (float)(266.5157269 * c * std::pow(d, -1.5));
      o[6 + 2 * i + 1] = (float)(pi * pi / d * l2e);
      o[12 + 2 * i] = (float)(300.8242834 * a);
      o[12 + 2 * i + 1] = (float)(2.0 * pi * std::sqrt(b));
      o[18 + 2 * i] = (float)(150.4121417 * c / d);
      o[18 + 2 * i + 1] = (float)(pi * pi / d * l2e);
    }
  }
}

// The innermost function of the potential kernels: one square root, six
// base-2 exponentials and a reciprocal per sample, taken as the hardware
// instructions (about 1 ulp each).  The library forms of sqrtf, exp2f and the
// division are correctly rounded over the whole range, denormals included,
// at 82 instructions per sample against about 30 here; their difference is
// ~1e-7 of the potential.  Exponentials below 2^-126 flush to zero: terms of
// that size are below the FP32 resolution of the O(1)..O(1e3) V sums.
__device__ __forceinline__ float vatom_dev(const float* c, float r2) {
  float r = __builtin_amdgcn_sqrtf(r2);
  r = fmaxf(r, 0.25f);
  const float rr = r * r;
  const float sl = c[0] * __builtin_amdgcn_exp2f(-c[1] * r) + c[2] * __builtin_amdgcn_exp2f(-c[3] * r) +
                   c[4] * __builtin_amdgcn_exp2f(-c[5] * r);
  const float sg = c[6] * __builtin_amdgcn_exp2f(-c[7] * rr) + c[8] * __builtin_amdgcn_exp2f(-c[9] * rr) +
                   c[10] * __builtin_amdgcn_exp2f(-c[11] * rr);
  return sl * __builtin_amdgcn_rcpf(r) + sg;
}

// Abramowitz & Stegun 9.8.1-9.8.6 in FP32 (the forms of U.cu:41-111)
__device__ __forceinline__ float bessi0_dev(float x) {
  const float ax = fabsf(x);
  if (ax <= 3.75f) {
    float t = x / 3.75f;
    t = t * t;
    return 1.0f + t * (3.5156229f + t * (3.0899424f + t * (1.2067492f + t * (0.2659732f + t * (0.0360768f + t * 0.0045813f)))));
  }
  const float t = 3.75f / ax;
  const float s = 0.39894228f + t * (0.01328592f + t * (0.00225319f + t * (-0.00157565f + t * (0.00916281f +
                  t * (-0.02057706f + t * (0.02635537f + t * (-0.01647633f + t * 0.00392377f)))))));
  return expf(ax) * s / sqrtf(ax);
}

__device__ __forceinline__ float bessk0_dev(float x) {
  const float ax = fabsf(x);
  if (ax > 0.f && ax <= 2.0f) {
    float x2 = ax * 0.5f;
    x2 = x2 * x2;
    const float s = -0.57721566f + x2 * (0.42278420f + x2 * (0.23069756f + x2 * (0.03488590f + x2 * (0.00262698f +
                    x2 * (0.00010750f + x2 * 0.00000740f)))));
    return -logf(ax * 0.5f) * bessi0_dev(x) + s;
  } else if (ax > 2.0f) {
    const float x2 = 2.0f / ax;
    const float s = 1.25331414f + x2 * (-0.07832358f + x2 * (0.02189568f + x2 * (-0.01062446f + x2 * (0.00587872f +
                    x2 * (-0.00251540f + x2 * 0.00053208f)))));
    return expf(-ax) * s / sqrtf(ax);
  }
  return 1.0e20f;
}

__device__ __forceinline__ float vzatom_dev(const float* c, float r2) {
  float r = sqrtf(r2);
  r = fmaxf(r, 0.25f);
  const float rr = r * r;
  const float sl = c[12] * bessk0_dev(c[13] * r) + c[14] * bessk0_dev(c[15] * r) + c[16] * bessk0_dev(c[17] * r);
  const float sg = c[18] * exp2f(-c[19] * rr) + c[20] * exp2f(-c[21] * rr) + c[22] * exp2f(-c[23] * rr);
  return sl + sg;
}

// FP32 distance decisions of the reference kernels (U.cu:637-640, 784, 931):
// pixel position fma(x, ps, offset), squared distance as fma(dy, dy, dx*dx)
// and fma(dz, dz, rxy).  Written out explicitly, with contraction off, so
// the r^2 < 5 cutoff is decided the same way on every build.
__device__ __forceinline__ float pix_pos(int i, float ps, float off) {
#pragma clang fp contract(off)
  return fmaf((float)i, ps, off);
}
__device__ __forceinline__ float dist2_xy(float px, float py, float ax, float ay) {
#pragma clang fp contract(off)
  const float dx = px - ax, dy = py - ay;
  return fmaf(dy, dy, dx * dx);
}
__device__ __forceinline__ float dist2_z(float rxy, float z, float az) {
#pragma clang fp contract(off)
  const float dd = z - az;
  return fmaf(dd, dd, rxy);
}
__device__ __forceinline__ float sample_z(float zf, int m, float dz) {
#pragma clang fp contract(off)
  return zf - ((float)m * dz) / 20.0f;
}
// binnedAtomPKernelRS samples z - m dz / 10 with an integer dz of 1 (U.cu:700)
__device__ __forceinline__ float sample_z_rs(float zf, int m) {
#pragma clang fp contract(off)
  return zf - (float)m / 10.0f;
}

constexpr int PT = 16;      // pixel tile edge of the grouped launches (serving)
// pixel tiles of 256 threads selectable per handle (PotArgs::tile, chosen by
// msl_autotune_potential): 16x16, 32x8, 64x4, 8x32.  Every shape sums each
// pixel's atoms in the same (slab, bin row, bin, CSR) order, so all give the
// same bits; they differ in the bin window each workgroup stages.
template <int V> struct PotTile;
template <> struct PotTile<0> { static constexpr int X = 16, Y = 16; };
template <> struct PotTile<1> { static constexpr int X = 32, Y = 8; };
template <> struct PotTile<2> { static constexpr int X = 64, Y = 4; };
template <> struct PotTile<3> { static constexpr int X = 8, Y = 32; };
constexpr int kPotTiles = 4;
constexpr int ACAP = 1024;  // atoms staged in LDS per chunk

// binnedAtomPKernel3 (U.cu:582-653), binnedAtomPKernelConventional3
// (U.cu:730-795) and binnedAtomPKernelFD (U.cu:883-946).  One workgroup
// owns a 16x16 pixel tile of one slice; the atoms of every CSR bin within
// sqrt(5) A of the tile (slab window per mode) are staged through LDS in
// chunks, then each thread sums its pixel's contributions.
// the potential arguments as device code reads them from a table in memory
// (pointers in the global address space, same layout)
struct PotArgsG {
  MSL_POT_ARGS(MSL_GLOBAL)
};
static_assert(sizeof(PotArgsG) == sizeof(PotArgs), "layouts differ");

// pixel tile (tx, ty) of TX x TY pixels (TX TY = 256 threads) of slice
// `slice` (relative to first_slice)
template <int TX, int TY, class A>
__device__ __forceinline__ void potential_tile(const A& a, int tx, int ty, int slice) {
  static_assert(TX * TY == 256, "one thread per pixel of a 256-thread workgroup");
  __shared__ float zt[103 * 24];
  __shared__ float sx[ACAP], sy[ACAP], sz[ACAP];
  __shared__ int sZ[ACAP];
  __shared__ int rs[64], rl[65];
  __shared__ float zsm[21];  // the slice's z samples (modes 0 and 3), index m
  const int tid = threadIdx.x;
  for (int i = tid; i < 103 * 24; i += 256) zt[i] = a.ztab[i];
  const int ix = tx * TX + (tid % TX), iy = ty * TY + (tid / TX);
  float zf, h;
  {
#pragma clang fp contract(off)
    h = a.MaxZ - a.MinZ;
    // compat=physical: slice i starts at the top of the specimen (i - 1 slices down)
    const int phys = (a.mode < 2 && a.physical) ? 1 : 0;
    zf = (a.mode >= 2) ? a.zplane : h - a.dz * (float)(a.first_slice + slice - phys);
  }
  int k0, k1;
  if (a.mode == 1) {
    k0 = (int)floorf(h - zf);
    k1 = (int)floorf(h - zf + a.dz) - 1;
  } else {
    const int kc = (int)floorf((h - zf) / 1.0f);
    k0 = kc - a.loadslicesz;
    k1 = kc + a.loadslicesz;
    if (a.mode == 0 && a.physical) {
      // down to the slab of the lowest sample, +-3 (every atom within sqrt(5))
#pragma clang fp contract(off)
      const float zb = zf - a.dz;
      k1 = (int)floorf((h - zb) / 1.0f) + a.loadslicesz;
    }
  }
  k0 = max(k0, 0);
  k1 = min(k1, a.nslab - 1);
  const float offx = a.mode == 2 ? 0.f : a.offx, offy = a.mode == 2 ? 0.f : a.offy;
  const float px = pix_pos(ix, a.ps, offx), py = pix_pos(iy, a.ps, offy);
  // bin window of the tile: every bin that may hold an atom within sqrt(5)
  const float wx = (a.MaxX - a.MinX) / 50.f, wy = (a.MaxY - a.MinY) / 50.f, reach = 2.2360680f;
  const float tx0 = offx + (float)(tx * TX) * a.ps, tx1 = offx + (float)(tx * TX + TX - 1) * a.ps;
  const float ty0 = offy + (float)(ty * TY) * a.ps, ty1 = offy + (float)(ty * TY + TY - 1) * a.ps;
  const int i0 = max(0, (int)floorf((tx0 - reach) / wx) - 1), i1 = min(49, (int)floorf((tx1 + reach) / wx) + 1);
  const int j0 = max(0, (int)floorf((ty0 - reach) / wy) - 1), j1 = min(49, (int)floorf((ty1 + reach) / wy) + 1);
  const float h20 = a.dz / 20.0f;
  // the sample positions depend on the slice only: computed once per
  // workgroup (each an FP32 division, as the reference takes it) instead of
  // once per atom and pixel
  if (tid >= 1 && tid <= 20) zsm[tid] = a.mode == 3 ? (tid <= 10 ? sample_z_rs(zf, tid) : 0.f) : sample_z(zf, tid, a.dz);
  float sum = 0.f;
  __syncthreads();
  for (int k = k0; k <= k1; ++k) {
    for (int jb = j0; jb <= j1; jb += 64) {
      const int nr = min(64, j1 - jb + 1);
      if (i0 > i1) break;
      if (tid < 64) {
        // the first wave loads the window's bin-row runs and scans their
        // lengths (inclusive, shuffles across the 64 lanes)
        int len = 0;
        if (tid < nr) {
          const int bb = k * 2500 + (jb + tid) * 50;
          const int r0 = a.start[bb + i0];
          rs[tid] = r0;
          len = a.start[bb + i1 + 1] - r0;
        }
        for (int d = 1; d < 64; d <<= 1) {
          const int v = __shfl_up(len, d, 64);
          if (tid >= d) len += v;
        }
        if (tid < nr) rl[tid + 1] = len;
        if (tid == 0) rl[0] = 0;
      }
      __syncthreads();
      const int tot = rl[nr];
      for (int c0 = 0; c0 < tot; c0 += ACAP) {
        const int cn = min(ACAP, tot - c0);
        for (int q = tid; q < cn; q += 256) {
          const int g = c0 + q;
          // the run holding g: the largest r with rl[r] <= g (rl ascending)
          int lo = 0, hi = nr - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rl[mid] <= g) lo = mid;
            else hi = mid - 1;
          }
          const int r = lo;
          const int src = rs[r] + (g - rl[r]);
          sx[q] = a.bx[src];
          sy[q] = a.by[src];
          sz[q] = a.bz[src];
          sZ[q] = a.bZ[src];
        }
        __syncthreads();
        for (int l = 0; l < cn; ++l) {
          const float rxy = dist2_xy(px, py, sx[l], sy[l]);
          if (rxy >= 5.0f) continue;
          const float* c = zt + (sZ[l] - 1) * 24;
          if (a.mode == 0) {
            const float zr = sz[l];
            // samples m with |zf - m dz / 20 - zr| < sqrt(5 - rxy), widened by
            // one sample on each side; inside that range each sample takes
            // the exact FP32 cutoff test, so the set summed is the 20-sample
            // loop's (the samples outside fail it by more than rounding, and
            // by far more than the hardware square root's last bit)
            const float hz = __builtin_amdgcn_sqrtf(5.0f - rxy), u = 20.0f / a.dz;
            const int mlo = max(1, (int)floorf((zf - zr - hz) * u) - 1);
            const int mhi = min(20, (int)ceilf((zf - zr + hz) * u) + 1);
            // the atom's coefficients in registers for all its samples
            float cr[12];
#pragma unroll
            for (int i = 0; i < 12; ++i) cr[i] = c[i];
            for (int m = mlo; m <= mhi; ++m) {
              const float r2 = dist2_z(rxy, zsm[m], zr);
              if (r2 < 5.0f) sum += h20 * vatom_dev(cr, r2);
            }
          } else if (a.mode == 1) {
            sum += vzatom_dev(c, rxy);
          } else if (a.mode == 3) {  // real-space slab of 1 A, 10 samples
            const float zr = sz[l];
            float cr[12];
#pragma unroll
            for (int i = 0; i < 12; ++i) cr[i] = c[i];
            for (int m = 1; m <= 10; ++m) {
              const float r2 = dist2_z(rxy, zsm[m], zr);
              if (r2 < 5.0f) sum += 0.1f * vatom_dev(cr, r2);
            }
          } else {
            const float r2 = dist2_z(rxy, zf, sz[l]);
            if (r2 < 5.0f) sum += vatom_dev(c, r2);
          }
        }
        __syncthreads();
      }
      __syncthreads();  // rs / rl are rewritten by the next window row
    }
  }
  if (ix >= a.nx || iy >= a.ny) return;
  const long long o = (long long)slice * a.out_sstride + (long long)iy * a.nx + ix;
  if (a.out_kind == 0) {
    float sn, cs;
    sincosf(a.sigma * sum, &sn, &cs);
    ((float2*)a.out)[o] = make_float2(cs, sn);
  } else if (a.out_kind == 1) {
    ((float2*)a.out)[o] = make_float2(sum, 0.f);
  } else {
    ((float*)a.out)[o] = sum;
  }
}

template <int V>
__global__ void __launch_bounds__(256) k_potential(PotArgs a) {
  potential_tile<PotTile<V>::X, PotTile<V>::Y>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// slice `first` of njobs jobs in one launch: workgroup id -> (job, tile)
// through the block offsets blk0[0..njobs] (ascending)
__global__ void __launch_bounds__(256) k_potential_group(const PotArgs* tab, const long long* blk0, int njobs,
                                                         int first) {
  const long long id = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (blk0[mid] <= id) lo = mid;
    else hi = mid - 1;
  }
  PotArgsG a = reinterpret_cast<const PotArgsG*>(tab)[lo];
  a.first_slice = first;
  const int tx = nblk(a.nx, PT);
  const long long local = id - blk0[lo];
  potential_tile<PT, PT>(a, (int)(local % tx), (int)(local / tx), 0);
}

template <int V> static hipError_t potential_v(const PotArgs& a, hipStream_t s) {
  dim3 grid(nblk(a.nx, PotTile<V>::X), nblk(a.ny, PotTile<V>::Y), a.nslices);
  hipLaunchKernelGGL(k_potential<V>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_potential(const PotArgs& a, hipStream_t s) {
  if (a.nslices <= 0) return hipSuccess;
  switch (a.tile) {
    case 1: return potential_v<1>(a, s);
    case 2: return potential_v<2>(a, s);
    case 3: return potential_v<3>(a, s);
    default: return potential_v<0>(a, s);
  }
}

int potential_tile_variants() { return kPotTiles; }

int potential_tiles(int nx, int ny) { return nblk(nx, PT) * nblk(ny, PT); }

hipError_t launch_potential_group(const PotArgs* tab, const long long* blk0, int njobs, long long nblocks, int first,
                                  hipStream_t s) {
  if (njobs <= 0 || nblocks <= 0) return hipSuccess;
  if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_potential_group, dim3((unsigned)nblocks), dim3(256), 0, s, tab, blk0, njobs, first);
  return hipGetLastError();
}

// ============================================================== probe
// STEMinitializingKernel (U.cu:354-383): exp(i(theta - chi)) inside
// k < apert.  theta = 2 pi (kx posx + ky posy) ps is reduced in FP64 so the
// ramp stays accurate for positions far from the origin.
__global__ void k_probe_kspace(ProbeArgs a) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, b = blockIdx.z;
  if (a.band_cols > 0) {
    if (x >= 2 * a.band_cols) return;
    if (x >= a.band_cols) x += a.nx - 2 * a.band_cols;
  } else if (x >= a.nx) {
    return;
  }
  const float k0x = a.kx[x], k0y = a.ky[y];
  const float k = sqrtf(k0x * k0x + k0y * k0y);
  float2 v = make_float2(0.f, 0.f);
  if (k < a.apert) {
    const int px = a.posx ? a.posx[b] : a.px0, py = a.posy ? a.posy[b] : a.py0;
    const double k2 = (double)k0x * k0x + (double)k0y * k0y, wl = a.wavel, pi = 3.14159265;
    const double chi = pi * wl * k2 * ((double)a.Cs * wl * wl * k2 * 0.5 + (double)a.df);
    const double th = 2.0 * pi * ((double)k0x * px * (double)a.ps + (double)k0y * py * (double)a.ps);
    double ph = th - chi;
    ph -= 6.283185307179586 * floor(ph / 6.283185307179586);
    float sn, cs;
    sincosf((float)ph, &sn, &cs);
    v = make_float2(cs, sn);
  }
  a.out[(long long)b * a.bstride + (long long)y * a.nx + x] = v;
}

hipError_t launch_probe_kspace(const ProbeArgs& a, hipStream_t s) {
  if (a.band_cols < 0 || 2 * a.band_cols > a.nx) return hipErrorInvalidValue;
  // probes on grid.z, at most 65535 per launch
  constexpr int kMaxGridZ = 65535;
  for (int b0 = 0; b0 < a.nprobe; b0 += kMaxGridZ) {
    ProbeArgs c = a;
    c.nprobe = a.nprobe - b0 < kMaxGridZ ? a.nprobe - b0 : kMaxGridZ;
    c.out = a.out + (long long)b0 * a.bstride;
    if (a.posx) {
      c.posx = a.posx + b0;
      c.posy = a.posy + b0;
    }
    const int cols = a.band_cols > 0 ? 2 * a.band_cols : a.nx;
    hipLaunchKernelGGL(k_probe_kspace, dim3(nblk(cols, 256), a.ny, c.nprobe), dim3(256), 0, s, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ============================================================== reductions
constexpr int RB = 256;

__device__ float block_sum(float v, float* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const float r = sh[0];
  __syncthreads();
  return r;
}

// per wave: partial[b][blk] = sum |w| over a fixed 4096-element chunk
constexpr int L1_CHUNK = 4096;
__global__ void __launch_bounds__(RB) k_abs_partial(const float2* w, long long n, long long bstride, float* partial,
                                                    int nchunk, int l2) {
  __shared__ float sh[RB];
  const int b = blockIdx.y, blk = blockIdx.x;
  const float2* p = w + (long long)b * bstride;
  float acc = 0.f;
  const long long s0 = (long long)blk * L1_CHUNK;
  for (int i = threadIdx.x; i < L1_CHUNK; i += RB) {
    const long long j = s0 + i;
    if (j < n) {
      const float i2 = p[j].x * p[j].x + p[j].y * p[j].y;
      acc += l2 ? i2 : sqrtf(i2);
    }
  }
  const float r = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[(long long)b * nchunk + blk] = r;
}

__global__ void __launch_bounds__(RB) k_abs_final(float* partial, int nchunk, float* norm, int l2) {
  __shared__ float sh[RB];
  const int b = blockIdx.x;
  float acc = 0.f;
  for (int i = threadIdx.x; i < nchunk; i += RB) acc += partial[(long long)b * nchunk + i];
  const float r = block_sum(acc, sh);
  if (threadIdx.x == 0) norm[b] = l2 ? sqrtf(r) : r;
}

__global__ void k_div_by(float2* w, long long n, long long bstride, const float* norm) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= n) return;
  const float inv = 1.0f / norm[b];
  float2* p = w + (long long)b * bstride;
  p[i] = cscale(p[i], inv);
}

hipError_t launch_l1_normalise(float2* w, long long n, int nbatch, long long bstride, float* scratch, hipStream_t s) {
  return launch_probe_normalise(w, n, nbatch, bstride, scratch, 0, s);
}

hipError_t launch_probe_normalise(float2* w, long long n, int nbatch, long long bstride, float* scratch, int l2,
                                  hipStream_t s) {
  const int nchunk = (int)nblk(n, L1_CHUNK);
  float* partial = scratch;
  float* norm = scratch + (long long)nbatch * nchunk;
  hipLaunchKernelGGL(k_abs_partial, dim3(nchunk, nbatch), dim3(RB), 0, s, w, n, bstride, partial, nchunk, l2);
  hipLaunchKernelGGL(k_abs_final, dim3(nbatch), dim3(RB), 0, s, partial, nchunk, norm, l2);
  hipLaunchKernelGGL(k_div_by, dim3(nblk(n, 256), nbatch), dim3(256), 0, s, w, n, bstride, norm);
  return hipGetLastError();
}

__global__ void k_fill(float2* w, long long n, float2 v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = v;
}
hipError_t launch_fill(float2* w, long long n, float2 v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, s, w, n, v);
  return hipGetLastError();
}

__global__ void k_hash_fill(float2* w, long long n, unsigned seed, int unit) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = (unsigned)i * 2654435761u ^ seed;
  h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
  const float u = (float)(h & 0xffffff) * (1.0f / 16777216.0f), v = (float)(h >> 8) * (1.0f / 16777216.0f);
  if (unit) {
    float sn, cs;
    sincosf(6.2831853f * u, &sn, &cs);
    w[i] = make_float2(cs, sn);
  } else {
    w[i] = make_float2(2.f * u - 1.f, 2.f * v - 1.f);
  }
}
hipError_t launch_hash_fill(float2* w, long long n, unsigned seed, int unit, hipStream_t s) {
  hipLaunchKernelGGL(k_hash_fill, dim3(nblk(n, 256)), dim3(256), 0, s, w, n, seed, unit);
  return hipGetLastError();
}

// plain streaming copy (16-byte vectors), used to calibrate traffic counters
__global__ void k_copy16(const float4* __restrict__ src, float4* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
hipError_t launch_copy16(const void* src, void* dst, long long bytes, hipStream_t s) {
  const long long n = bytes / 16;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::min<long long>(nblk(n, 256), 65536)), dim3(256), 0, s,
                     (const float4*)src, (float4*)dst, n);
  return hipGetLastError();
}

__global__ void k_scale(float2* w, long long n, float f) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = cscale(w[i], f);
}
hipError_t launch_scale(float2* w, long long n, float f, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(nblk(n, 256)), dim3(256), 0, s, w, n, f);
  return hipGetLastError();
}

__global__ void k_cmul(const float2* a, const float2* b, float2* o, long long n, float f) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = cscale(cmul(a[i], b[i]), f);
}
hipError_t launch_cmul(const float2* a, const float2* b, float2* o, long long n, float f, hipStream_t s) {
  hipLaunchKernelGGL(k_cmul, dim3(nblk(n, 256)), dim3(256), 0, s, a, b, o, n, f);
  return hipGetLastError();
}

__global__ void k_reduce_rows(const float* partial, int nbatch, int ntiles, float* out, int acc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  float s = 0.f;
  for (int i = 0; i < ntiles; ++i) s += partial[(long long)b * ntiles + i];
  out[b] = acc ? out[b] + s : s;
}
hipError_t launch_reduce_rows(const float* partial, int nbatch, int ntiles, float* out, int acc, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_rows, dim3(nblk(nbatch, 256)), dim3(256), 0, s, partial, nbatch, ntiles, out, acc);
  return hipGetLastError();
}

__global__ void k_abs(const float2* w, float* o, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = sqrtf(w[i].x * w[i].x + w[i].y * w[i].y);
}
hipError_t launch_abs(const float2* w, float* o, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_abs, dim3(nblk(n, 256)), dim3(256), 0, s, w, o, n);
  return hipGetLastError();
}

__global__ void k_intensity(const float2* w, float* o, long long n, float f) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (w[i].x * w[i].x + w[i].y * w[i].y) * f;
}
hipError_t launch_intensity(const float2* w, float* o, long long n, float f, hipStream_t s) {
  hipLaunchKernelGGL(k_intensity, dim3(nblk(n, 256)), dim3(256), 0, s, w, o, n, f);
  return hipGetLastError();
}

// ============================================================== real-space multislice
// One term of MultisliceStepRS's series (RSMS with DifferentialXKernel /
// DifferentialYKernel, U.cu:798-877): out = w + i pre (Lx w + Ly w + V w)
// with L (4/3 d1 - d2/12) / ps^2 per axis, d1 = f[+1] + f[-1] - f[0],
// d2 = f[+2] + f[-2] - f[0], and neighbours outside the grid reading element
// 0 of the row / column; the first term (w == null) is 1 + i pre V.
__global__ void k_rs_term(const float2* w, float2* out, const float2* V, int nx, int ny, float ps, float pre) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= nx) return;
  const long long row = (long long)y * nx, id = row + x;
  const float v = V[id].x;
  if (!w) {
    out[id] = make_float2(1.0f, pre * v);
    return;
  }
  const int xp1 = x > 0 ? x - 1 : 0, xp2 = x > 1 ? x - 2 : 0, xn1 = x < nx - 1 ? x + 1 : 0, xn2 = x < nx - 2 ? x + 2 : 0;
  const int yp1 = y > 0 ? y - 1 : 0, yp2 = y > 1 ? y - 2 : 0, yn1 = y < ny - 1 ? y + 1 : 0, yn2 = y < ny - 2 ? y + 2 : 0;
  const float2 f = w[id];
  const float2 a1 = w[row + xn1], b1 = w[row + xp1], a2 = w[row + xn2], b2 = w[row + xp2];
  const float2 c1 = w[(long long)yn1 * nx + x], e1 = w[(long long)yp1 * nx + x];
  const float2 c2 = w[(long long)yn2 * nx + x], e2 = w[(long long)yp2 * nx + x];
  const float s = 1.0f / (ps * ps);
  const float dxx = s * (4.0f * ((a1.x + b1.x) - f.x) / 3.0f - ((a2.x + b2.x) - f.x) / 12.0f);
  const float dxy = s * (4.0f * ((a1.y + b1.y) - f.y) / 3.0f - ((a2.y + b2.y) - f.y) / 12.0f);
  const float dyx = s * (4.0f * ((c1.x + e1.x) - f.x) / 3.0f - ((c2.x + e2.x) - f.x) / 12.0f);
  const float dyy = s * (4.0f * ((c1.y + e1.y) - f.y) / 3.0f - ((c2.y + e2.y) - f.y) / 12.0f);
  const float r1x = -1.0f * pre * (dxy + dyy + v * f.y);
  const float r1y = pre * (dxx + dyx + v * f.x);
  out[id] = make_float2(f.x + r1x, f.y + r1y);
}
hipError_t launch_rs_term(const float2* w, float2* out, const float2* V, int nx, int ny, float ps, float pre,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_rs_term, dim3((nx + 255) / 256, ny), dim3(256), 0, s, w, out, V, nx, ny, ps, pre);
  return hipGetLastError();
}

// ============================================================== detector model
// Radial curve of DQEFunc (U.cu:1120-1132): signed frequency indices, radius
// sqrtf(x'^2 + y'^2) / binning in FP32, entry floor(radius) up to 724, else
// 0.0001.  F *= scale * (root ? sqrtf(curve) : curve)  (DQEKernel takes the
// square root, NTFKernel the curve itself, U.cu:1134-1186).
__global__ void k_detector_filter(float2* F, int nx, int ny, const float* tab, int binning, int root, float scale) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= nx) return;
  const int xd = x < nx / 2 ? x : x - nx, yd = y < ny / 2 ? y : y - ny;
  const float r = sqrtf((float)(xd * xd + yd * yd)) / (float)binning;
  const int i = (int)floorf(r);
  const float v = i <= 724 ? tab[i] : 0.0001f;
  const float f = (root ? sqrtf(v) : v) * scale;
  float2& e = F[(long long)y * nx + x];
  e = make_float2(e.x * f, e.y * f);
}
hipError_t launch_detector_filter(float2* F, int nx, int ny, const float* tab, int binning, int root, float scale,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_detector_filter, dim3((nx + 255) / 256, ny), dim3(256), 0, s, F, nx, ny, tab, binning, root,
                     scale);
  return hipGetLastError();
}

// Philox4x64-10 (Salmon et al., SC'11) on the device; the same generator as
// the host's frozen-phonon displacements.
__device__ void philox4x64_dev(unsigned long long c[4], unsigned long long k0, unsigned long long k1) {
  const unsigned long long M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const unsigned long long W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  for (int r = 0; r < 10; ++r) {
    const unsigned long long hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    const unsigned long long hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    const unsigned long long n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

// Shot noise of the detector (U.cu:2375-2396), seeded: a = |z| zscale,
// m = dose a, counts = round(m + sqrt(|m|) g) with g the Box-Muller normal of
// Philox block (i, 0xDE7EC7, 3, 0) under key (seed, 0x4D534C), clamped to
// +-3; out = (counts, 0).
__global__ void k_shot_noise(const float2* z, float2* out, long long n, float zscale, float dose,
                             unsigned long long seed) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned long long c[4] = {(unsigned long long)i, 0xDE7EC7ULL, 3ULL, 0ULL};
  philox4x64_dev(c, seed, 0x4D534CULL);
  const double u1 = ((double)(c[0] >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(c[1] >> 11) * (1.0 / 9007199254740992.0);
  double g = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
  g = g < -3.0 ? -3.0 : (g > 3.0 ? 3.0 : g);
  const float2 v = z[i];
  const double a = (double)(sqrtf(v.x * v.x + v.y * v.y) * zscale);
  const double m = (double)dose * a;
  out[i] = make_float2((float)round(m + sqrt(fabs(m)) * g), 0.f);
}
hipError_t launch_shot_noise(const float2* z, float2* out, long long n, float zscale, float dose,
                             unsigned long long seed, hipStream_t s) {
  hipLaunchKernelGGL(k_shot_noise, dim3(nblk(n, 256)), dim3(256), 0, s, z, out, n, zscale, dose, seed);
  return hipGetLastError();
}

__global__ void k_real_to_complex(const float* in, float2* out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(in[i], 0.f);
}
hipError_t launch_real_to_complex(const float* in, float2* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_real_to_complex, dim3(nblk(n, 256)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

// ============================================================== CTF
// imagingKernel2, U.cu:1044-1080.  chi and the envelope arguments are
// evaluated in FP64 (they reach ~1e2 rad at large Cs).
__global__ void k_ctf(CtfArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, d = blockIdx.z;
  if (x >= a.nx) return;
  const long long id = (long long)y * a.nx + x;
  const float k0x = a.kx[x], k0y = a.ky[y];
  const float k2f = k0x * k0x + k0y * k0y;
  const float objap2 = ((a.obj * 0.001f) / a.wavel) * ((a.obj * 0.001f) / a.wavel);
  float2 out = make_float2(0.f, 0.f);
  if (k2f < objap2) {
    const double lam = a.wavel, k2 = (double)k0x * k0x + (double)k0y * k0y, k = sqrt(k2);
    const double df = a.dfs[d], cs = a.css ? a.css[d] : a.Cs, b = a.beta, dl = a.delta, p1 = 3.141593;
    const double factor = b * b / (4.0 * lam * lam);
    const double esa = p1 * k * lam * 2.0 * df + 2.0 * p1 * lam * lam * lam * cs * k2 * k;
    const double es = exp(-factor * esa * esa);
    const double ed = exp(-0.25 * dl * dl * p1 * p1 * k2 * k2 * lam * lam);
    const double phi = atan2((double)k0y, (double)k0x);
    const double chi = 3.1415926 * lam * k2 *
                       (0.5 * cs * lam * lam * k2 + df + (double)a.A1 * sin(2.0 * (phi - (double)a.A1phi)) +
                        2.0 * (double)a.A2 * lam * k * sin(3.0 * (phi - (double)a.A2phi)) / 3.0);
    double c = chi - 6.283185307179586 * floor(chi / 6.283185307179586);
    float sn, cs_;
    sincosf((float)c, &sn, &cs_);
    const float e = (float)(es * ed) * a.scale;
    const float2 v = a.in[id];
    out = make_float2(e * (v.x * cs_ + v.y * sn), e * (-v.x * sn + v.y * cs_));
  }
  a.out[(long long)d * a.nx * a.ny + id] = out;
}
hipError_t launch_ctf(const CtfArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_ctf, dim3(nblk(a.nx, 256), a.ny, a.ndf), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ============================================================== detectors
// CentreKernel + conjugationKernel + host sum of AddTDSWaves (U.cu:2734-2755)
__global__ void k_tds_accum(const float2* F, float* tds, int nx, int ny, int intensity, float scale) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= nx) return;
  const int x2 = x >= nx / 2 ? x - nx / 2 : x + nx / 2;
  const int y2 = y >= ny / 2 ? y - ny / 2 : y + ny / 2;
  const float2 v = F[(long long)y2 * nx + x2];
  const float i2 = v.x * v.x + v.y * v.y;
  tds[(long long)y * nx + x] += intensity ? scale * i2 : sqrtf(i2);
}
hipError_t launch_tds_accum(const float2* F, float* tds, int nx, int ny, hipStream_t s) {
  return launch_tds_accum_mode(F, tds, nx, ny, 0, 1.0f, s);
}
hipError_t launch_tds_accum_mode(const float2* F, float* tds, int nx, int ny, int intensity, float scale,
                                 hipStream_t s) {
  hipLaunchKernelGGL(k_tds_accum, dim3(nblk(nx, 256), ny), dim3(256), 0, s, F, tds, nx, ny, intensity, scale);
  return hipGetLastError();
}

__global__ void k_sum_maps(const float* in, int nb, long long n, float* out, int acc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += in[(long long)b * n + i];
  out[i] = acc ? out[i] + s : s;
}
hipError_t launch_sum_maps(const float* in, int nb, long long n, float* out, int acc, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_maps, dim3(nblk(n, 256)), dim3(256), 0, s, in, nb, n, out, acc);
  return hipGetLastError();
}

// ============================================================== FD
// finiteDifferenceKernel, U.cu:995-1023 (pi written 3.14159f there)
__global__ void k_fd_update(const float2* grad, const float2* psim, float2* psip, const float2* psi, const float2* V,
                            long long n, long long vper, float sigma, float wavel, float dz) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float a = 2.f * 3.14159f * dz / wavel;
  const float den = 1.f + a * a;
  const float2 rcp = make_float2(1.f / den, -a / den);                     // 1 / (1 + i a)
  const float2 ratio = make_float2((1.f - a * a) / den, -2.f * a / den);   // (1 - i a) / (1 + i a)
  const float2 p = psi[i], g = grad[i], m = psim[i];
  const float vv = dz * dz * 4.f * 3.14159f * sigma * V[i % vper].x / wavel;
  const float2 A = make_float2(2.f * p.x - dz * dz * g.x - vv * p.x, 2.f * p.y - dz * dz * g.y - vv * p.y);
  psip[i] = csub(cmul(rcp, A), cmul(ratio, m));
}
hipError_t launch_fd_update(const float2* grad, const float2* psim, float2* psip, const float2* psi, const float2* V,
                            long long n, long long vper, float sigma, float wavel, float dz, hipStream_t s) {
  if (vper <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fd_update, dim3(nblk(n, 256)), dim3(256), 0, s, grad, psim, psip, psi, V, n, vper, sigma, wavel,
                     dz);
  return hipGetLastError();
}

// ============================================================== transmission
// t = exp(i sigma V) of one slice, computed once and read by the row pass
// for every probe of the differentiable path
__global__ void k_transmission(const float* V, float2* t, long long n, float sigma) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sn, cs;
  sincosf(sigma * V[i], &sn, &cs);
  t[i] = make_float2(cs, sn);
}
hipError_t launch_transmission(const float* V, float2* t, long long n, float sigma, hipStream_t s) {
  hipLaunchKernelGGL(k_transmission, dim3(nblk(n, 256)), dim3(256), 0, s, V, t, n, sigma);
  return hipGetLastError();
}

}  // namespace msl
