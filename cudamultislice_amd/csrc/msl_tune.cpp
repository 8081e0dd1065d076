// msl_tune.cpp -- pass-layout variants: timed benchmark, validation against
// the reference layout, and the autotuner that picks a variant per length.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/msl.h"
#include "msl_internal.hpp"

using namespace msl;

namespace {

struct Dev {
  void* p = nullptr;
  explicit Dev(size_t bytes) {
    if (hipMalloc(&p, bytes) != hipSuccess) p = nullptr;
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  template <class T> T* as() const { return (T*)p; }
};

struct Bench {
  int n, nb;
  std::vector<float> kx;
  Dev W, t, k, pxy, tw, rows;
  float kmax;
  Bench(int n_, int nb_)
      : n(n_), nb(nb_), W((size_t)nb_ * n_ * n_ * 8), t((size_t)n_ * n_ * 8), k((size_t)n_ * 4),
        pxy((size_t)n_ * 8), tw((size_t)n_ * 8), rows((size_t)n_ * 8) {}
  bool ok() const { return W.p && t.p && k.p && pxy.p && tw.p && rows.p; }
  hipError_t init(hipStream_t s) {
    // frequencies / propagator of a 0.1 A grid at 300 kV, 2 A slices
    std::vector<float> kk(n);
    std::vector<float2> p(n), w(n);
    const float ks = 0.1f * n;
    for (int i = 0; i < n; ++i) kk[i] = (i <= n / 2 ? (float)i : (float)(i - n)) / ks;
    kmax = 0.5f * (n / 2 - 1) / ks;
    kx = kk;
    for (int i = 0; i < n; ++i) {
      const double a = 3.14159265 * 2.0 * 0.0197 * (double)kk[i] * kk[i];
      p[i] = make_float2((float)std::cos(a), (float)-std::sin(a));
      const double b = -2.0 * M_PI * i / n;
      w[i] = make_float2((float)std::cos(b), (float)std::sin(b));
    }
    hipError_t e;
    if ((e = hipMemcpyAsync(k.p, kk.data(), n * 4, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(pxy.p, p.data(), n * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(tw.p, w.data(), n * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = launch_hash_fill(t.as<float2>(), (long long)n * n, 7u, 1, s)) != hipSuccess) return e;
    if ((e = make_prop_rows(k.as<float>(), k.as<float>(), n, n, kmax, rows.as<int2>(), s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
  hipError_t fill(float2* w, int nw, hipStream_t s) { return launch_hash_fill(w, (long long)nw * n * n, 11u, 0, s); }
  // kind 0: row IFFT * t * FFT (scaled 1/n); kind 1: column FFT * P / n^2 * IFFT;
  // band-pruned as in a multislice run wherever the length allows it
  hipError_t run(int kind, int var, float2* w, int nw, hipStream_t s) {
    const long long bs = (long long)n * n;
    const int tiles_lo = band_tiles_lo(n);
    if (kind == 0) {
      RowArgs a{};
      a.prune_in = a.prune_out = tiles_lo > 0;
      a.src = w; a.dst = w; a.src_bstride = a.dst_bstride = bs; a.nx = n; a.ny = n; a.nbatch = nw;
      a.t = t.as<float2>(); a.t_group = nw; a.t_bstride = 0; a.scale = 1.0f / n; a.tw = tw.as<float2>();
      return launch_row_v(ROW_INV_MUL_FWD, RM_T, a, s, var);
    }
    ColArgs c{};
    c.src = w; c.dst = w; c.src_bstride = c.dst_bstride = bs; c.nx = n; c.ny = n; c.nbatch = nw;
    c.kx = k.as<float>(); c.ky = k.as<float>(); c.px = pxy.as<float2>(); c.py = pxy.as<float2>();
    c.kmax = kmax; c.scale = 1.0f / ((float)n * n); c.tw = tw.as<float2>(); c.prop_rows = rows.as<int2>();
    c.tiles_lo = tiles_lo;
    c.zb_rows = prop_rows_quarter(kx.data(), n, kmax) ? 1 : 0;
    return launch_col_v(COL_FWD_MUL_INV, CM_PROP, c, s, var);
  }
};

}  // namespace

extern "C" {

msl_status msl_copy_device(void* dst, const void* src, long long bytes, void* stream) {
  if (!dst || !src || bytes <= 0 || bytes % 16) return MSL_ERR_ARG;
  if (launch_copy16(src, dst, bytes, (hipStream_t)stream) != hipSuccess) return MSL_ERR_HIP;
  return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? MSL_OK : MSL_ERR_HIP;
}

msl_status msl_set_pass_variant(int kind, int n, int variant) {
  if ((kind != 0 && kind != 1) || !fft_size_supported(n) || variant < 0 || variant >= num_pass_variants(kind))
    return MSL_ERR_ARG;
  set_pass_variant(kind, n, variant);
  return MSL_OK;
}

msl_status msl_get_pass_variant(int kind, int n, int* variant) {
  if ((kind != 0 && kind != 1) || !fft_size_supported(n) || !variant) return MSL_ERR_ARG;
  *variant = pass_variant(kind, n);
  return MSL_OK;
}

// Times `reps` launches of one pass variant on an n x n batch of nbatch
// waves (HIP events on a private stream) and, when maxdiff is given, checks
// the variant against variant 0 on two waves (max |diff| / max |ref|).
msl_status msl_bench_pass(int kind, int variant, int n, int nbatch, int reps, float* ms, float* maxdiff) {
  if ((kind != 0 && kind != 1) || !fft_size_supported(n) || nbatch < 1 || reps < 1 || !ms) return MSL_ERR_ARG;
  if (variant < 0 || variant >= num_pass_variants(kind)) return MSL_ERR_ARG;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MSL_ERR_HIP;
  msl_status st = MSL_OK;
  {
    Bench b(n, nbatch);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!b.ok()) st = MSL_ERR_OOM;
    else if (b.init(s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = MSL_ERR_HIP;
    if (st == MSL_OK && maxdiff) {
      // enough waves for the layouts that need several per workgroup
      const int nv = std::min(n <= 1024 ? 16 : 2, nbatch);
      const size_t cnt = (size_t)nv * n * n;
      Dev A(cnt * 8), B(cnt * 8);
      std::vector<float2> ha(cnt), hb(cnt);
      if (!A.p || !B.p || b.fill(A.as<float2>(), nv, s) != hipSuccess || b.fill(B.as<float2>(), nv, s) != hipSuccess ||
          b.run(kind, 0, A.as<float2>(), nv, s) != hipSuccess || b.run(kind, variant, B.as<float2>(), nv, s) != hipSuccess ||
          hipMemcpyAsync(ha.data(), A.p, cnt * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipMemcpyAsync(hb.data(), B.p, cnt * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        st = MSL_ERR_HIP;
      } else {
        double md = 0, mr = 0;
        for (size_t i = 0; i < cnt; ++i) {
          md = std::max(md, (double)std::hypot(ha[i].x - hb[i].x, ha[i].y - hb[i].y));
          mr = std::max(mr, (double)std::hypot(ha[i].x, ha[i].y));
        }
        *maxdiff = (float)(mr > 0 ? md / mr : md);
      }
    }
    if (st == MSL_OK) {
      float2* w = b.W.as<float2>();
      if (b.fill(w, nbatch, s) != hipSuccess || b.run(kind, variant, w, nbatch, s) != hipSuccess ||
          hipEventRecord(e0, s) != hipSuccess) {
        st = MSL_ERR_HIP;
      } else {
        for (int r = 0; r < reps && st == MSL_OK; ++r)
          if (b.run(kind, variant, w, nbatch, s) != hipSuccess) st = MSL_ERR_HIP;
        float t = 0;
        if (st == MSL_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&t, e0, e1) != hipSuccess))
          st = MSL_ERR_HIP;
        *ms = t / reps;
      }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  (void)hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  return st;
}

// Autotuner: for length n and a batch of nbatch waves, time every row and
// column variant that agrees with variant 0 to 1e-5 and select the fastest.
// out (optional): [0..7] row ms per variant, [8..15] column ms per variant,
// -1 for variants rejected or not built.
msl_status msl_autotune_passes(int n, int nbatch, int reps, int* row_var, int* col_var, float* out) {
  if (!fft_size_supported(n) || nbatch < 1 || reps < 1) return MSL_ERR_ARG;
  int best[2] = {0, 0};
  if (out)
    for (int i = 0; i < 16; ++i) out[i] = -1.f;
  for (int kind = 0; kind < 2; ++kind) {
    float bms = 1e30f;
    for (int v = 0; v < num_pass_variants(kind); ++v) {
      float ms = -1, md = 0;
      msl_status st = msl_bench_pass(kind, v, n, nbatch, reps, &ms, &md);
      if (st != MSL_OK) return st;
      if (!(md < 1e-5f)) ms = -1;
      if (out && v < 8) out[8 * kind + v] = ms;
      if (ms > 0 && ms < bms) {
        bms = ms;
        best[kind] = v;
      }
    }
    set_pass_variant(kind, n, best[kind]);
  }
  if (row_var) *row_var = best[0];
  if (col_var) *col_var = best[1];
  return MSL_OK;
}

}  // extern "C"
