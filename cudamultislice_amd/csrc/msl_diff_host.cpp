// msl_diff_host.cpp -- differentiable multislice with slice checkpointing.
//
// Forward: psi_i = IFFT(P . FFT(exp(i sigma V_i) psi_{i-1})) / (nx ny), the
// conventional step of U.cu:2085-2124 with the potential given per slice.
// Backward: adjoint passes in reverse slice order; every K-th intermediate
// wave is checkpointed by the forward and the waves inside a segment are
// recomputed from its checkpoint (the last segment is kept from the forward).
// The passes that produce a segment store its real-space waves psi_{i-1}, so
// the adjoint row pass transforms only the gradient.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/msl.h"
#include "msl_diff.hpp"
#include "msl_internal.hpp"

using namespace msl;

namespace {

struct DErr : std::runtime_error {
  int code;
  DErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void hc(hipError_t e, const char* w) {
  if (e == hipSuccess) return;
  throw DErr(e == hipErrorOutOfMemory ? MSL_ERR_OOM : MSL_ERR_HIP, std::string(w) + ": " + hipGetErrorString(e));
}
#define HC(x) hc((x), #x)

template <class T> struct Buf {
  T* p = nullptr;
  size_t n = 0;
  ~Buf() { if (p) (void)hipFree(p); }
  void alloc(size_t c) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = c;
    if (c) HC(hipMalloc(&p, c * sizeof(T)));
  }
};

}  // namespace

struct msl_diff {
  int device = 0;
  std::string err;
  int nx = 0, ny = 0, ns = 0, B = 0, K = 1, S = 1;
  float ps = 0, kV = 0, dz = 0, wavel = 0, sigma = 0, kmax = 0, nf = 0;
  std::vector<float> hkx, hky;
  Buf<float> kx, ky, partial;
  Buf<float2> twx, twy, px, py, W, G, exitw, ckpt, seg;
  Buf<float2> tall;    // exp(i sigma V_i) of every slice, [ns][ny][nx], written by the forward
  Buf<int2> prow;  // rows of each column inside |k| < kmax
  const float* V = nullptr;
  const float2* psi0 = nullptr;
  bool have_forward = false;
  int ngroups = 1;      // partial gradient maps allocated (largest adjoint group count)
  int chunk = 1;        // probes per adjoint chunk (see msl_diff_backward)
  int prune_tiles = 0;  // band pruning (see the propagation passes)
  bool zb_rows = false;  // rows ny/4 .. 3ny/4 - 1 outside the propagator band
  long long np() const { return (long long)nx * ny; }
};

namespace {

// checkpoint z_j (j a positive multiple of K below ns)
float2* ckpt_slot(msl_diff* d, int j) { return d->ckpt.p + (size_t)(j / d->K - 1) * d->B * d->np(); }
// real-space wave w_i = psi_{i-1} of slice i in the segment starting at base
float2* wslot(msl_diff* d, int i, int base) { return d->seg.p + (size_t)(i - base - 1) * d->B * d->np(); }

RowArgs rargs(msl_diff* d, const float2* src, float2* dst, bool pin = false, bool pout = false) {
  RowArgs a{};
  a.prune_in = pin && d->prune_tiles > 0;
  a.prune_out = pout && d->prune_tiles > 0;
  a.src = src; a.dst = dst; a.src_bstride = a.dst_bstride = d->np();
  a.nx = d->nx; a.ny = d->ny; a.nbatch = d->B; a.t_group = d->B; a.t_bstride = 0; a.scale = 1.f;
  a.tw = d->twx.p; a.sigma = d->sigma;
  return a;
}
ColArgs cargs(msl_diff* d, const float2* src, float2* dst, float scale) {
  ColArgs a{};
  a.src = src; a.dst = dst; a.src_bstride = a.dst_bstride = d->np();
  a.nx = d->nx; a.ny = d->ny; a.nbatch = d->B; a.kx = d->kx.p; a.ky = d->ky.p; a.px = d->px.p; a.py = d->py.p;
  a.kmax = d->kmax; a.scale = scale; a.tw = d->twy.p; a.prop_rows = d->prow.p;
  a.tiles_lo = d->prune_tiles;
  a.zb_rows = d->zb_rows ? 1 : 0;
  return a;
}

// probes per line and workgroups per row of the adjoint row pass for a
// batch of nb probes
// (up to 16 probes per line: one workgroup per row holds up to 64 probes
// at 1024 and writes the gradient rows itself, with no partial maps; on the
// 1024^2 x 100-slice x 64-probe workload fwd+bwd 94.1 ms against 99.5 ms with
// 4 probes per line and partial maps)
void adj_groups(int nx, int nb, bool real_z, int& per_line, int& ngroups) {
  const int lines = adj_lines(nx, real_z);
  per_line = std::max(1, std::min(16, (nb + lines - 1) / lines));
  ngroups = (nb + lines * per_line - 1) / (lines * per_line);
}

// The backward can run the adjoint of a segment chunk by chunk of probes,
// each chunk through all of the segment's slices (msl_diff_set_chunk).  With
// the partial-map reduction chunks of 32 probes were faster (their gradient
// waves stay in the Infinity Cache between passes: 100.2 vs 102.1 ms at
// 1024^2 x 64 probes); with one workgroup per row writing the gradient
// directly the whole batch in one chunk is (94.1 vs 95.7 ms), the default.
int auto_chunk(int nx, int ny, int nprobe) {
  (void)nx;
  (void)ny;
  return nprobe;
}

// transmission exp(i sigma V_i) of slice i (1-based)
float2* tslot(msl_diff* d, int i) { return d->tall.p + (size_t)(i - 1) * d->np(); }

// slice i (1-based) forward: z_{i-1} (psi0 in real space for i = 1) -> z_i
// at dst, through W (dst may be W: both passes work in place).  With wout
// the row pass also stores w_i = IF_x(z_{i-1}) there.  With make_t the
// slice's transmission exp(i sigma V_i) is evaluated once into its slot and
// read by the row pass for every probe (evaluated per line, the same sincosf
// would run once per probe); the backward's recomputation and adjoint passes
// read the slots the forward wrote.
void fwd_slice(msl_diff* d, int i, const float2* zprev, float2* dst, float2* wout, bool make_t, hipStream_t s) {
  if (make_t) HC(launch_transmission(d->V + (size_t)(i - 1) * d->np(), tslot(d, i), d->np(), d->sigma, s));
  RowArgs r = rargs(d, zprev, d->W.p, i > 1, true);
  r.t = tslot(d, i);
  r.t_group = d->B;
  r.t_bstride = 0;
  if (i > 1) r.aux = wout;
  HC(launch_row(i == 1 ? ROW_FWD : ROW_INV_MUL_FWD, RM_T, r, s));
  HC(launch_col(COL_FWD_MUL_INV, CM_PROP, cargs(d, d->W.p, dst, d->nf * d->nf), s));
}

}  // namespace

extern "C" {

msl_status msl_diff_create(msl_diff** out, int device, int nx, int ny, float ps, float kV, float dz, int nslices,
                           int nprobe, int K) {
  if (!out) return MSL_ERR_ARG;
  *out = nullptr;
  if (!fft_size_supported(nx) || !fft_size_supported(ny)) return MSL_ERR_UNSUPPORTED;
  if (!(ps > 0) || !(kV > 0) || !(dz > 0) || nslices < 1 || nprobe < 1) return MSL_ERR_ARG;
  msl_diff* d = new msl_diff;
  try {
    d->device = device;
    HC(hipSetDevice(device));
    d->nx = nx; d->ny = ny; d->ps = ps; d->kV = kV; d->dz = dz; d->ns = nslices; d->B = nprobe;
    if (K < 0) {
      // the longest segment whose waves fit in half the free device memory:
      // longer segments recompute fewer slices in the backward (none when
      // K = nslices); the wave count 3 + (S - 1) + K is smallest near sqrt(ns)
      size_t fr = 0, tot = 0;
      HC(hipMemGetInfo(&fr, &tot));
      const double wave = (double)nprobe * nx * ny * 8.0;
      auto waves = [&](int k) { return 3.0 + (nslices + k - 1) / k - 1 + k; };
      int best = 1;
      for (int k = 1; k <= nslices; ++k)
        if (waves(k) < waves(best)) best = k;
      for (int k = nslices; k >= 1; --k)
        if (waves(k) * wave <= 0.5 * (double)fr) {
          best = std::max(best, k);
          break;
        }
      K = best;
    }
    d->K = std::max(1, std::min(K == 0 ? nslices : K, nslices));
    d->S = (nslices + d->K - 1) / d->K;
    // FP32 physics of U.cu:1791-1799
    const float Pi = 3.1415926f, V = kV, e = 1.6e-019f;
    d->wavel = 6.63e-034f * 3e+008f / std::sqrt((e * V * 1000.f * (2.f * 9.11e-031f * 9e+016f + e * V * 1000.f))) * 1e+010f;
    d->sigma = (2.f * Pi / (d->wavel * V * 1000.f)) * ((9.11e-031f * 9e+016f + e * V * 1000.f) / (2.f * 9.11e-031f * 9e+016f + e * V * 1000.f));
    // frequencies and kmax, U.cu:1877-1953
    const float ksx = ps * (float)nx, ksy = ps * (float)ny;
    const int imx = (int)std::floor(nx / 2 + 0.5), imy = (int)std::floor(ny / 2 + 0.5);
    d->hkx.resize(nx); d->hky.resize(ny);
    for (int i = 1; i <= nx; ++i) d->hkx[i - 1] = ((i - 1) > imx) ? (float)((i - 1) - nx) / ksx : (float)(i - 1) / ksx;
    for (int i = 1; i <= ny; ++i) d->hky[i - 1] = ((i - 1) > imy) ? (float)((i - 1) - ny) / ksy : (float)(i - 1) / ksy;
    const float hx = d->hkx[imx - 1] * 1 / 2, hy = d->hky[imy - 1] * 1 / 2;
    const float kmx = (float)std::pow(hx, 2), kmy = (float)std::pow(hy, 2);
    d->kmax = std::sqrt(kmy <= kmx ? kmy : kmx);
    d->nf = 1.0f / std::sqrt((float)(nx * ny));
    d->prune_tiles = band_tiles_lo(nx);
    for (int x = nx / 4; x < 3 * nx / 4 && d->prune_tiles > 0; ++x)
      if (std::sqrt(d->hkx[x] * d->hkx[x] + 0.0f) < d->kmax) d->prune_tiles = 0;
    d->zb_rows = prop_rows_quarter(d->hky.data(), ny, d->kmax);
    auto tw = [](int n) {
      std::vector<float2> t(n);
      for (int k = 0; k < n; ++k) {
        const double a = -2.0 * M_PI * k / n;
        t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
      return t;
    };
    auto pr = [&](const std::vector<float>& k) {
      std::vector<float2> t(k.size());
      for (size_t i = 0; i < k.size(); ++i) {
        const double a = 3.14159265 * (double)dz * (double)d->wavel * (double)k[i] * (double)k[i];
        t[i] = make_float2((float)std::cos(a), (float)-std::sin(a));
      }
      return t;
    };
    auto up = [&](auto& buf, const auto& v) {
      buf.alloc(v.size());
      HC(hipMemcpy(buf.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
    };
    up(d->kx, d->hkx); up(d->ky, d->hky); up(d->twx, tw(nx)); up(d->twy, tw(ny));
    up(d->px, pr(d->hkx)); up(d->py, pr(d->hky));
    d->prow.alloc(nx);
    HC(make_prop_rows(d->kx.p, d->ky.p, nx, ny, d->kmax, d->prow.p, nullptr));
    const size_t wave = (size_t)nprobe * nx * ny;
    d->W.alloc(wave);
    d->G.alloc(wave);
    d->tall.alloc((size_t)nslices * nx * ny);
    d->exitw.alloc(wave);
    if (d->S > 1) d->ckpt.alloc((size_t)(d->S - 1) * wave);
    d->seg.alloc((size_t)d->K * wave);
    // columns outside the band are never written by pruned passes: keep them
    // finite from the start
    for (Buf<float2>* b : {&d->W, &d->G, &d->exitw, &d->ckpt, &d->seg})
      if (b->p) HC(hipMemset(b->p, 0, b->n * sizeof(float2)));
    // partial maps for the larger group count of the two adjoint variants
    int pl = 1, g0 = 1, g1 = 1;
    adj_groups(nx, nprobe, true, pl, g0);
    adj_groups(nx, nprobe, false, pl, g1);
    d->ngroups = std::max(g0, g1);
    d->partial.alloc((size_t)d->ngroups * nx * ny);
    d->chunk = auto_chunk(nx, ny, nprobe);
  } catch (const DErr& e) {
    const int c = e.code;
    delete d;
    return c;
  }
  *out = d;
  return MSL_OK;
}

void msl_diff_destroy(msl_diff* d) {
  if (!d) return;
  (void)hipSetDevice(d->device);
  (void)hipDeviceSynchronize();  // work queued on callers' streams may still use the buffers
  delete d;
}

const char* msl_diff_last_error(const msl_diff* d) { return d ? d->err.c_str() : "null handle"; }

msl_status msl_diff_set_chunk(msl_diff* d, int probes) {
  if (!d || probes < 0) return MSL_ERR_ARG;
  d->chunk = probes == 0 ? auto_chunk(d->nx, d->ny, d->B) : std::min(probes, d->B);
  return MSL_OK;
}

msl_status msl_diff_info(const msl_diff* d, float* wl, float* sg, float* km, long long* ws) {
  if (!d) return MSL_ERR_ARG;
  if (wl) *wl = d->wavel;
  if (sg) *sg = d->sigma;
  if (km) *km = d->kmax;
  if (ws) {
    const long long w = (long long)d->B * d->np() * 8;
    *ws = w * (3 + (d->S - 1) + d->K) + (long long)d->ngroups * d->np() * 4 + (long long)d->ns * d->np() * 8;
  }
  return MSL_OK;
}

namespace {
// keep = false: intensities only; no checkpoints, no stored real-space waves,
// and the backward refuses until a forward with keep = true
msl_status diff_forward(msl_diff* d, const float* V, const void* psi0, float* inten, void* stream, bool keep);
}  // namespace

msl_status msl_diff_forward(msl_diff* d, const float* V, const void* psi0, float* inten, void* stream) {
  return diff_forward(d, V, psi0, inten, stream, true);
}

msl_status msl_diff_forward_only(msl_diff* d, const float* V, const void* psi0, float* inten, void* stream) {
  return diff_forward(d, V, psi0, inten, stream, false);
}

namespace {
msl_status diff_forward(msl_diff* d, const float* V, const void* psi0, float* inten, void* stream, bool keep) {
  if (!d || !V || !psi0 || !inten) return MSL_ERR_ARG;
  try {
    HC(hipSetDevice(d->device));
    // the caller's stream; NULL is the null stream, ordered with the other
    // default-stream work of the process (as torch's default stream)
    const hipStream_t s = (hipStream_t)stream;
    d->V = V;
    d->psi0 = (const float2*)psi0;
    // z_K, z_2K, ... go to the checkpoints; the last segment's real-space
    // waves w_i stay in seg for the backward pass
    const int last_base = (d->S - 1) * d->K;
    const float2* zprev = d->psi0;
    d->have_forward = false;
    for (int i = 1; i <= d->ns; ++i) {
      float2* dst = (keep && i % d->K == 0 && i < d->ns) ? ckpt_slot(d, i) : d->W.p;
      fwd_slice(d, i, zprev, dst, keep && i > last_base ? wslot(d, i, last_base) : nullptr, true, s);
      zprev = dst;
    }
    HC(launch_row(ROW_INV, RM_NONE, rargs(d, zprev, d->exitw.p, true, false), s));
    HC(launch_intensity(d->exitw.p, inten, (long long)d->B * d->np(), 1.0f, s));
    d->have_forward = keep;
  } catch (const DErr& e) {
    d->err = e.what();
    return e.code;
  }
  d->err.clear();
  return MSL_OK;
}
}  // namespace

msl_status msl_diff_backward(msl_diff* d, const float* dLdI, float* gradV, void* stream) {
  if (!d || !dLdI || !gradV) return MSL_ERR_ARG;
  if (!d->have_forward) {
    d->err = "forward first";
    return MSL_ERR_STATE;
  }
  try {
    HC(hipSetDevice(d->device));
    // the caller's stream; NULL is the null stream, ordered with the other
    // default-stream work of the process (as torch's default stream)
    const hipStream_t s = (hipStream_t)stream;
    const long long np = d->np();
    // G at the exit, 2 dL/dI psi_n, through the adjoint of the final IF_x
    // (one row pass forms and transforms it)
    RowArgs sa = rargs(d, d->exitw.p, d->G.p, false, true);
    sa.seed = dLdI;
    HC(launch_row(ROW_FWD, RM_SEED, sa, s));
    for (int sg = d->S - 1; sg >= 0; --sg) {
      const int base = sg * d->K;
      const int top = std::min(base + d->K, d->ns);
      const bool kept = sg == d->S - 1;  // w_i of the last segment come from the forward
      // z_{top-1} for the adjoint of slice top in recomputed segments
      const float2* ztop = base == 0 ? d->psi0 : ckpt_slot(d, base);
      if (!kept) {
        // recompute z_{base+1} .. z_{top-1} from the checkpoint, storing w_i
        for (int j = base + 1; j < top; ++j) {
          fwd_slice(d, j, ztop, d->W.p, wslot(d, j, base), false, s);
          ztop = d->W.p;
        }
      }
      // the adjoint of the segment, one chunk of probes at a time through
      // all its slices; gradient rows of later chunks are added in chunk
      // order (a fixed summation order)
      for (int c0 = 0; c0 < d->B; c0 += d->chunk) {
        const int nb = std::min(d->chunk, d->B - c0);
        const long long off = (long long)c0 * np;
        for (int i = top; i > base; --i) {
          ColArgs ca = cargs(d, d->G.p + off, d->G.p + off, d->nf * d->nf);
          ca.nbatch = nb;
          HC(launch_col(COL_FWD_MUL_INV, CM_PROPCONJ, ca, s));
          AdjArgs a{};
          a.G = d->G.p + off;
          a.Gout = d->G.p + off;
          bool real_z = true;
          if (i == 1) a.Z = d->psi0 + off;
          else if (kept || i < top) a.Z = wslot(d, i, base) + off;
          else {
            a.Z = ztop + off;  // z_{top-1}: w_top is transformed inside the adjoint pass
            real_z = false;
          }
          int per_line = 1, ngroups = 1;
          adj_groups(d->nx, nb, real_z, per_line, ngroups);
          a.t = tslot(d, i);
          a.sigma = d->sigma;
          a.bstride = np;
          a.nx = d->nx; a.ny = d->ny; a.nbatch = nb;
          a.per_line = per_line;
          a.ngroups = ngroups;
          a.partial = d->partial.p;
          a.grad = gradV + (size_t)(i - 1) * np;
          a.accumulate = c0 > 0 ? 1 : 0;
          a.tw = d->twx.p;
          a.prune = d->prune_tiles > 0;
          HC(launch_row_adjoint(a, real_z, s));
          // one workgroup per row wrote the gradient rows itself
          if (ngroups > 1)
            HC(launch_sum_maps(d->partial.p, ngroups, np, gradV + (size_t)(i - 1) * np, c0 > 0 ? 1 : 0, s));
        }
      }
    }
  } catch (const DErr& e) {
    d->err = e.what();
    return e.code;
  }
  d->err.clear();
  return MSL_OK;
}

}  // extern "C"
