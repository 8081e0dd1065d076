// msl_host.cpp -- host side of libmsl: the simulation handle behind the
// C-ABI in include/msl.h.  Mirrors the state and call sequence of the
// reference's UnmanagedMultisliceSimulation (U.cuh:62-199, U.cu), with
// per-handle device memory and stream instead of process-wide __constant__
// globals, and status codes instead of exit(1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/msl.h"
#include "fparams_table.h"
#include "msl_internal.hpp"

#include "msl_host.hpp"

using namespace msl;
using namespace mslh;

namespace mslh {

bool pow2_ok(int n) { return fft_size_supported(n); }
// any side from 2 to 8192 (the generic transforms take every length)
bool grid_ok(int n) { return n >= 2 && n <= 8192; }

// complete symbol table, Z = 1..103
const char* kSymbols[103] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
    "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe",
    "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
    "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
    "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr"};

// GetZNum, U.cu:1188-1243: 25 symbols, everything else is hydrogen
int reference_znum(const std::string& s) {
  static const char* sym[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Na", "Mg", "Si", "P",
                              "S", "Cl", "Ca", "Cr", "Br", "Fe", "Sr", "Ru", "La", "Sm", "Ta", "W"};
  static const int z[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 14, 15, 16, 17, 20, 24, 35, 26, 38, 44, 57, 62, 73, 74};
  for (int i = 0; i < 25; ++i)
    if (s == sym[i]) return z[i];
  return 1;
}

int full_znum(const std::string& s) {
  for (int i = 0; i < 103; ++i)
    if (s == kSymbols[i]) return i + 1;
  return 0;
}

// Philox4x64-10 (Salmon et al., SC'11) and the Box-Muller mapping used for
// frozen-phonon displacements: atom i, axis a of configuration c uses block
// counter (i, c, a, 0) under key (seed, 0x4D534C).
void philox4x64(const uint64_t ci[4], const uint64_t ki[2], uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  uint64_t c0 = ci[0], c1 = ci[1], c2 = ci[2], c3 = ci[3], k0 = ki[0], k1 = ki[1];
  for (int r = 0; r < 10; ++r) {
    const unsigned __int128 p0 = (unsigned __int128)M0 * c0, p1 = (unsigned __int128)M1 * c2;
    const uint64_t n0 = (uint64_t)(p1 >> 64) ^ c1 ^ k0, n1 = (uint64_t)p1;
    const uint64_t n2 = (uint64_t)(p0 >> 64) ^ c3 ^ k1, n3 = (uint64_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double philox_normal(uint64_t seed, uint64_t config, uint64_t atom, uint64_t axis) {
  const uint64_t ctr[4] = {atom, config, axis, 0}, key[2] = {seed, 0x4D534CULL};
  uint64_t o[4];
  philox4x64(ctr, key, o);
  const double u1 = ((double)(o[0] >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(o[1] >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

void tds_disp(uint64_t seed, int config, int n, float sigma, float* dx, float* dy, float* dz) {
  for (int i = 0; i < n; ++i) {
    dx[i] = (float)((double)sigma * philox_normal(seed, (uint64_t)config, (uint64_t)i, 0));
    dy[i] = (float)((double)sigma * philox_normal(seed, (uint64_t)config, (uint64_t)i, 1));
    dz[i] = (float)((double)sigma * philox_normal(seed, (uint64_t)config, (uint64_t)i, 2));
  }
}

std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  return t;
}

}  // namespace mslh

namespace mslh {

void set_device(msl_sim* s) { HIPC(hipSetDevice(s->device)); }

// every transfer is ordered on the handle's (non-blocking) stream and waited
// for, so it sees all kernels queued before it
void d2h(msl_sim* s, void* dst, const void* src, size_t bytes) {
  HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
}
void h2d(msl_sim* s, void* dst, const void* src, size_t bytes) {
  HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
}

// ------------------------------------------------------------- pass helpers
RowArgs row_args(msl_sim* s, const float2* src, float2* dst, int nbatch, long long bs) {
  RowArgs a{};
  a.src = src;
  a.dst = dst;
  a.src_bstride = a.dst_bstride = bs;
  a.nx = s->resX;
  a.ny = s->resY;
  a.nbatch = nbatch;
  a.t_group = 1;
  a.scale = 1.0f;
  a.tw = s->twx.p;
  return a;
}

ColArgs col_args(msl_sim* s, const float2* src, float2* dst, int nbatch, long long bs) {
  ColArgs a{};
  a.src = src;
  a.dst = dst;
  a.src_bstride = a.dst_bstride = bs;
  a.nx = s->resX;
  a.ny = s->resY;
  a.nbatch = nbatch;
  a.kx = s->kx.p;
  a.ky = s->ky.p;
  a.px = s->px.p;
  a.py = s->py.p;
  a.prop_rows = s->prow.p;
  a.kmax = s->kmax;
  a.scale = 1.0f;
  a.tw = s->twy.p;
  a.zb_rows = s->zb_rows ? 1 : 0;
  return a;
}

void row(msl_sim* s, int mode, int mul, const RowArgs& a) { HIPC(launch_row(mode, mul, a, s->stream)); }
void col(msl_sim* s, int mode, int mul, const ColArgs& a) { HIPC(launch_col(mode, mul, a, s->stream)); }

// separable Fresnel factors exp(-i pi dz lambda k^2) per axis (CreatePropsKernel,
// U.cu:399-428: cos / -sin of pi dz lambda (kx^2 + ky^2))
void ensure_prop(msl_sim* s, float dz) {
  if (s->prop_dz == dz && s->px.p) return;
  auto tab = [&](const std::vector<float>& k) {
    std::vector<float2> t(k.size());
    for (size_t i = 0; i < k.size(); ++i) {
      const double a = 3.14159265 * (double)dz * (double)s->wavel * (double)k[i] * (double)k[i];
      t[i] = make_float2((float)std::cos(a), (float)-std::sin(a));
    }
    return t;
  };
  auto tx = tab(s->hkx), ty = tab(s->hky);
  s->px.ensure(tx.size());
  s->py.ensure(ty.size());
  HIPC(hipMemcpyAsync(s->px.p, tx.data(), tx.size() * sizeof(float2), hipMemcpyHostToDevice, s->stream));
  HIPC(hipMemcpyAsync(s->py.p, ty.data(), ty.size() * sizeof(float2), hipMemcpyHostToDevice, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  s->prop_dz = dz;
}

// generic transform of nbatch waves in place, then the scale
void gen_fft2(msl_sim* s, float2* w, int nbatch, bool inverse, float scale) {
  s->gscr.ensure((size_t)nbatch * s->npix());
  HIPC(launch_gen_fft2(w, s->gscr.p, s->resX, s->resY, nbatch, inverse ? 1 : 0, s->twx.p, s->twy.p, s->stream));
  if (scale != 1.0f)
    HIPC(launch_gen_kmul(w, s->resX, s->resY, nbatch, CM_SCALE, nullptr, nullptr, nullptr, nullptr, 0.f, scale,
                         s->stream));
}
// w *= (k-space multiplier mul) * scale, one wave (generic grids)
void gen_kmul(msl_sim* s, float2* w, int mul, float scale) {
  HIPC(launch_gen_kmul(w, s->resX, s->resY, 1, mul, s->kx.p, s->ky.p, s->px.p, s->py.p, s->kmax, scale, s->stream));
}
void copy_wave(msl_sim* s, float2* dst, const float2* src) {
  if (dst != src)
    HIPC(hipMemcpyAsync(dst, src, s->npix() * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
}

// in-place unnormalised 2D transforms of nbatch waves
void fft2(msl_sim* s, float2* w, int nbatch, bool inverse, float scale = 1.0f) {
  if (s->generic) {
    gen_fft2(s, w, nbatch, inverse, scale);
    return;
  }
  const long long bs = (long long)s->npix();
  if (!inverse) {
    row(s, ROW_FWD, RM_NONE, row_args(s, w, w, nbatch, bs));
    ColArgs c = col_args(s, w, w, nbatch, bs);
    c.scale = scale;
    col(s, COL_FWD, CM_SCALE, c);
  } else {
    ColArgs c = col_args(s, w, w, nbatch, bs);
    c.scale = scale;
    col(s, COL_INV, CM_SCALE, c);
    row(s, ROW_INV, RM_NONE, row_args(s, w, w, nbatch, bs));
  }
}

// w <- IFFT(P(dz) FFT(w)) nf^2, the propagation of U.cu:1986-1993
void propagate(msl_sim* s, float2* src, float2* dst, float dz) {
  ensure_prop(s, dz);
  if (s->generic) {
    copy_wave(s, dst, src);
    gen_fft2(s, dst, 1, false, 1.0f);
    gen_kmul(s, dst, CM_PROP, s->nf * s->nf);
    gen_fft2(s, dst, 1, true, 1.0f);
    return;
  }
  const long long bs = (long long)s->npix();
  row(s, ROW_FWD, RM_NONE, row_args(s, src, dst, 1, bs));
  ColArgs c = col_args(s, dst, dst, 1, bs);
  c.scale = s->nf * s->nf;
  col(s, COL_FWD_MUL_INV, CM_PROP, c);
  row(s, ROW_INV, RM_NONE, row_args(s, dst, dst, 1, bs));
}

void ensure_ztab(msl_sim* s) {
  if (s->ztab.p) return;
  std::vector<float> zt(103 * 24);
  kirkland_transform(s->fparams.data(), zt.data());
  s->ztab.ensure(zt.size());
  HIPC(hipMemcpyAsync(s->ztab.p, zt.data(), zt.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
}

// LoadAtomFile extents, U.cu:1323-1355
void set_atoms(msl_sim* s, std::vector<int>&& Z, std::vector<float>&& x, std::vector<float>&& y,
               std::vector<float>&& z) {
  require(!x.empty(), MSL_ERR_ARG, "no atoms");
  for (int v : Z) require(v >= 1 && v <= 103, MSL_ERR_ARG, "atomic number outside 1..103");
  size_t mx = 0, nx = 0, my = 0, ny = 0, mz = 0, nz = 0;
  for (size_t i = 1; i < x.size(); ++i) {
    if (x[i] > x[mx]) mx = i;
    if (y[i] > y[my]) my = i;
    if (z[i] > z[mz]) mz = i;
    if (x[i] < x[nx]) nx = i;
    if (y[i] < y[ny]) ny = i;
    if (z[i] < z[nz]) nz = i;
  }
  s->MaxX = x[mx] + 2; s->MinX = x[nx] - 2;
  s->MaxY = y[my] + 2; s->MinY = y[ny] - 2;
  s->MaxZ = z[mz] + 2; s->MinZ = z[nz] - 2;
  s->Z = std::move(Z); s->x = std::move(x); s->y = std::move(y); s->z = std::move(z);
  s->have_atoms = true;
  s->binned = false;
  s->tds_config = 0;
}

// UploadBinnedAtoms / UploadBinnedAtomsTDS (U.cu:1358-1716): FP32 bin ids on
// the device, stable counting sort into the (slab, y, x) CSR.
void bin_atoms(msl_sim* s, const float* dx, const float* dy, const float* dz) {
  require(s->have_atoms, MSL_ERR_STATE, "no atoms loaded");
  const int n = (int)s->x.size();
  std::vector<float> X(n), Y(n), Zp(n);
  for (int i = 0; i < n; ++i) {
    X[i] = s->x[i] + (dx ? dx[i] : 0.0f);
    Y[i] = s->y[i] + (dy ? dy[i] : 0.0f);
    Zp[i] = s->z[i] + (dz ? dz[i] : 0.0f);
  }
  int nslab = (int)std::ceil((s->MaxZ - s->MinZ) / 1.0f);
  nslab += (nslab == 0);
  const int nb = nslab * 2500;
  s->in_x.ensure(n); s->in_y.ensure(n); s->in_z.ensure(n); s->in_Z.ensure(n);
  s->bx.ensure(n); s->by.ensure(n); s->bz.ensure(n); s->bZ.ensure(n);
  s->key.ensure(n); s->perm.ensure(n);
  s->count.ensure(nb + 1); s->fill.ensure(nb); s->start.ensure(nb + 1); s->bad.ensure(1);
  s->scan_tmp.ensure((nb + 1 + 1023) / 1024 + 1);
  hipStream_t st = s->stream;
  HIPC(hipMemcpyAsync(s->in_x.p, X.data(), n * sizeof(float), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(s->in_y.p, Y.data(), n * sizeof(float), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(s->in_z.p, Zp.data(), n * sizeof(float), hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(s->in_Z.p, s->Z.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
  HIPC(hipMemsetAsync(s->count.p, 0, (nb + 1) * sizeof(int), st));
  HIPC(hipMemsetAsync(s->fill.p, 0, nb * sizeof(int), st));
  HIPC(hipMemsetAsync(s->bad.p, 0, sizeof(int), st));
  BinArgs a{};
  a.x = s->in_x.p; a.y = s->in_y.p; a.z = s->in_z.p; a.Z = s->in_Z.p; a.n = n;
  a.MaxX = s->MaxX; a.MinX = s->MinX; a.MaxY = s->MaxY; a.MinY = s->MinY; a.MaxZ = s->MaxZ; a.MinZ = s->MinZ;
  a.nslab = nslab;
  a.key = s->key.p; a.count = s->count.p; a.start = s->start.p; a.fill = s->fill.p; a.perm = s->perm.p;
  a.bx = s->bx.p; a.by = s->by.p; a.bz = s->bz.p; a.bZ = s->bZ.p; a.bad = s->bad.p; a.scan_tmp = s->scan_tmp.p;
  HIPC(launch_bin(a, st));
  int bad = 0;
  HIPC(hipMemcpyAsync(&bad, s->bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  require(!bad, MSL_ERR_RANGE, "an atom lies outside the 50x50 bin grid (displacement larger than the 2 A margin?)");
  s->nslab = nslab;
  s->nbinned = n;
  s->binned = true;
}

void calc_vars(msl_sim* s, float ps, float df, int rx, int ry, float sizeZ, float ox, float oy, bool stem) {
  require(s->have_micro, MSL_ERR_STATE, "ApplyMicroscopeParameters first");
  require(grid_ok(rx) && grid_ok(ry), MSL_ERR_UNSUPPORTED, "resolution must be 2..8192 per side");
  // powers of two 16..2048 run the fused LDS passes, other sizes the generic ones
  s->generic = !(pow2_ok(rx) && pow2_ok(ry));
  require(ps > 0, MSL_ERR_ARG, "pixel scale must be positive");
  s->resX = rx; s->resY = ry; s->df = df; s->ps = ps;
  s->offx = stem ? 0.f : ox; s->offy = stem ? 0.f : oy;
  s->ksizex = ps * (float)rx;
  s->ksizey = ps * (float)ry;
  // FP32 physics in the operation order of U.cu:1791-1834
  const float Pi = 3.1415926f, V = s->kV, echarge = 1.6e-019f;
  const float wl = 6.63e-034f * 3e+008f / std::sqrt((echarge * V * 1000.f * (2.f * 9.11e-031f * 9e+016f + echarge * V * 1000.f))) * 1e+010f;
  const float s2 = (2.f * Pi / (wl * V * 1000.f)) * ((9.11e-031f * 9e+016f + echarge * V * 1000.f) / (2.f * 9.11e-031f * 9e+016f + echarge * V * 1000.f));
  s->wavel = wl;
  s->sigma2 = s2;
  const float V2 = V * 1000.f;
  const float p1 = (float)rx / (2.f * s->ksizex), p2 = (float)ry / (2.f * s->ksizey);
  const float ke2 = 0.666666f * (p1 * p1 + p2 * p2);
  const float qA = (ke2 * ke2 * 16.f * Pi * Pi * Pi * Pi) - (32.f * Pi * Pi * Pi * ke2 * s2 * V2 / wl) + (16.f * Pi * Pi * s2 * s2 * V2 * V2 / (wl * wl));
  const float qB = 16.f * Pi * Pi * (ke2 - (s2 * V2 / (Pi * wl)) - (1.f / (4.f * wl * wl)));
  const float qC = 3.f;
  const float disc = qB * qB - 4.f * qA * qC;
  float msd = (-qB + std::sqrt(disc)) / (2.f * qA);
  msd = (float)(0.99 * (double)std::sqrt(msd));
  if (stem && msd > 0.06) msd = 0.06f;
  int nFD = (int)std::ceil(sizeZ / msd);
  nFD += (nFD == 0);
  s->numberOfSlices = nFD;
  s->dzFD = msd;
  if (s->MaxX > s->MinX && s->MaxY > s->MinY) {
    s->loadblocksx = (int)std::ceil(std::sqrt(5.0f) / ((s->MaxX - s->MinX) / 50.f));
    s->loadblocksy = (int)std::ceil(std::sqrt(5.0f) / ((s->MaxY - s->MinY) / 50.f));
  }
  s->loadslicesz = (int)std::ceil(std::sqrt(5.0f) / 1.0f);
  s->have_calc = true;
  s->have_freqs = false;
  s->have_waves = false;
}

// PreCalculateFrequencies, U.cu:1877-1964
void precalc(msl_sim* s) {
  require(s->have_calc, MSL_ERR_STATE, "SetCalculationVariables first");
  const int rx = s->resX, ry = s->resY;
  const int imidx = (int)std::floor(rx / 2 + 0.5), imidy = (int)std::floor(ry / 2 + 0.5);
  s->hkx.resize(rx);
  s->hky.resize(ry);
  for (int i = 1; i <= rx; ++i) s->hkx[i - 1] = ((i - 1) > imidx) ? (float)((i - 1) - rx) / s->ksizex : (float)(i - 1) / s->ksizex;
  for (int i = 1; i <= ry; ++i) s->hky[i - 1] = ((i - 1) > imidy) ? (float)((i - 1) - ry) / s->ksizey : (float)(i - 1) / s->ksizey;
  const float hx = s->hkx[imidx - 1] * 1 / 2, hy = s->hky[imidy - 1] * 1 / 2;
  const float kmx = (float)std::pow(hx, 2), kmy = (float)std::pow(hy, 2);
  s->kmax = std::sqrt((kmy <= kmx) ? kmy : kmx);
  // compat=physical: 2/3 of the Nyquist frequency
  if (s->compat & MSL_COMPAT_PHYSICAL) s->kmax = (2.0f / 3.0f) * std::min(s->hkx[imidx], s->hky[imidy]);
  s->nf = 1.0f / std::sqrt((float)(rx * ry));
  // band pruning is exact when every column N/4 <= x < 3N/4 fails the
  // propagator's |k| < kmax test for any ky (then so does the band limit's)
  s->prune_tiles = band_tiles_lo(rx);
  for (int x = rx / 4; x < 3 * rx / 4 && s->prune_tiles > 0; ++x)
    if (std::sqrt(s->hkx[x] * s->hkx[x] + 0.0f) < s->kmax) s->prune_tiles = 0;
  s->zb_rows = prop_rows_quarter(s->hky.data(), ry, s->kmax);
  s->kx.ensure(rx);
  s->ky.ensure(ry);
  HIPC(hipMemcpyAsync(s->kx.p, s->hkx.data(), rx * sizeof(float), hipMemcpyHostToDevice, s->stream));
  HIPC(hipMemcpyAsync(s->ky.p, s->hky.data(), ry * sizeof(float), hipMemcpyHostToDevice, s->stream));
  s->prow.ensure(rx);
  HIPC(make_prop_rows(s->kx.p, s->ky.p, rx, ry, s->kmax, s->prow.p, s->stream));
  auto tx = twiddles(rx), ty = twiddles(ry);
  s->twx.ensure(rx);
  s->twy.ensure(ry);
  HIPC(hipMemcpyAsync(s->twx.p, tx.data(), rx * sizeof(float2), hipMemcpyHostToDevice, s->stream));
  HIPC(hipMemcpyAsync(s->twy.p, ty.data(), ry * sizeof(float2), hipMemcpyHostToDevice, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  s->prop_dz = std::nanf("");
  s->have_freqs = true;
}

void alloc_waves(msl_sim* s) {
  const size_t n = s->npix();
  s->PsiMinus.ensure(n); s->Psi.ensure(n); s->PsiPlus.ensure(n);
  s->V.ensure(n); s->G.ensure(n); s->B.ensure(n);
  s->red.ensure(4096);
}

void potential(msl_sim* s, int mode, int first, int nslices, float dz, float zplane, int out_kind, void* out,
               long long sstride) {
  const PotArgs p = pot_args(s, mode, first, nslices, dz, zplane, out_kind, out, sstride);
  HIPC(launch_potential(p, s->stream));
}

PotArgs pot_args(msl_sim* s, int mode, int first, int nslices, float dz, float zplane, int out_kind, void* out,
                 long long sstride) {
  require(s->binned, MSL_ERR_STATE, "UploadBinnedAtoms first");
  ensure_ztab(s);
  PotArgs p{};
  p.bx = s->bx.p; p.by = s->by.p; p.bz = s->bz.p; p.bZ = s->bZ.p; p.start = s->start.p; p.nslab = s->nslab;
  p.ztab = s->ztab.p;
  p.MaxX = s->MaxX; p.MinX = s->MinX; p.MaxY = s->MaxY; p.MinY = s->MinY; p.MaxZ = s->MaxZ; p.MinZ = s->MinZ;
  p.nx = s->resX; p.ny = s->resY; p.ps = s->ps; p.offx = s->offx; p.offy = s->offy;
  p.dz = dz; p.first_slice = first; p.nslices = nslices; p.zplane = zplane;
  p.loadslicesz = s->loadslicesz; p.sigma = s->sigma2; p.mode = mode; p.out_kind = out_kind;
  p.physical = (s->compat & MSL_COMPAT_PHYSICAL) ? 1 : 0;
  p.tile = s->pot_tile;
  p.out = out; p.out_sstride = sstride;
  return p;
}

// host argmax / argmin of |psi| with first-occurrence ties (cublasIsamax /
// Isamin semantics, U.cu:2261-2262)
void extrema(const std::vector<float2>& w, float& mx, float& mn) {
  size_t imx = 0, imn = 0;
  float vmx = -1, vmn = 3.4e38f;
  for (size_t i = 0; i < w.size(); ++i) {
    const float a = std::sqrt(w[i].x * w[i].x + w[i].y * w[i].y);
    if (a > vmx) { vmx = a; imx = i; }
    if (a < vmn) { vmn = a; imn = i; }
  }
  mx = std::sqrt(w[imx].x * w[imx].x + w[imx].y * w[imx].y);
  mn = std::sqrt(w[imn].x * w[imn].x + w[imn].y * w[imn].y);
}

// STEM probe of U.cu:2672-2705 for one position: k-space synthesis,
// unnormalised inverse FFT, division by sum |psi|
void make_probe(msl_sim* s, float2* out, int posx, int posy) {
  ProbeArgs p{};
  p.out = out; p.bstride = (long long)s->npix(); p.nx = s->resX; p.ny = s->resY; p.nprobe = 1;
  p.kx = s->kx.p; p.ky = s->ky.p; p.px0 = posx; p.py0 = posy;
  p.apert = (float)((double)s->obj * 0.001 / (double)s->wavel);
  p.ps = s->ps; p.df = s->df; p.Cs = s->Cs; p.wavel = s->wavel;
  HIPC(launch_probe_kspace(p, s->stream));
  fft2(s, out, 1, true);
  HIPC(launch_probe_normalise(out, (long long)s->npix(), 1, (long long)s->npix(), s->red.p,
                              (s->compat & MSL_COMPAT_PHYSICAL) ? 1 : 0, s->stream));
}

float download_scalar(msl_sim* s, const float* d) {
  float v = 0;
  HIPC(hipMemcpyAsync(&v, d, sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  return v;
}

// Event pairs around launches of one chunk, summed per kernel kind once the
// chunk has finished.  kind: 0 row pass, 1 column pass, 2 detector pass,
// 3 potential.
struct KTimer {
  msl_sim* s;
  std::vector<int> kinds;
  explicit KTimer(msl_sim* s_) : s(s_) {}
  void begin(int kind) {
    if (!s->profiling) return;
    const size_t i = kinds.size() * 2;
    while (s->events.size() < i + 2) {
      hipEvent_t e;
      HIPC(hipEventCreate(&e));
      s->events.push_back(e);
    }
    HIPC(hipEventRecord(s->events[i], s->stream));
    kinds.push_back(kind);
  }
  void end() {
    if (!s->profiling) return;
    HIPC(hipEventRecord(s->events[kinds.size() * 2 - 1], s->stream));
  }
  void flush() {
    if (!s->profiling || kinds.empty()) return;
    HIPC(hipStreamSynchronize(s->stream));
    for (size_t k = 0; k < kinds.size(); ++k) {
      float ms = 0;
      HIPC(hipEventElapsedTime(&ms, s->events[2 * k], s->events[2 * k + 1]));
      s->kstats[2 * kinds[k]] += ms;
      s->kstats[2 * kinds[k] + 1] += 1;
    }
    kinds.clear();
  }
};

// annular |Psi_unitary|^2 of one real-space wave (copied into scratch)
float adf_of(msl_sim* s, const float2* w, float inner, float outer) {
  const long long bs = (long long)s->npix();
  if (s->generic) {
    copy_wave(s, s->B.p, w);
    gen_fft2(s, s->B.p, 1, false, 1.0f);
    s->gpart.ensure(gen_annulus_blocks(bs));
    HIPC(launch_gen_annulus(s->B.p, s->resX, s->resY, s->kx.p, s->ky.p, inner * 1e-3f / s->wavel,
                            outer * 1e-3f / s->wavel, s->nf * s->nf, s->gpart.p, s->red.p + 2048, s->stream));
    return download_scalar(s, s->red.p + 2048);
  }
  row(s, ROW_FWD, RM_NONE, row_args(s, w, s->B.p, 1, bs));
  ColArgs c = col_args(s, s->B.p, nullptr, 1, bs);
  c.kin = inner * 1e-3f / s->wavel;
  c.kout = outer * 1e-3f / s->wavel;
  c.det_scale = s->nf * s->nf;
  c.partial = s->red.p;
  col(s, COL_FWD_DETECT, CM_SCALE, c);
  const int ntiles = s->resX / col_tile(s->resY);
  HIPC(launch_reduce_rows(s->red.p, 1, ntiles, s->red.p + 2048, 0, s->stream));
  return download_scalar(s, s->red.p + 2048);
}

// the finite-difference step of U.cu:2050-2083 on the fused passes (grids of
// powers of two): band limit of V, Laplacian of psi, update, band limit of psi+
void fd_step_passes(msl_sim* s) {
  const long long bs = (long long)s->npix();
  // band-limit V: FFT, |k| <= kmax, IFFT, * nf^2
  row(s, ROW_FWD, RM_NONE, row_args(s, s->V.p, s->B.p, 1, bs));
  ColArgs c = col_args(s, s->B.p, s->B.p, 1, bs);
  c.scale = s->nf * s->nf;
  col(s, COL_FWD_MUL_INV, CM_BAND, c);
  row(s, ROW_INV, RM_NONE, row_args(s, s->B.p, s->V.p, 1, bs));
  // spectral Laplacian: FFT, * nf (-4 pi^2 k^2), IFFT, * nf
  row(s, ROW_FWD, RM_NONE, row_args(s, s->Psi.p, s->G.p, 1, bs));
  c = col_args(s, s->G.p, s->G.p, 1, bs);
  c.scale = s->nf * s->nf;
  col(s, COL_FWD_MUL_INV, CM_GRAD, c);
  row(s, ROW_INV, RM_NONE, row_args(s, s->G.p, s->G.p, 1, bs));
  HIPC(launch_fd_update(s->G.p, s->PsiMinus.p, s->PsiPlus.p, s->Psi.p, s->V.p, bs, bs, s->sigma2, s->wavel, s->dzFD,
                        s->stream));
  // band-limit psi+
  row(s, ROW_FWD, RM_NONE, row_args(s, s->PsiPlus.p, s->B.p, 1, bs));
  c = col_args(s, s->B.p, s->B.p, 1, bs);
  c.scale = s->nf * s->nf;
  col(s, COL_FWD_MUL_INV, CM_BAND, c);
  row(s, ROW_INV, RM_NONE, row_args(s, s->B.p, s->PsiPlus.p, 1, bs));
  // rotate: psi- <- psi, psi <- psi+
  std::swap(s->PsiMinus.p, s->Psi.p);
  std::swap(s->Psi.p, s->PsiPlus.p);
}

}  // namespace mslh


extern "C" {

const char* msl_status_string(msl_status st) {
  switch (st) {
    case MSL_OK: return "ok";
    case MSL_ERR_ARG: return "invalid argument";
    case MSL_ERR_STATE: return "call out of sequence";
    case MSL_ERR_IO: return "file error";
    case MSL_ERR_HIP: return "HIP runtime error";
    case MSL_ERR_UNSUPPORTED: return "unsupported size or mode";
    case MSL_ERR_OOM: return "out of memory";
    case MSL_ERR_RANGE: return "value out of range";
    case MSL_ERR_CANCELLED: return "cancelled";
  }
  return "unknown status";
}

msl_status msl_create(msl_sim** out, int device) {
  if (!out) return MSL_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MSL_ERR_HIP;
  if (device < 0 || device >= ndev) return MSL_ERR_ARG;
  msl_sim* s = new (std::nothrow) msl_sim;
  if (!s) return MSL_ERR_OOM;
  s->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return MSL_ERR_HIP;
  }
  *out = s;
  return MSL_OK;
}

void msl_destroy(msl_sim* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) {
    (void)hipStreamSynchronize(s->stream);
    if (s->own_stream) (void)hipStreamDestroy(s->stream);
  }
  delete s;
}

const char* msl_last_error(const msl_sim* s) { return s ? s->err.c_str() : "null handle"; }

msl_status msl_set_compat(msl_sim* s, unsigned flags) {
  if (!s) return MSL_ERR_ARG;
  s->compat = flags;
  return MSL_OK;
}

msl_status msl_set_potential_mode(msl_sim* s, int mode) {
  if (!s) return MSL_ERR_ARG;
  if (mode != MSL_POT_INTEGRATED && mode != MSL_POT_PROJECTED) {
    s->err = "unknown potential mode";
    return MSL_ERR_ARG;
  }
  s->potmode = mode;
  return MSL_OK;
}

msl_status msl_cancel(msl_sim* s) {
  if (!s) return MSL_ERR_ARG;
  s->cancel = 1;
  return MSL_OK;
}

// LoadAtomFile, U.cu:1293-1356: count, rest of line 1, then count lines of
// "symbol x y z".  A second line that does not parse as an atom is taken as
// the standard XYZ comment line.
msl_status msl_load_atom_file(msl_sim* s, const char* path) {
  API_BEGIN(s)
  require(path != nullptr, MSL_ERR_ARG, "null path");
  std::ifstream in(path);
  require((bool)in, MSL_ERR_IO, std::string("cannot open ") + path);
  std::string line;
  require((bool)std::getline(in, line), MSL_ERR_IO, "empty XYZ file");
  int n = 0;
  {
    std::istringstream ls(line);
    require((bool)(ls >> n) && n > 0, MSL_ERR_IO, "XYZ: bad atom count");
  }
  std::vector<std::string> rest;
  while (std::getline(in, line)) rest.push_back(line);
  size_t li = 0;
  if (!rest.empty()) {
    std::istringstream ls(rest[0]);
    std::string sym;
    float a, b, c;
    if (!(ls >> sym >> a >> b >> c)) li = 1;
  }
  std::vector<int> Z;
  std::vector<float> X, Y, Zp;
  for (int i = 0; i < n; ++i) {
    while (li < rest.size() && rest[li].find_first_not_of(" \t\r") == std::string::npos) ++li;
    require(li < rest.size(), MSL_ERR_IO, "XYZ: fewer atoms than declared");
    std::istringstream ls(rest[li++]);
    std::string sym;
    float a, b, c;
    require((bool)(ls >> sym >> a >> b >> c), MSL_ERR_IO, "XYZ: malformed atom line");
    int zn = (s->compat & MSL_COMPAT_REFERENCE_SYMBOLS) ? reference_znum(sym) : full_znum(sym);
    require(zn >= 1, MSL_ERR_IO, "XYZ: unknown element symbol " + sym);
    Z.push_back(zn); X.push_back(a); Y.push_back(b); Zp.push_back(c);
  }
  set_atoms(s, std::move(Z), std::move(X), std::move(Y), std::move(Zp));
  API_END(s)
}

msl_status msl_set_atoms(msl_sim* s, int n, const int* Z, const float* x, const float* y, const float* z) {
  API_BEGIN(s)
  require(n > 0 && Z && x && y && z, MSL_ERR_ARG, "bad atom arrays");
  set_atoms(s, std::vector<int>(Z, Z + n), std::vector<float>(x, x + n), std::vector<float>(y, y + n),
            std::vector<float>(z, z + n));
  API_END(s)
}

msl_status msl_upload_binned_atoms(msl_sim* s) {
  API_BEGIN(s)
  bin_atoms(s, nullptr, nullptr, nullptr);
  API_END(s)
}

msl_status msl_set_tds(msl_sim* s, uint64_t seed, float sigma) {
  if (!s) return MSL_ERR_ARG;
  s->tds_seed = seed;
  s->tds_sigma = sigma;
  s->tds_config = 0;
  return MSL_OK;
}

msl_status msl_upload_binned_atoms_tds(msl_sim* s) {
  API_BEGIN(s)
  require(s->have_atoms, MSL_ERR_STATE, "no atoms loaded");
  const int n = (int)s->x.size();
  std::vector<float> dx(n), dy(n), dz(n);
  tds_disp(s->tds_seed, s->tds_config, n, s->tds_sigma, dx.data(), dy.data(), dz.data());
  bin_atoms(s, dx.data(), dy.data(), dz.data());
  s->tds_config++;
  API_END(s)
}

msl_status msl_upload_binned_atoms_displaced(msl_sim* s, const float* dx, const float* dy, const float* dz) {
  API_BEGIN(s)
  bin_atoms(s, dx, dy, dz);
  API_END(s)
}

msl_status msl_tds_displacements(uint64_t seed, int config, int n, float sigma, float* dx, float* dy, float* dz) {
  if (n < 0 || !dx || !dy || !dz) return MSL_ERR_ARG;
  tds_disp(seed, config, n, sigma, dx, dy, dz);
  return MSL_OK;
}

// GetParameterisation, U.cu:1718-1747
msl_status msl_get_parameterisation(msl_sim* s, const char* path) {
  API_BEGIN(s)
  if (path && *path) {
    std::ifstream in(path);
    require((bool)in, MSL_ERR_IO, std::string("cannot open ") + path);
    std::vector<float> v;
    float f;
    while (in >> f) v.push_back(f);
    require(v.size() >= 103 * 12, MSL_ERR_IO, "parameter file holds fewer than 103 x 12 values");
    v.resize(103 * 12);
    s->fparams = v;
  } else {
    s->fparams.assign(kFparams, kFparams + 103 * 12);
  }
  s->ztab.release();
  s->have_params = true;
  API_END(s)
}

msl_status msl_get_fparams(const msl_sim* s, float* out) {
  if (!s || !out) return MSL_ERR_ARG;
  std::copy(s->fparams.begin(), s->fparams.end(), out);
  return MSL_OK;
}

msl_status msl_apply_microscope_parameters(msl_sim* s, float kV, float df, float Mod2f, float Arg2f, float Cs,
                                           float beta, float delta, float aperture) {
  API_BEGIN(s)
  require(kV > 0, MSL_ERR_ARG, "voltage must be positive");
  s->kV = kV; s->df = df; s->A1 = Mod2f; s->A1phi = Arg2f; s->Cs = Cs; s->beta = beta; s->delta = delta;
  s->obj = aperture;
  s->A2 = 0; s->A2phi = 0;  // Mod3fold / Arg3fold are fixed at 0 (U.cu:2554-2555)
  s->have_micro = true;
  API_END(s)
}

msl_status msl_set_calculation_variables(msl_sim* s, float ps, float df, int resX, int resY, float sizeX, float sizeY,
                                         float sizeZ, float offx, float offy) {
  API_BEGIN(s)
  (void)sizeX; (void)sizeY;
  calc_vars(s, ps, df, resX, resY, sizeZ, offx, offy, false);
  API_END(s)
}

msl_status msl_stem_set_calculation_variables(msl_sim* s, float ps, float df, int resX, int resY, float sizeX,
                                              float sizeY, float sizeZ) {
  API_BEGIN(s)
  (void)sizeX; (void)sizeY;
  calc_vars(s, ps, df, resX, resY, sizeZ, 0, 0, true);
  API_END(s)
}

msl_status msl_precalculate_frequencies(msl_sim* s) {
  API_BEGIN(s)
  precalc(s);
  API_END(s)
}

msl_status msl_initialise_wavefunctions(msl_sim* s) {
  API_BEGIN(s)
  require(s->have_freqs, MSL_ERR_STATE, "PreCalculateFrequencies first");
  require(std::isfinite(s->dzFD) && s->dzFD > 0, MSL_ERR_RANGE, "no stable finite-difference step for these conditions");
  s->conventionaldz = 1.0f;
  alloc_waves(s);
  HIPC(launch_fill(s->PsiMinus.p, (long long)s->npix(), make_float2(1.f, 0.f), s->stream));
  propagate(s, s->PsiMinus.p, s->Psi.p, s->dzFD);
  s->have_waves = true;
  s->fd_ready = true;
  API_END(s)
}

msl_status msl_initialise_wavefunctions_dz(msl_sim* s, float dz) {
  API_BEGIN(s)
  require(s->have_freqs, MSL_ERR_STATE, "PreCalculateFrequencies first");
  require(dz > 0, MSL_ERR_ARG, "slice thickness must be positive");
  s->conventionaldz = dz;
  alloc_waves(s);
  HIPC(launch_fill(s->PsiMinus.p, (long long)s->npix(), make_float2(1.f, 0.f), s->stream));
  HIPC(launch_fill(s->Psi.p, (long long)s->npix(), make_float2(1.f, 0.f), s->stream));
  s->have_waves = true;
  s->fd_ready = std::isfinite(s->dzFD) && s->dzFD > 0;
  if (s->fd_ready) propagate(s, s->PsiMinus.p, s->Psi.p, s->dzFD);
  API_END(s)
}

msl_status msl_stem_initialise_wavefunctions(msl_sim* s, int posx, int posy) {
  API_BEGIN(s)
  require(s->have_freqs, MSL_ERR_STATE, "PreCalculateFrequencies first");
  alloc_waves(s);
  make_probe(s, s->PsiMinus.p, posx, posy);
  HIPC(hipMemcpyAsync(s->Psi.p, s->PsiMinus.p, s->npix() * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  s->fd_ready = std::isfinite(s->dzFD) && s->dzFD > 0;
  if (s->fd_ready) propagate(s, s->PsiMinus.p, s->Psi.p, s->dzFD);
  s->have_waves = true;
  API_END(s)
}

msl_status msl_set_conventional_dz(msl_sim* s, float dz) {
  if (!s || !(dz > 0)) return MSL_ERR_ARG;
  s->conventionaldz = dz;
  return MSL_OK;
}

// MultisliceStepConv, U.cu:2085-2124: t from the potential kernel, then
// IFFT(P FFT(t psi) nf) nf as one row pass, one column pass, one row pass
msl_status msl_multislice_step_conv(msl_sim* s, int it) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "InitialiseWavefunctions first");
  require(it >= 1, MSL_ERR_ARG, "iterations count from 1");
  require(s->cancel == 0, MSL_ERR_CANCELLED, "cancelled");
  const long long bs = (long long)s->npix();
  potential(s, s->potmode, it, 1, s->conventionaldz, 0.f, 0, s->V.p, bs);
  ensure_prop(s, s->conventionaldz);
  const float2* src = (it == 1) ? s->PsiMinus.p : s->Psi.p;
  if (s->generic) {
    HIPC(launch_gen_cmul(src, s->V.p, s->PsiPlus.p, bs, s->stream));
    gen_fft2(s, s->PsiPlus.p, 1, false, 1.0f);
    gen_kmul(s, s->PsiPlus.p, CM_PROP, s->nf * s->nf);
    gen_fft2(s, s->PsiPlus.p, 1, true, 1.0f);
    copy_wave(s, s->Psi.p, s->PsiPlus.p);
  } else {
    const bool prune = s->prune_tiles > 0;
    RowArgs r = row_args(s, src, s->PsiPlus.p, 1, bs);
    r.t = s->V.p;
    r.prune_out = prune;
    row(s, ROW_FWD, RM_T, r);
    ColArgs c = col_args(s, s->PsiPlus.p, s->PsiPlus.p, 1, bs);
    c.scale = s->nf * s->nf;
    c.tiles_lo = s->prune_tiles;
    col(s, COL_FWD_MUL_INV, CM_PROP, c);
    RowArgs ri = row_args(s, s->PsiPlus.p, s->Psi.p, 1, bs);
    ri.prune_in = prune;
    row(s, ROW_INV, RM_NONE, ri);
  }
  API_END(s)
}

// MultisliceStep (finite difference), U.cu:2040-2083
msl_status msl_multislice_step(msl_sim* s, int it) {
  API_BEGIN(s)
  require(s->have_waves && s->fd_ready, MSL_ERR_STATE, "InitialiseWavefunctions first (finite-difference start)");
  require(s->cancel == 0, MSL_ERR_CANCELLED, "cancelled");
  const long long bs = (long long)s->npix();
  float zc;
  {
    const float h = s->MaxZ - s->MinZ;
    const float d = s->dzFD * (float)it;
    zc = h - d;
  }
  potential(s, 2, 0, 1, 1.0f, zc, 1, s->V.p, bs);
  if (s->generic) {
    gen_fft2(s, s->V.p, 1, false, 1.0f);
    gen_kmul(s, s->V.p, CM_BAND, s->nf * s->nf);
    gen_fft2(s, s->V.p, 1, true, 1.0f);
    copy_wave(s, s->G.p, s->Psi.p);
    gen_fft2(s, s->G.p, 1, false, 1.0f);
    gen_kmul(s, s->G.p, CM_GRAD, s->nf * s->nf);
    gen_fft2(s, s->G.p, 1, true, 1.0f);
    HIPC(launch_fd_update(s->G.p, s->PsiMinus.p, s->PsiPlus.p, s->Psi.p, s->V.p, bs, bs, s->sigma2, s->wavel,
                          s->dzFD, s->stream));
    gen_fft2(s, s->PsiPlus.p, 1, false, 1.0f);
    gen_kmul(s, s->PsiPlus.p, CM_BAND, s->nf * s->nf);
    gen_fft2(s, s->PsiPlus.p, 1, true, 1.0f);
    std::swap(s->PsiMinus.p, s->Psi.p);
    std::swap(s->Psi.p, s->PsiPlus.p);
  } else {
    fd_step_passes(s);
  }
  API_END(s)
}

// MultisliceStepRS, U.cu:2126-2178: real potential of the 1 A slab below
// z = H - it, the 10-term series from 1 (RSMS), then PsiPlus = result;
// after the first step PsiMinus = PsiPlus * Psi; Psi = PsiPlus.
msl_status msl_multislice_step_rs(msl_sim* s, int it) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "InitialiseWavefunctions first");
  require(it >= 1, MSL_ERR_ARG, "iterations count from 1");
  require(s->cancel == 0, MSL_ERR_CANCELLED, "cancelled");
  const long long n = (long long)s->npix();
  float zc;
  {
    const float h = s->MaxZ - s->MinZ;
    const float d = 1.0f * (float)it;
    zc = h - d;
  }
  potential(s, 3, 0, 1, 1.0f, zc, 1, s->V.p, n);
  float2 *a = s->PsiPlus.p, *b = s->G.p;
  for (int m = 1; m <= 10; ++m) {
    const float pre = (1.0f / (float)m) * (s->wavel * 1.0f / (4.0f * 3.1415927f));
    if (m == 1) {
      HIPC(launch_rs_term(nullptr, a, s->V.p, s->resX, s->resY, s->ps, pre, s->stream));
    } else {
      HIPC(launch_rs_term(a, b, s->V.p, s->resX, s->resY, s->ps, pre, s->stream));
      std::swap(a, b);
    }
  }
  if (a != s->PsiPlus.p)
    HIPC(hipMemcpyAsync(s->PsiPlus.p, a, n * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  if (it != 1) HIPC(launch_cmul(s->PsiPlus.p, s->Psi.p, s->PsiMinus.p, n, 1.0f, s->stream));
  HIPC(hipMemcpyAsync(s->Psi.p, s->PsiPlus.p, n * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  API_END(s)
}

msl_status msl_get_slices(const msl_sim* s, int conventional, int* out) {
  if (!s || !out) return MSL_ERR_ARG;
  if (conventional) *out = (int)std::ceil((s->MaxZ - s->MinZ) / s->conventionaldz);
  else *out = s->numberOfSlices;
  return MSL_OK;
}

msl_status msl_get_exit_wave(msl_sim* s, float* out) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "no wave");
  std::vector<float2> ew(s->npix());
  HIPC(hipMemcpyAsync(ew.data(), s->Psi.p, ew.size() * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  extrema(ew, s->maxposEW, s->minposEW);
  if (out) std::memcpy(out, ew.data(), ew.size() * sizeof(float2));
  API_END(s)
}

// GetSTEMExitWave, U.cu:2180-2229: unnormalised FFT, fftshift
msl_status msl_get_stem_exit_wave(msl_sim* s, float* out) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "no wave");
  HIPC(hipMemcpyAsync(s->B.p, s->Psi.p, s->npix() * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  fft2(s, s->B.p, 1, false);
  std::vector<float2> F(s->npix()), ew(s->npix());
  HIPC(hipMemcpyAsync(F.data(), s->B.p, F.size() * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  const int nx = s->resX, ny = s->resY;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const int x2 = x >= nx / 2 ? x - nx / 2 : x + nx / 2, y2 = y >= ny / 2 ? y - ny / 2 : y + ny / 2;
      ew[(size_t)y * nx + x] = F[(size_t)y2 * nx + x2];
    }
  extrema(ew, s->maxposEW, s->minposEW);
  if (out) std::memcpy(out, ew.data(), ew.size() * sizeof(float2));
  API_END(s)
}

// Diffraction log display of the centred spectrum C that GetSTEMExitWave
// reads back (the diffpattern branch of conjugationKernel, U.cu:230-231, the
// mode U.cu:2198 marks "Should be 1"): for pixel (x, y), with (x3, y3) the
// half-grid shift of (x, y) and (x4, y4) = (nx-1-x3, ny-1-y3),
// out = log(1 + 3000 |C[y3][x3] + conj(C[y4][x4])|^2), in float.
msl_status msl_get_diffraction_display(msl_sim* s, float* out) {
  API_BEGIN(s)
  require(out != nullptr, MSL_ERR_ARG, "null output");
  require(s->have_waves, MSL_ERR_STATE, "no wave");
  const int nx = s->resX, ny = s->resY;
  std::vector<float2> C((size_t)nx * ny);
  const msl_status st = msl_get_stem_exit_wave(s, reinterpret_cast<float*>(C.data()));
  if (st != MSL_OK) return st;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const int x3 = x < nx / 2 ? x + nx / 2 : x - nx / 2, y3 = y < ny / 2 ? y + ny / 2 : y - ny / 2;
      const float2 a = C[(size_t)y3 * nx + x3], b = C[(size_t)(ny - 1 - y3) * nx + (nx - 1 - x3)];
      const float re = a.x + b.x, im = a.y - b.y;
      out[(size_t)y * nx + x] = std::log(1.0f + 3000.0f * (re * re + im * im));
    }
  API_END(s)
}

msl_status msl_alloc_tds(msl_sim* s) {
  API_BEGIN(s)
  require(s->have_freqs, MSL_ERR_STATE, "PreCalculateFrequencies first");
  s->tds.ensure(s->npix());
  HIPC(hipMemsetAsync(s->tds.p, 0, s->npix() * sizeof(float), s->stream));
  s->have_tds = true;
  API_END(s)
}

msl_status msl_add_tds_waves(msl_sim* s) {
  API_BEGIN(s)
  require(s->have_tds && s->have_waves, MSL_ERR_STATE, "AllocTDS and a wave first");
  HIPC(hipMemcpyAsync(s->B.p, s->Psi.p, s->npix() * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  fft2(s, s->B.p, 1, false);
  const bool phys = (s->compat & MSL_COMPAT_PHYSICAL) != 0;
  HIPC(launch_tds_accum_mode(s->B.p, s->tds.p, s->resX, s->resY, phys ? 1 : 0, s->nf * s->nf, s->stream));
  API_END(s)
}

msl_status msl_get_tds(msl_sim* s, float* out) {
  API_BEGIN(s)
  require(s->have_tds, MSL_ERR_STATE, "AllocTDS first");
  require(out != nullptr, MSL_ERR_ARG, "null output");
  HIPC(hipMemcpyAsync(out, s->tds.p, s->npix() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  API_END(s)
}

namespace {
// CTF image(s) of the current wave, U.cu:2322-2343 with imagingKernel2
void ctf_images(msl_sim* s, int ndf, const float* dfs, float* d_out, const float* css = nullptr) {
  const long long n = (long long)s->npix();
  HIPC(hipMemcpyAsync(s->B.p, s->Psi.p, n * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  fft2(s, s->B.p, 1, false, s->nf);
  DBuf<float2> K;
  DBuf<float> dd, dc;
  K.ensure((size_t)ndf * n);
  dd.ensure(ndf);
  HIPC(hipMemcpyAsync(dd.p, dfs, ndf * sizeof(float), hipMemcpyHostToDevice, s->stream));
  if (css) {
    dc.ensure(ndf);
    HIPC(hipMemcpyAsync(dc.p, css, ndf * sizeof(float), hipMemcpyHostToDevice, s->stream));
  }
  CtfArgs c{};
  c.in = s->B.p; c.out = K.p; c.nx = s->resX; c.ny = s->resY; c.ndf = ndf; c.kx = s->kx.p; c.ky = s->ky.p;
  c.dfs = dd.p; c.css = css ? dc.p : nullptr; c.Cs = s->Cs; c.A1 = s->A1; c.A1phi = s->A1phi; c.A2 = s->A2; c.A2phi = s->A2phi; c.obj = s->obj;
  c.wavel = s->wavel; c.beta = s->beta / 1000.f; c.delta = s->delta; c.scale = 1.0f;
  HIPC(launch_ctf(c, s->stream));
  fft2(s, K.p, ndf, true, s->nf);
  HIPC(launch_intensity(K.p, d_out, n * ndf, 1.0f, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
}

// Detector model of SimulateImage (U.cu:2345-2428) on the CTF intensity in
// s->image: FFT nf, sqrt(DQE) filter, IFFT nf, modulus, seeded shot noise
// with rounding to counts, FFT nf, NTF filter, IFFT times 64 nf (8 CCD
// counts per electron, squared, U.cu:2360, 2423).  Result in s->detimg.
void detector_image(msl_sim* s, float dose) {
  const long long n = (long long)s->npix();
  if (!s->dqe_tab.p) {
    s->dqe_tab.ensure(725);
    s->ntf_tab.ensure(725);
    HIPC(hipMemcpyAsync(s->dqe_tab.p, kDetectorDQE, sizeof(kDetectorDQE), hipMemcpyHostToDevice, s->stream));
    HIPC(hipMemcpyAsync(s->ntf_tab.p, kDetectorNTF, sizeof(kDetectorNTF), hipMemcpyHostToDevice, s->stream));
  }
  s->detimg.ensure(n);
  float2* W = s->B.p;
  const int binning = 4;  // DQEKernel / NTFKernel launches, U.cu:2367, 2421
  HIPC(launch_real_to_complex(s->image.p, W, n, s->stream));
  fft2(s, W, 1, false);
  HIPC(launch_detector_filter(W, s->resX, s->resY, s->dqe_tab.p, binning, 1, s->nf, s->stream));
  fft2(s, W, 1, true);
  HIPC(launch_shot_noise(W, s->detimg.p, n, s->nf, dose, s->detector_seed, s->stream));
  fft2(s, s->detimg.p, 1, false);
  HIPC(launch_detector_filter(s->detimg.p, s->resX, s->resY, s->ntf_tab.p, binning, 0, s->nf, s->stream));
  fft2(s, s->detimg.p, 1, true, 64.f * s->nf);
  s->have_detimg = true;
}
}  // namespace

msl_status msl_simulate_image(msl_sim* s, float dose) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "no wave");
  s->image.ensure(s->npix());
  const float df = s->df;
  ctf_images(s, 1, &df, s->image.p);
  s->have_detimg = false;
  std::vector<float> im(s->npix());
  if (s->detector) {
    detector_image(s, dose);
    std::vector<float2> z(s->npix());
    d2h(s, z.data(), s->detimg.p, z.size() * sizeof(float2));
    for (size_t i = 0; i < z.size(); ++i) im[i] = std::sqrt(z[i].x * z[i].x + z[i].y * z[i].y);
  } else {
    d2h(s, im.data(), s->image.p, im.size() * sizeof(float));
  }
  // display range and contrast (U.cu:2430-2491) on the CTF intensity:
  // standard deviation over the top-left min(512, N) square
  size_t imx = 0, imn = 0;
  for (size_t i = 0; i < im.size(); ++i) {
    if (std::fabs(im[i]) > std::fabs(im[imx])) imx = i;
    if (std::fabs(im[i]) < std::fabs(im[imn])) imn = i;
  }
  s->maxposIm = std::fabs(im[imx]);
  s->minposIm = std::fabs(im[imn]);
  const int w = std::min(512, s->resX), h = std::min(512, s->resY);
  double mean = 0, var = 0;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) mean += std::fabs(im[(size_t)y * s->resX + x]);
  mean /= (double)w * h;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const double d = std::fabs(im[(size_t)y * s->resX + x]) - mean;
      var += d * d;
    }
  s->imagecontrast = (float)std::sqrt(var / ((double)w * h));
  s->have_image = true;
  API_END(s)
}

msl_status msl_set_detector(msl_sim* s, int enabled, uint64_t seed) {
  if (!s) return MSL_ERR_ARG;
  s->detector = enabled != 0;
  s->detector_seed = seed;
  s->have_detimg = false;
  return MSL_OK;
}

msl_status msl_get_detector_image(msl_sim* s, float* out) {
  API_BEGIN(s)
  require(s->have_detimg, MSL_ERR_STATE, "SimulateImage with the detector model first");
  require(out != nullptr, MSL_ERR_ARG, "null output");
  d2h(s, out, s->detimg.p, s->npix() * sizeof(float2));
  API_END(s)
}

// GetImValue's readout noise, U.cu:2497-2512: 4 + 1.33 g with g a standard
// normal clamped to +-3 (draw `counter` of the detector's Philox stream)
msl_status msl_readout_noise(const msl_sim* s, long long counter, float* out) {
  if (!s || !out || counter < 0) return MSL_ERR_ARG;
  double g = philox_normal(s->detector_seed, 0xDE7EC7ULL, (uint64_t)counter, 4);
  g = g < -3 ? -3 : (g > 3 ? 3 : g);
  *out = (float)(4.0 + g * 1.33);
  return MSL_OK;
}

msl_status msl_get_image(msl_sim* s, float* out) {
  API_BEGIN(s)
  require(s->have_image, MSL_ERR_STATE, "SimulateImage first");
  require(out != nullptr, MSL_ERR_ARG, "null output");
  d2h(s, out, s->image.p, s->npix() * sizeof(float));
  API_END(s)
}

msl_status msl_ctf_sweep(msl_sim* s, int ndf, const float* dfs, float* out) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "no wave");
  require(ndf > 0 && dfs && out, MSL_ERR_ARG, "bad defocus list");
  DBuf<float> d;
  d.ensure((size_t)ndf * s->npix());
  ctf_images(s, ndf, dfs, d.p);
  d2h(s, out, d.p, (size_t)ndf * s->npix() * sizeof(float));
  API_END(s)
}

// the GUI's batch Cs / defocus dialog (BatchProcessing.xaml.cs:29-135, run
// handler left empty by the reference): a grid of CTF images of the current
// wave, computed in chunks of at most 64 images
msl_status msl_cs_df_sweep(msl_sim* s, float cs0, float cs1, int ncs, float df0, float df1, int ndf, float* out,
                           const char* folder) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "no wave");
  require(ncs >= 1 && ndf >= 1 && (out || folder), MSL_ERR_ARG, "bad sweep");
  auto val = [](float a, float b, int n, int i) { return n > 1 ? a + (b - a) * (float)i / (float)(n - 1) : a; };
  const long long np = (long long)s->npix();
  const int total = ncs * ndf, chunk = 64;
  std::vector<float> cs, df, img((size_t)std::min(total, chunk) * np);
  for (int i = 0; i < ncs; ++i)
    for (int j = 0; j < ndf; ++j) {
      cs.push_back(val(cs0, cs1, ncs, i));
      df.push_back(val(df0, df1, ndf, j));
    }
  DBuf<float> d;
  d.ensure((size_t)std::min(total, chunk) * np);
  for (int k0 = 0; k0 < total; k0 += chunk) {
    const int m = std::min(chunk, total - k0);
    ctf_images(s, m, df.data() + k0, d.p, cs.data() + k0);
    d2h(s, img.data(), d.p, (size_t)m * np * sizeof(float));
    if (out) std::memcpy(out + (size_t)k0 * np, img.data(), (size_t)m * np * sizeof(float));
    for (int k = 0; folder && k < m; ++k) {
      char name[96];
      std::snprintf(name, sizeof(name), "/Cs_%g_df_%g.tiff", cs[k0 + k], df[k0 + k]);
      const msl_status st = msl_write_tiff((std::string(folder) + name).c_str(), s->resX, s->resY,
                                           img.data() + (size_t)k * np);
      require(st == MSL_OK, st, std::string("cannot write into ") + folder);
    }
  }
  API_END(s)
}

msl_status msl_adf(msl_sim* s, float inner, float outer, float* out) {
  API_BEGIN(s)
  require(s->have_waves, MSL_ERR_STATE, "no wave");
  require(out != nullptr, MSL_ERR_ARG, "null output");
  *out = adf_of(s, s->Psi.p, inner, outer);
  API_END(s)
}

msl_status msl_get_info(const msl_sim* s, msl_info* o) {
  if (!s || !o) return MSL_ERR_ARG;
  o->MaximumX = s->MaxX; o->MinimumX = s->MinX; o->MaximumY = s->MaxY; o->MinimumY = s->MinY;
  o->MaximumZ = s->MaxZ; o->MinimumZ = s->MinZ;
  o->numberOfSlices = s->numberOfSlices; o->kmax = s->kmax; o->conventionaldz = s->conventionaldz;
  o->maxposIm = s->maxposIm; o->maxposEW = s->maxposEW; o->minposEW = s->minposEW; o->minposIm = s->minposIm;
  o->imagecontrast = s->imagecontrast; o->wavelength = s->wavel; o->sigma2 = s->sigma2; o->dzFD = s->dzFD;
  o->natoms = (int)s->x.size(); o->nslab = s->nslab; o->resX = s->resX; o->resY = s->resY;
  o->band_tiles = s->prune_tiles;
  return MSL_OK;
}

msl_status msl_get_binned(const msl_sim* cs, int* st, float* x, float* y, float* z, int* Z) {
  msl_sim* s = const_cast<msl_sim*>(cs);
  API_BEGIN(s)
  require(s->binned, MSL_ERR_STATE, "not binned");
  const int n = s->nbinned, nb = s->nslab * 2500 + 1;
  if (st) d2h(s, st, s->start.p, nb * sizeof(int));
  if (x) d2h(s, x, s->bx.p, n * sizeof(float));
  if (y) d2h(s, y, s->by.p, n * sizeof(float));
  if (z) d2h(s, z, s->bz.p, n * sizeof(float));
  if (Z) d2h(s, Z, s->bZ.p, n * sizeof(int));
  API_END(s)
}

msl_status msl_get_bins(const msl_sim* cs, int* keys) {
  msl_sim* s = const_cast<msl_sim*>(cs);
  API_BEGIN(s)
  require(s->binned && keys, MSL_ERR_STATE, "not binned");
  d2h(s, keys, s->key.p, s->nbinned * sizeof(int));
  API_END(s)
}

msl_status msl_potential_slice(msl_sim* s, int it, float dz, int real, float* out) {
  API_BEGIN(s)
  require(s->have_freqs, MSL_ERR_STATE, "PreCalculateFrequencies first");
  require(out != nullptr && dz > 0, MSL_ERR_ARG, "bad arguments");
  DBuf<float> d;
  d.ensure(s->npix() * 2);
  potential(s, s->potmode, it, 1, dz, 0.f, real ? 2 : 0, d.p, (long long)s->npix());
  d2h(s, out, d.p, s->npix() * (real ? 1 : 2) * sizeof(float));
  API_END(s)
}

msl_status msl_set_potential_tile(msl_sim* s, int variant) {
  if (!s) return MSL_ERR_ARG;
  if (variant < 0 || variant >= potential_tile_variants()) return MSL_ERR_ARG;
  s->pot_tile = variant;
  return MSL_OK;
}

// Times every pixel-tile shape of the potential kernel on the handle's atoms
// and grid (nslices slices of thickness dz, its potential mode), keeps the
// fastest whose potentials equal the 16x16 tile's bit for bit.
msl_status msl_autotune_potential(msl_sim* s, float dz, int nslices, int* best, float* ms_per_variant) {
  API_BEGIN(s)
  require(s->have_freqs, MSL_ERR_STATE, "PreCalculateFrequencies first");
  require(s->binned, MSL_ERR_STATE, "UploadBinnedAtoms first");
  require(dz > 0 && nslices >= 1, MSL_ERR_ARG, "bad slice set");
  const int nv = potential_tile_variants();
  const long long np = (long long)s->npix();
  DBuf<float> d;
  d.ensure((size_t)nslices * np);
  std::vector<float> ref((size_t)nslices * np), got((size_t)nslices * np);
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  const int saved = s->pot_tile;
  int pick = 0;
  float tbest = 0.f;
  try {
    for (int v = 0; v < nv; ++v) {
      s->pot_tile = v;
      potential(s, s->potmode, 1, nslices, dz, 0.f, 2, d.p, np);  // warm-up
      HIPC(hipEventRecord(e0, s->stream));
      potential(s, s->potmode, 1, nslices, dz, 0.f, 2, d.p, np);
      HIPC(hipEventRecord(e1, s->stream));
      HIPC(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, e0, e1));
      d2h(s, v == 0 ? ref.data() : got.data(), d.p, ref.size() * sizeof(float));
      const bool same = v == 0 || std::memcmp(ref.data(), got.data(), ref.size() * sizeof(float)) == 0;
      if (ms_per_variant) ms_per_variant[v] = same ? ms : -1.f;
      if (same && (v == 0 || ms < tbest)) {
        pick = v;
        tbest = ms;
      }
    }
  } catch (...) {
    s->pot_tile = saved;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s->pot_tile = pick;
  if (best) *best = pick;
  API_END(s)
}

msl_status msl_set_wave(msl_sim* s, int which, const float* in) {
  API_BEGIN(s)
  require(s->have_waves && in, MSL_ERR_STATE, "no waves");
  float2* d = which == 0 ? s->Psi.p : which == 1 ? s->PsiMinus.p : s->PsiPlus.p;
  h2d(s, d, in, s->npix() * sizeof(float2));
  API_END(s)
}

msl_status msl_get_wave(msl_sim* s, int which, float* out) {
  API_BEGIN(s)
  require(s->have_waves && out, MSL_ERR_STATE, "no waves");
  const float2* d = which == 0 ? s->Psi.p : which == 1 ? s->PsiMinus.p : s->PsiPlus.p;
  HIPC(hipStreamSynchronize(s->stream));
  d2h(s, out, d, s->npix() * sizeof(float2));
  API_END(s)
}

msl_status msl_set_profiling(msl_sim* s, int on) {
  if (!s) return MSL_ERR_ARG;
  s->profiling = on != 0;
  std::fill(s->kstats, s->kstats + 8, 0.0);
  return MSL_OK;
}

msl_status msl_kernel_stats(const msl_sim* s, double* out8) {
  if (!s || !out8) return MSL_ERR_ARG;
  std::copy(s->kstats, s->kstats + 8, out8);
  return MSL_OK;
}

// Batch width: times the propagation of one chunk of B probes over a few
// slices (probe synthesis, row / column passes, fused detector) for B = 64,
// 128, ... up to the probes given (and ~32 GiB of waves), and keeps the width
// with the least time per probe-slice for batch = 0 runs on this grid.
msl_status msl_autotune_batch(msl_sim* s, const msl_stem_batch_desc* d, int nslices, int* best, int* widths,
                              float* us_per_probe_slice, int* ncand) {
  API_BEGIN(s)
  require(d != nullptr && d->nprobe > 0 && d->posx && d->posy, MSL_ERR_ARG, "bad batch descriptor");
  require(nslices >= 1, MSL_ERR_ARG, "nslices must be positive");
  require(s->have_freqs, MSL_ERR_STATE, "STEMSetCalculationVariables and PreCalculateFrequencies first");
  const long long np = (long long)s->resX * s->resY;
  const int cap = (int)std::min<long long>(std::min(d->nprobe, 65535), (32LL << 30) / (np * 8));
  std::vector<int> cand;
  for (int b = 64; b < cap; b *= 2) cand.push_back(b);
  cand.push_back(std::max(1, cap));
  std::vector<float> adf((size_t)cap), cost(cand.size());
  const int saved = s->tuned_batch;
  s->tuned_batch = 0;
  int pick = 0;
  for (size_t k = 0; k < cand.size(); ++k) {
    msl_stem_batch_desc t = *d;
    t.nprobe = cand[k];
    t.nconfig = 1;
    t.sigma = 0.f;
    t.nslices = nslices;
    t.batch = cand[k];
    t.adf_out = adf.data();
    t.pacbed_out = nullptr;
    t.exit_out = nullptr;
    // the first run of a width also builds its graph and warms the caches;
    // the cost is the faster of the next two
    cost[k] = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      const msl_status st = msl_stem_batched(s, &t);
      if (st != MSL_OK) {
        s->tuned_batch = saved;
        return st;
      }
      if (rep > 0) cost[k] = std::min(cost[k], s->timings[2] * 1000.f / ((float)cand[k] * (float)nslices));
    }
    if (cost[k] < cost[pick]) pick = (int)k;
  }
  // the widest candidate within 2 % of the fastest: a run of many chunks
  // gains more from longer launches (shorter ramp and tail per probe) than
  // one chunk over a few slices shows (512^2 SrTiO3: 128 and 256 probes
  // timed 1.177 and 1.189 us per probe-slice here, and 1.005e6 against
  // 1.084e6 probe-slice/s over 16 configurations)
  for (size_t k = 0; k < cand.size(); ++k)
    if (cost[k] <= 1.02f * cost[pick] && cand[k] > cand[pick]) pick = (int)k;
  s->tuned_batch = cand[pick];
  s->tuned_batch_n = s->resX;
  if (best) *best = cand[pick];
  if (ncand) {
    const int n = std::min<int>(*ncand, (int)cand.size());
    for (int k = 0; k < n; ++k) {
      if (widths) widths[k] = cand[k];
      if (us_per_probe_slice) us_per_probe_slice[k] = cost[k];
    }
    *ncand = (int)cand.size();
  }
  API_END(s)
}

msl_status msl_last_timings(const msl_sim* s, float* out4) {
  if (!s || !out4) return MSL_ERR_ARG;
  std::copy(s->timings, s->timings + 4, out4);
  return MSL_OK;
}

// batched 2D FFT on caller device memory
msl_status msl_fft2_device(int nx, int ny, int nbatch, void* data, int inverse, void* stream) {
  if (!fft_size_supported(nx) || !fft_size_supported(ny) || nbatch < 0 || !data) return MSL_ERR_ARG;
  try {
    const hipStream_t st = (hipStream_t)stream;
    const auto tx = twiddles(nx), ty = twiddles(ny);
    DBuf<float2> twx, twy;
    twx.ensure(nx);
    twy.ensure(ny);
    HIPC(hipMemcpyAsync(twx.p, tx.data(), nx * sizeof(float2), hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync(twy.p, ty.data(), ny * sizeof(float2), hipMemcpyHostToDevice, st));
    float2* w = (float2*)data;
    const long long bs = (long long)nx * ny;
    RowArgs r{};
    r.src = w; r.dst = w; r.src_bstride = r.dst_bstride = bs; r.nx = nx; r.ny = ny; r.nbatch = nbatch;
    r.t_group = 1; r.scale = 1.f; r.tw = twx.p;
    ColArgs c{};
    c.src = w; c.dst = w; c.src_bstride = c.dst_bstride = bs; c.nx = nx; c.ny = ny; c.nbatch = nbatch;
    c.scale = 1.f; c.tw = twy.p;
    if (!inverse) {
      HIPC(launch_row(ROW_FWD, RM_NONE, r, st));
      HIPC(launch_col(COL_FWD, CM_SCALE, c, st));
    } else {
      HIPC(launch_col(COL_INV, CM_SCALE, c, st));
      HIPC(launch_row(ROW_INV, RM_NONE, r, st));
    }
    HIPC(hipStreamSynchronize(st));
  } catch (const Error& e) {
    return e.code;
  }
  return MSL_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batched STEM
namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" msl_status msl_stem_batched(msl_sim* s, const msl_stem_batch_desc* d) {
  API_BEGIN(s)
  require(d != nullptr && d->nprobe > 0 && d->posx && d->posy && d->adf_out, MSL_ERR_ARG, "bad batch descriptor");
  require(s->have_freqs, MSL_ERR_STATE, "STEMSetCalculationVariables and PreCalculateFrequencies first");
  require(s->have_atoms, MSL_ERR_STATE, "no atoms");
  require(d->dz > 0, MSL_ERR_ARG, "slice thickness must be positive");
  require(!s->generic, MSL_ERR_UNSUPPORTED, "batched STEM needs power-of-two sides 16..2048");
  s->cancel = 0;
  const double t0 = now_ms();
  const int nx = s->resX, ny = s->resY;
  const long long np = (long long)nx * ny;
  const int nslices = d->nslices > 0 ? d->nslices : (int)std::ceil((s->MaxZ - s->MinZ) / d->dz);
  const int nconf = std::max(1, d->nconfig);
  const bool displace = d->sigma > 0.f;
  // probes per launch: the tuned width for this grid (msl_autotune_batch), else
  // all of them when they fit in ~16 GiB of waves
  const int auto_b = (s->tuned_batch > 0 && s->tuned_batch_n == nx)
                         ? s->tuned_batch
                         : (int)std::min<long long>(d->nprobe, (16LL << 30) / (np * 8));
  int B = d->batch > 0 ? d->batch : auto_b;
  B = std::max(1, std::min(std::min(B, d->nprobe), 65535));  // grid.y limit of the column pass
  const int C = col_tile(ny), ntiles = nx / C;
  alloc_waves(s);
  ensure_ztab(s);
  ensure_prop(s, d->dz);
  DBuf<float2> tall, W;
  DBuf<float> partial, adf;
  DBuf<int> pos;
  DBuf<float> inten, pacbed;
  tall.ensure((size_t)nslices * np);
  W.ensure((size_t)B * np);
  partial.ensure((size_t)B * ntiles);
  adf.ensure(d->nprobe);
  pos.ensure((size_t)2 * B);
  HIPC(hipMemsetAsync(adf.p, 0, d->nprobe * sizeof(float), s->stream));
  if (d->pacbed_out) {
    inten.ensure((size_t)B * np);
    pacbed.ensure(np);
    HIPC(hipMemsetAsync(pacbed.p, 0, np * sizeof(float), s->stream));
  }
  // probe normalisation: sum |psi| of the probe at the origin (U.cu:2698-2705;
  // sqrt(sum |psi|^2) under compat=physical); integer-pixel shifts leave it
  // unchanged
  float2* tmp = s->B.p;
  {
    ProbeArgs p{};
    p.out = tmp; p.bstride = np; p.nx = nx; p.ny = ny; p.nprobe = 1; p.kx = s->kx.p; p.ky = s->ky.p;
    p.apert = (float)((double)s->obj * 0.001 / (double)s->wavel);
    p.ps = s->ps; p.df = s->df; p.Cs = s->Cs; p.wavel = s->wavel;
    HIPC(launch_probe_kspace(p, s->stream));
    fft2(s, tmp, 1, true);
  }
  HIPC(launch_probe_normalise(tmp, np, 1, np, s->red.p, (s->compat & MSL_COMPAT_PHYSICAL) ? 1 : 0, s->stream));
  const int l1chunks = (int)((np + 4095) / 4096);
  const float norm = download_scalar(s, s->red.p + l1chunks);
  require(norm > 0 && std::isfinite(norm), MSL_ERR_RANGE, "probe has zero amplitude (aperture too small?)");
  const bool want_exit = d->exit_out != nullptr;
  const bool prune = s->prune_tiles > 0;
  // The probes' spectra vanish outside the aperture (k_probe_kspace's FP32
  // test k < apert, and k >= |kx|): when every column the band pruning drops
  // lies outside it, by a margin, probe synthesis, the column IFFT and the
  // first row pass move only the band's columns.  The others are then never
  // written, and no pass reads them (the |Psi|^2 map, which reads every
  // column, keeps the full set-up).
  const float apert = (float)((double)s->obj * 0.001 / (double)s->wavel);
  bool probe_in_band = prune && !d->pacbed_out;
  for (int x = nx / 4; x < 3 * nx / 4 && probe_in_band; ++x)
    if (!(std::sqrt(s->hkx[x] * s->hkx[x]) > apert * 1.0001f)) probe_in_band = false;
  const int band_cols = probe_in_band ? nx / 4 : 0;
  // detector tiles per probe: the |Psi|^2 map needs every column written
  const int det_tiles = (prune && !d->pacbed_out) ? 2 * s->prune_tiles : ntiles;
  // One chunk of nb probes whose positions are in pos: probe synthesis, the
  // slice loop and the fused detector.  Every pointer it uses is fixed for
  // the whole call, so the sequence can be captured once per chunk size and
  // replayed (no memsets inside).
  auto enqueue_chunk = [&](int nb, KTimer* kt) {
    // probes in k-space, then IFFT_y with the 1/sum|psi| scale -> (kx, y) domain
    ProbeArgs p{};
    p.out = W.p; p.bstride = np; p.nx = nx; p.ny = ny; p.nprobe = nb; p.kx = s->kx.p; p.ky = s->ky.p;
    p.posx = pos.p; p.posy = pos.p + B;
    p.apert = apert;
    p.ps = s->ps; p.df = s->df; p.Cs = s->Cs; p.wavel = s->wavel;
    p.band_cols = band_cols;
    HIPC(launch_probe_kspace(p, s->stream));
    ColArgs ci = col_args(s, W.p, W.p, nb, np);
    ci.scale = 1.0f / norm;
    if (probe_in_band) ci.tiles_lo = s->prune_tiles;
    col(s, COL_INV, CM_SCALE, ci);
    for (int i = 1; i <= nslices; ++i) {
      RowArgs r = row_args(s, W.p, W.p, nb, np);
      r.t = tall.p + (size_t)(i - 1) * np;
      r.t_group = nb;
      r.t_bstride = 0;
      // the probe is band-limited when its aperture lies inside the band
      r.prune_in = prune && (i > 1 || probe_in_band);
      r.prune_out = prune;
      if (kt) kt->begin(0);
      row(s, ROW_INV_MUL_FWD, RM_T, r);
      if (kt) kt->end();
      ColArgs c = col_args(s, W.p, W.p, nb, np);
      c.tiles_lo = s->prune_tiles;
      if (i < nslices || want_exit) {
        c.scale = s->nf * s->nf;
        if (kt) kt->begin(1);
        col(s, COL_FWD_MUL_INV, CM_PROP, c);
        if (kt) kt->end();
      } else {
        c.dst = nullptr;
        c.scale = 1.0f;
        c.kin = d->adf_inner * 1e-3f / s->wavel;
        c.kout = d->adf_outer * 1e-3f / s->wavel;
        c.det_scale = s->nf * s->nf;
        c.partial = partial.p;
        c.inten = d->pacbed_out ? inten.p : nullptr;
        if (d->pacbed_out) c.tiles_lo = 0;
        if (kt) kt->begin(2);
        col(s, COL_FWD_DETECT, CM_PROP, c);
        if (kt) kt->end();
      }
    }
  };
  // graphs replace the eager launches unless exit waves are copied out per
  // chunk or per-kernel events are being recorded
  const bool graph_mode = d->use_graph != 0 && !want_exit && !s->profiling;
  struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    int nb = -1;
    void reset() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
      e = nullptr;
      g = nullptr;
      nb = -1;
    }
    ~Graph() { reset(); }
  } graph;
  double tb = 0, tp = 0, tprop = 0;
  // displacements of the configuration being binned, and the next
  // configuration's, drawn on the host while this one propagates (the GPU
  // would otherwise wait for them between configurations)
  std::vector<float> dx, dy, dzv, ndx, ndy, ndz;
  bool have_next = false;
  for (int cfg = 0; cfg < nconf; ++cfg) {
    require(s->cancel == 0, MSL_ERR_CANCELLED, "cancelled");
    double a0 = now_ms();
    if (displace) {
      const int n = (int)s->x.size();
      if (have_next) {
        dx.swap(ndx); dy.swap(ndy); dzv.swap(ndz);
        have_next = false;
      } else {
        dx.resize(n); dy.resize(n); dzv.resize(n);
        tds_disp(d->seed, cfg, n, d->sigma, dx.data(), dy.data(), dzv.data());
      }
      bin_atoms(s, dx.data(), dy.data(), dzv.data());
    } else if (cfg == 0 || !s->binned) {
      bin_atoms(s, nullptr, nullptr, nullptr);
    }
    HIPC(hipStreamSynchronize(s->stream));
    double a1 = now_ms();
    KTimer kt(s);
    kt.begin(3);
    potential(s, s->potmode, 1, nslices, d->dz, 0.f, 0, tall.p, np);
    kt.end();
    kt.flush();
    HIPC(hipStreamSynchronize(s->stream));
    double a2 = now_ms();
    tb += a1 - a0;
    tp += a2 - a1;
    for (int p0 = 0; p0 < d->nprobe; p0 += B) {
      require(s->cancel == 0, MSL_ERR_CANCELLED, "cancelled");
      const int nb = std::min(B, d->nprobe - p0);
      HIPC(hipMemcpyAsync(pos.p, d->posx + p0, nb * sizeof(int), hipMemcpyHostToDevice, s->stream));
      HIPC(hipMemcpyAsync(pos.p + B, d->posy + p0, nb * sizeof(int), hipMemcpyHostToDevice, s->stream));
      if (graph_mode) {
        if (graph.nb != nb) {
          graph.reset();
          HIPC(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
          try {
            enqueue_chunk(nb, nullptr);
          } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          HIPC(hipStreamEndCapture(s->stream, &graph.g));
          HIPC(hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0));
          graph.nb = nb;
        }
        HIPC(hipGraphLaunch(graph.e, s->stream));
      } else {
        enqueue_chunk(nb, &kt);
        kt.flush();
      }
      if (displace && p0 == 0 && cfg + 1 < nconf) {
        const int n = (int)s->x.size();
        ndx.resize(n); ndy.resize(n); ndz.resize(n);
        tds_disp(d->seed, cfg + 1, n, d->sigma, ndx.data(), ndy.data(), ndz.data());
        have_next = true;
      }
      if (want_exit) {
        RowArgs ri = row_args(s, W.p, W.p, nb, np);
        ri.prune_in = prune;
        row(s, ROW_INV, RM_NONE, ri);
        HIPC(hipMemcpyAsync(d->exit_out + ((size_t)cfg * d->nprobe + p0) * np * 2, W.p, (size_t)nb * np * sizeof(float2),
                            hipMemcpyDeviceToHost, s->stream));
        // detector on the stored exit waves
        for (int q = 0; q < nb; ++q) {
          const float v = adf_of(s, W.p + (size_t)q * np, d->adf_inner, d->adf_outer);
          float cur = 0;
          d2h(s, &cur, adf.p + p0 + q, sizeof(float));
          cur += v;
          h2d(s, adf.p + p0 + q, &cur, sizeof(float));
        }
      } else {
        HIPC(launch_reduce_rows(partial.p, nb, det_tiles, adf.p + p0, 1, s->stream));
        if (d->pacbed_out) HIPC(launch_sum_maps(inten.p, nb, np, pacbed.p, 1, s->stream));
      }
    }
    HIPC(hipStreamSynchronize(s->stream));
    tprop += now_ms() - a2;
  }
  std::vector<float> h(d->nprobe);
  d2h(s, h.data(), adf.p, d->nprobe * sizeof(float));
  for (int i = 0; i < d->nprobe; ++i) d->adf_out[i] = h[i] / (float)nconf;
  if (d->pacbed_out) {
    std::vector<float> pc(np);
    d2h(s, pc.data(), pacbed.p, np * sizeof(float));
    const float sc = 1.0f / ((float)nconf * (float)d->nprobe);
    for (long long i = 0; i < np; ++i) d->pacbed_out[i] = pc[i] * sc;
  }
  s->timings[0] = (float)tb;
  s->timings[1] = (float)tp;
  s->timings[2] = (float)tprop;
  s->timings[3] = (float)(now_ms() - t0);
  API_END(s)
}
