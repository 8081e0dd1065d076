// msl_host.hpp -- internal to libmsl: the simulation handle and the host
// helpers shared by the C-ABI translation units (not installed, not part of
// the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/msl.h"
#include "fparams_table.h"
#include "mtf_tables.h"
#include "msl_internal.hpp"

namespace mslh {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw Error(MSL_ERR_OOM, std::string(what) + ": out of device memory");
  throw Error(MSL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

inline void require(bool ok, int code, const std::string& msg) {
  if (!ok) throw Error(code, msg);
}

template <class T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    HIPC(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
};

}  // namespace mslh

// ------------------------------------------------------------------ handle
struct msl_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: the stream belongs to a caller (job handles of a serving run)
  std::string err;
  unsigned compat = MSL_COMPAT_REFERENCE_SYMBOLS;
  int potmode = MSL_POT_INTEGRATED;
  // pixel tile shape of the potential kernel (msl_autotune_potential); 8x32
  // by default (on MI355X 0.30 against 0.32 ms for 16x16 per 8 slices of the
  // 512^2 SrTiO3 workload, equal on 1024^2 Si)
  int pot_tile = 3;
  volatile int cancel = 0;

  // structure (U.cuh:88-99, 127-132)
  std::vector<int> Z;
  std::vector<float> x, y, z;
  float MaxX = 0, MinX = 0, MaxY = 0, MinY = 0, MaxZ = 0, MinZ = 0;
  bool have_atoms = false;

  // parameterisation
  std::vector<float> fparams = std::vector<float>(kFparams, kFparams + 103 * 12);
  bool have_params = false;
  mslh::DBuf<float> ztab;

  // binned atoms on the device
  mslh::DBuf<float> in_x, in_y, in_z, bx, by, bz;
  mslh::DBuf<int> in_Z, bZ, key, count, start, fill, perm, bad, scan_tmp;
  int nslab = 0, nbinned = 0;
  bool binned = false;
  uint64_t tds_seed = 0x4D534CULL;
  float tds_sigma = 0.075f;  // TDSRand, U.cu:34
  int tds_config = 0;

  // microscope (U.cuh:101-111)
  float kV = 0, df = 0, A1 = 0, A1phi = 0, A2 = 0, A2phi = 0, Cs = 0, beta = 0, delta = 0, obj = 0;
  bool have_micro = false;

  // calculation variables (U.cuh:113-125)
  int resX = 0, resY = 0;
  float ps = 0, offx = 0, offy = 0, ksizex = 0, ksizey = 0;
  float wavel = 0, sigma2 = 0, dzFD = 0;
  int numberOfSlices = 0, loadblocksx = 0, loadblocksy = 0, loadslicesz = 0;
  bool have_calc = false, have_freqs = false;
  float kmax = 0, nf = 0;
  int prune_tiles = 0;  // > 0: columns N/4 <= x < 3N/4 are outside the propagator band
  bool zb_rows = false;  // rows N/4 <= y < 3N/4 are outside the propagator band (every column)
  std::vector<float> hkx, hky;
  mslh::DBuf<float> kx, ky;
  mslh::DBuf<float2> twx, twy, px, py;
  mslh::DBuf<int2> prow;  // rows of each column inside |k| < kmax
  float prop_dz = std::nanf("");
  float conventionaldz = 1.0f;

  // waves
  mslh::DBuf<float2> PsiMinus, Psi, PsiPlus, V, G, B;
  // sides not powers of two: the generic transforms (k_generic.hip) and
  // their scratch; the batched, serving and differentiable paths refuse
  bool generic = false;
  mslh::DBuf<float2> gscr;
  mslh::DBuf<float> gpart;
  bool have_waves = false, fd_ready = false;
  mslh::DBuf<float> tds, image, red;
  // stochastic detector model (SimulateImage, U.cu:2345-2428)
  bool detector = false;
  uint64_t detector_seed = 0x4D534CULL;
  mslh::DBuf<float> dqe_tab, ntf_tab;
  mslh::DBuf<float2> detimg;
  bool have_detimg = false;
  bool have_tds = false, have_image = false;
  float maxposIm = 0, maxposEW = 0, minposEW = 0, minposIm = 0, imagecontrast = 0;
  float timings[4] = {0, 0, 0, 0};
  int tuned_batch = 0;  // probes per launch chosen by msl_autotune_batch (0: not tuned)
  int tuned_batch_n = 0;  // grid size it was tuned for
  // per-kernel HIP-event timing of batched runs: [row ms, row n, col ms,
  // col n, detect ms, detect n, potential ms, potential n]
  bool profiling = false;
  double kstats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<hipEvent_t> events;

  size_t npix() const { return (size_t)resX * resY; }
  ~msl_sim() {
    for (auto e : events) (void)hipEventDestroy(e);
  }
};

namespace mslh {

using msl::ColArgs;
using msl::RowArgs;

void set_device(msl_sim* s);
void d2h(msl_sim* s, void* dst, const void* src, size_t bytes);
void h2d(msl_sim* s, void* dst, const void* src, size_t bytes);
RowArgs row_args(msl_sim* s, const float2* src, float2* dst, int nbatch, long long bs);
ColArgs col_args(msl_sim* s, const float2* src, float2* dst, int nbatch, long long bs);
void row(msl_sim* s, int mode, int mul, const RowArgs& a);
void col(msl_sim* s, int mode, int mul, const ColArgs& a);
// separable Fresnel factors for slice thickness dz (px, py)
void ensure_prop(msl_sim* s, float dz);
void ensure_ztab(msl_sim* s);
void set_atoms(msl_sim* s, std::vector<int>&& Z, std::vector<float>&& x, std::vector<float>&& y,
               std::vector<float>&& z);
void bin_atoms(msl_sim* s, const float* dx, const float* dy, const float* dz);
void calc_vars(msl_sim* s, float ps, float df, int rx, int ry, float sizeZ, float ox, float oy, bool stem);
void precalc(msl_sim* s);
// potential / transmission of slices first .. first+nslices-1 (mode 0, 1) or
// of the plane zplane (mode 2) into out (out_kind 0 complex t, 1 complex V,
// 2 real V), slices sstride elements apart
void potential(msl_sim* s, int mode, int first, int nslices, float dz, float zplane, int out_kind, void* out,
               long long sstride);
// the launch arguments of the same, without launching
msl::PotArgs pot_args(msl_sim* s, int mode, int first, int nslices, float dz, float zplane, int out_kind, void* out,
                      long long sstride);
float download_scalar(msl_sim* s, const float* d);
std::vector<float2> twiddles(int n);

}  // namespace mslh

// ------------------------------------------------------------------ C-ABI bodies
#define API_BEGIN(h)                                        \
  if (!(h)) return MSL_ERR_ARG;                             \
  try {                                                     \
    set_device(h);
#define API_END(h)                                          \
  }                                                         \
  catch (const Error& e) {                                  \
    (h)->err = e.what();                                    \
    return e.code;                                          \
  }                                                         \
  catch (const std::bad_alloc&) {                           \
    (h)->err = "host allocation failed";                    \
    return MSL_ERR_OOM;                                     \
  }                                                         \
  catch (const std::exception& e) {                         \
    (h)->err = e.what();                                    \
    return MSL_ERR_ARG;                                     \
  }                                                         \
  (h)->err.clear();                                         \
  return MSL_OK;
