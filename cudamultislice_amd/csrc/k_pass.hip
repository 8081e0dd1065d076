// k_pass.hip -- row and column FFT passes with fused pointwise hooks.
//
// One multislice step over a batch of waves is two passes over HBM
// (transmission multiply of U.cu:2105-2109, FFT/propagator/IFFT of
// U.cu:2117-2122, restated):
//   row pass    IFFT_x -> * t(x, y) -> FFT_x               (ROW_INV_MUL_FWD)
//   column pass FFT_y  -> * P(kx, ky) / (nx ny) -> IFFT_y  (COL_FWD_MUL_INV)
// Between slices a wave stays in the (kx, y) domain; the leading IFFT_x of
// the next row pass returns it to real space in registers.
//
// Each pass exists in a few layouts (variants); the variant per length is
// chosen by msl_autotune_passes() from timed runs, or by the defaults below.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>
#include <vector>

#include "msl_fft.hpp"
#include "msl_internal.hpp"

namespace msl {

static int lg2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// ------------------------------------------------------------------ variants
// row:    0 workgroup barriers, unpadded LDS
//         1 wave-local exchanges (lines never straddle a wave), LDS pad 1/16
//         2 as 1 with the alternative plan (AltPlan: other radices and
//           threads per line)
// column: 0 [row][col] interleaved LDS, workgroup barriers
//         1 [col][row] LDS with pad 1/16, workgroup barriers
//         2 as 0 with the alternative plan
//         3 as 0 with half as many columns per workgroup (more workgroups
//           per CU to overlap the barriers)
//         4 as 0 with two LDS copies used by alternate exchanges (one
//           barrier per exchange instead of two)
// (a staged column layout with one wavefront per column and an LDS pad of
// 1/8 were measured slower or equal on MI355X and were removed)
// defaults from msl_autotune_passes on MI355X (band-pruned passes, 2 GiB of
// waves): rows of 512 take the alternative plan (1.27x); at 1024 it is 1.09x
// faster with a precomputed transmission but holds 256 VGPRs (one wave per
// SIMD) when the row pass evaluates exp(i sigma V) itself, so 1024 keeps 1
static int g_row_var[12] = {0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 1, 1};
static int g_col_var[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

int pass_variant(int kind, int n) { return kind == 0 ? g_row_var[lg2(n)] : g_col_var[lg2(n)]; }
void set_pass_variant(int kind, int n, int v) {
  if (kind == 0) g_row_var[lg2(n)] = v;
  else g_col_var[lg2(n)] = v;
}
int num_pass_variants(int kind) { return kind == 0 ? 3 : 5; }

template <int N> struct ColTile0 { static constexpr int C = (N >= 2048) ? 4 : (N >= 1024 ? 8 : 16); };

template <int N, int CV> struct ColCfg {
  using P = typename std::conditional<CV == 2, AltPlan<N>, Plan<N>>::type;
  static constexpr bool INTERLEAVED = CV != 1;
  static constexpr int T = P::T;
  static constexpr int C = CV == 3 ? ColTile0<N>::C / 2 : ColTile0<N>::C;
  static constexpr int SH = INTERLEAVED ? 0 : 4;
  // [col][row] line stride in float2: span + padding so consecutive columns
  // start two banks apart (stride = 2 mod 32 float2)
  static constexpr int SPAN = LdsLine<SH>::span(N);
  static constexpr int LSC = SPAN + ((2 - SPAN % 32) + 32) % 32;
  static constexpr bool DUAL = CV == 4;
  static constexpr int LDS = INTERLEAVED ? N * C * (DUAL ? 2 : 1) : C * LSC;
  static constexpr int THREADS = C * T;
  // no occupancy floor: forcing 4 waves per SIMD spills the long transforms
  static constexpr int MIN_WAVES = 1;
};

int col_tile(int ny) {
  switch (ny) {
    case 16: return ColTile0<16>::C;
    case 32: return ColTile0<32>::C;
    case 64: return ColTile0<64>::C;
    case 128: return ColTile0<128>::C;
    case 256: return ColTile0<256>::C;
    case 512: return ColTile0<512>::C;
    case 1024: return ColTile0<1024>::C;
    case 2048: return ColTile0<2048>::C;
  }
  return 0;
}

bool fft_size_supported(int n) { return col_tile(n) > 0; }

// column tiles at each end of kx holding the band x < N/4, x >= 3N/4
int band_tiles_lo(int nx) {
  const int C = col_tile(nx);
  if (C <= 0 || (nx / 4) % C != 0 || nx / 4 < C) return 0;
  return nx / 4 / C;
}

// ------------------------------------------------------------------ rows
template <int N, int RV> struct RowCfg {
  using P = typename std::conditional<RV == 2, AltPlan<N>, Plan<N>>::type;
  static constexpr int T = P::T;
  static constexpr bool WAVE = RV >= 1 && T <= 64;
  static constexpr int SH = RV >= 1 ? 4 : 0;
  static constexpr int LINES = (256 / T) > 0 ? 256 / T : 1;
  static constexpr int LS = LdsLine<SH>::span(N) + 1;
};

// Band-limited lines: the propagator of U.cu:419 is zero for |k| >= kmax,
// about half the Nyquist frequency, so between slices only the kx columns
// x < N/4 and x >= 3N/4 can be non-zero (the host checks this on the FP32
// k-grid before enabling it).  With PRUNE_IN the row pass loads only those
// canonical blocks (the rest is known zero); with PRUNE_OUT it stores only
// them (the column pass never reads the others).
// STORE_AUX: the pass also writes the real-space wave of ROW_INV_MUL_FWD
// (the adjoint reuses it instead of transforming again).
enum { PRUNE_IN = 1, PRUNE_OUT = 2, STORE_AUX = 4 };
template <int E> struct BandBlocks {
  static constexpr int LO = E / 4 > 0 ? E / 4 : 1;
  static __host__ __device__ constexpr bool active(int c) { return c < LO || c >= E - LO; }
};

template <int N, int MODE, int MUL, int FL, int RV>
__global__ void __launch_bounds__((RowCfg<N, RV>::LINES * RowCfg<N, RV>::T)) k_row(RowArgs a) {
  using Cfg = RowCfg<N, RV>;
  using P = typename Cfg::P;
  using S = typename std::conditional<Cfg::WAVE, WaveSync, BlockSync>::type;
  using LL = LdsLine<Cfg::SH>;
  constexpr int T = Cfg::T, E = N / T, LINES = Cfg::LINES, LS = Cfg::LS;
  __shared__ float2 lds[LINES * LS];
  const int tid = threadIdx.x, t = tid % T, ln = tid / T;
  const long long line = (long long)blockIdx.x * LINES + ln;
  const bool ok = line < (long long)a.nbatch * a.ny;
  const int b = ok ? (int)(line / a.ny) : 0;
  const int y = ok ? (int)(line - (long long)b * a.ny) : 0;
  float2 v[E];
  {
    const float2* row = a.src + (long long)b * a.src_bstride + (long long)y * N;
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const bool load = ok && (!(FL & PRUNE_IN) || BandBlocks<E>::active(c));
      v[c] = load ? row[t + c * T] : make_float2(0.f, 0.f);
    }
  }
  const LL L{lds + ln * LS, 1};
  if constexpr (MODE == ROW_INV || MODE == ROW_INV_MUL_FWD) fft_line<P, N, +1, S>(v, t, a.tw, L);
  if constexpr ((FL & STORE_AUX) != 0) {
    // read back only much later (by the adjoint pass): streamed past the caches
    if (ok) {
      float2* row = a.aux + (long long)b * a.dst_bstride + (long long)y * N;
#pragma unroll
      for (int c = 0; c < E; ++c) nt_store(v[c], &row[t + c * T]);
    }
  }
  if constexpr (MUL == RM_T) {
    if (ok) {
      const long long slot = a.t_index ? (long long)a.t_index[b] : (long long)(b / a.t_group);
      const float2* tr = a.t + slot * a.t_bstride + (long long)y * N;
#pragma unroll
      for (int c = 0; c < E; ++c) v[c] = cscale(cmul(v[c], tr[t + c * T]), a.scale);
    }
  } else if constexpr (MUL == RM_PHASE) {
    if (ok) {
      const float* vr = a.V + (long long)(b / a.t_group) * a.t_bstride + (long long)y * N;
#pragma unroll
      for (int c = 0; c < E; ++c) {
        float sn, cs;
        sincosf(a.sigma * vr[t + c * T], &sn, &cs);
        v[c] = cscale(cmul(v[c], make_float2(cs, sn)), a.scale);
      }
    }
  } else {
    if (a.scale != 1.0f) {
#pragma unroll
      for (int c = 0; c < E; ++c) v[c] = cscale(v[c], a.scale);
    }
  }
  if constexpr (MODE == ROW_FWD || MODE == ROW_INV_MUL_FWD) fft_line<P, N, -1, S>(v, t, a.tw, L);
  if (ok) {
    float2* row = a.dst + (long long)b * a.dst_bstride + (long long)y * N;
#pragma unroll
    for (int c = 0; c < E; ++c)
      if (!(FL & PRUNE_OUT) || BandBlocks<E>::active(c)) row[t + c * T] = v[c];
  }
}

// ------------------------------------------------------------------ columns
template <int MUL>
__device__ __forceinline__ float2 kmul(const ColArgs& a, int x, int r, float kxv, float kyv) {
  if constexpr (MUL == CM_SCALE) {
    return make_float2(a.scale, 0.f);
  } else if constexpr (MUL == CM_PROP || MUL == CM_PROPCONJ) {
    // CreatePropsKernel, U.cu:419: inside when sqrt(kx^2 + ky^2) < kmax
    const float k = sqrtf(kxv * kxv + kyv * kyv);
    if (!(k < a.kmax)) return make_float2(0.f, 0.f);
    float2 p = cmul(a.px[x], a.py[r]);
    if (MUL == CM_PROPCONJ) p.y = -p.y;
    return cscale(p, a.scale);
  } else if constexpr (MUL == CM_BAND) {
    // BandLimitKernel, U.cu:1037-1040: keeps |k| <= kmax
    const float k = sqrtf(kxv * kxv + kyv * kyv);
    return make_float2(k <= a.kmax ? a.scale : 0.f, 0.f);
  } else {  // CM_GRAD: gradKernel, U.cu:959-960
    return make_float2(a.scale * -4.f * 3.14159f * 3.14159f * (kxv * kxv + kyv * kyv), 0.f);
  }
}

template <int MODE> struct ColModeTraits {
  static constexpr bool FWD = MODE == COL_FWD || MODE == COL_FWD_MUL_INV || MODE == COL_FWD_MUL || MODE == COL_FWD_DETECT;
  static constexpr bool MULT = MODE == COL_FWD_MUL_INV || MODE == COL_FWD_MUL || MODE == COL_FWD_DETECT || MODE == COL_MUL_INV;
  static constexpr bool INV = MODE == COL_INV || MODE == COL_FWD_MUL_INV || MODE == COL_MUL_INV;
};

// FFT / multiply / IFFT of one column held canonically by T threads
template <class P, int N, int MODE, int MUL, class S, class LL>
__device__ __forceinline__ void col_body(float2 (&v)[N / P::T], int t, int x, const ColArgs& a, const LL& L) {
  using Tr = ColModeTraits<MODE>;
  constexpr int T = P::T, E = N / T;
  if constexpr (Tr::FWD) fft_line<P, N, -1, S>(v, t, a.tw, L);
  // frequencies are read only by modes that need them (scale-only passes may
  // run without k-grids)
  constexpr bool NEEDK = (Tr::MULT && MUL != CM_SCALE);
  float kxv = 0.f;
  if constexpr (NEEDK) kxv = a.kx[x];
  if constexpr (Tr::MULT) {
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int r = t + c * T;
      float kyv = 0.f;
      if constexpr (NEEDK) kyv = a.ky[r];
      v[c] = cmul(v[c], kmul<MUL>(a, x, r, kxv, kyv));
    }
  } else if (a.scale != 1.0f) {
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = cscale(v[c], a.scale);
  }
  if constexpr (Tr::INV) fft_line<P, N, +1, S>(v, t, a.tw, L);
}

// variants 0, 1: the C columns of a tile spread over the workgroup, thread
// (col, t) holds rows t + c T of column col; workgroup barriers
template <int N, int MODE, int MUL, int CV>
__global__ void __launch_bounds__((ColCfg<N, CV>::THREADS), (ColCfg<N, CV>::MIN_WAVES)) k_col(ColArgs a) {
  using Cfg = ColCfg<N, CV>;
  constexpr int T = Cfg::T, E = N / T, C = Cfg::C;
  __shared__ float2 lds[Cfg::LDS];
  const int col = threadIdx.x % C, t = threadIdx.x / C;
  // band-limited runs launch only the tiles at both ends of kx
  const int tile = (a.tiles_lo > 0 && (int)blockIdx.x >= a.tiles_lo) ? (int)blockIdx.x + a.nx / C - 2 * a.tiles_lo
                                                                      : (int)blockIdx.x;
  const int b = blockIdx.y;
  const int x = tile * C + col;
  const long long nx = a.nx;
  float2 v[E];
  {
    const float2* src = a.src + (long long)b * a.src_bstride + x;
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = src[(long long)(t + c * T) * nx];
  }
  using LL = LdsLine<Cfg::SH>;
  const LL L = Cfg::INTERLEAVED ? LL{lds + col, C} : LL{lds + col * Cfg::LSC, 1};
  if constexpr (Cfg::DUAL) {
    const DualLine<LL> D{L, LL{lds + N * C + col, C}};
    col_body<typename Cfg::P, N, MODE, MUL, BlockSync>(v, t, x, a, D);
  } else {
    col_body<typename Cfg::P, N, MODE, MUL, BlockSync>(v, t, x, a, L);
  }
  if constexpr (MODE == COL_FWD_DETECT) {
    const float kxv = a.kx[x];
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int r = t + c * T;
      const float kyv = a.ky[r];
      const float k = sqrtf(kxv * kxv + kyv * kyv);
      const float i2 = v[c].x * v[c].x + v[c].y * v[c].y;
      if (k >= a.kin && k < a.kout) acc += i2;
      if (a.inten) a.inten[(long long)b * a.ny * nx + (long long)r * nx + x] = i2;
    }
    // deterministic tree reduction over the workgroup
    __syncthreads();
    float* red = reinterpret_cast<float*>(lds);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = (C * T) / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[(long long)b * gridDim.x + blockIdx.x] = red[0] * a.det_scale;
    if (a.dst == nullptr) return;
  }
  {
    float2* dst = a.dst + (long long)b * a.dst_bstride + x;
#pragma unroll
    for (int c = 0; c < E; ++c) dst[(long long)(t + c * T) * nx] = v[c];
  }
}

// ------------------------------------------------------------------ dispatch
template <int N, int RV>
static hipError_t row_nv(int mode, int mul, const RowArgs& a, hipStream_t s) {
  constexpr int LINES = RowCfg<N, RV>::LINES, T = RowCfg<N, RV>::T;
  const long long nlines = (long long)a.nbatch * a.ny;
  if (nlines <= 0) return hipSuccess;
  dim3 grid((unsigned)((nlines + LINES - 1) / LINES)), block(LINES * T);
  const int fl = (a.prune_in ? PRUNE_IN : 0) | (a.prune_out ? PRUNE_OUT : 0) | (a.aux ? STORE_AUX : 0);
  if (a.aux && mode != ROW_INV_MUL_FWD) return hipErrorInvalidValue;
#define MSL_ROW(M, U, F) hipLaunchKernelGGL((k_row<N, M, U, F, RV>), grid, block, 0, s, a)
  if (mode == ROW_FWD && mul == RM_NONE && fl == 0) MSL_ROW(ROW_FWD, RM_NONE, 0);
  else if (mode == ROW_FWD && mul == RM_NONE && fl == PRUNE_OUT) MSL_ROW(ROW_FWD, RM_NONE, PRUNE_OUT);
  else if (mode == ROW_FWD && mul == RM_T && fl == 0) MSL_ROW(ROW_FWD, RM_T, 0);
  else if (mode == ROW_FWD && mul == RM_T && fl == PRUNE_OUT) MSL_ROW(ROW_FWD, RM_T, PRUNE_OUT);
  else if (mode == ROW_FWD && mul == RM_PHASE && fl == 0) MSL_ROW(ROW_FWD, RM_PHASE, 0);
  else if (mode == ROW_FWD && mul == RM_PHASE && fl == PRUNE_OUT) MSL_ROW(ROW_FWD, RM_PHASE, PRUNE_OUT);
  else if (mode == ROW_INV && mul == RM_NONE && fl == 0) MSL_ROW(ROW_INV, RM_NONE, 0);
  else if (mode == ROW_INV && mul == RM_NONE && fl == PRUNE_IN) MSL_ROW(ROW_INV, RM_NONE, PRUNE_IN);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == 0) MSL_ROW(ROW_INV_MUL_FWD, RM_T, 0);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == PRUNE_OUT) MSL_ROW(ROW_INV_MUL_FWD, RM_T, PRUNE_OUT);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == (PRUNE_IN | PRUNE_OUT))
    MSL_ROW(ROW_INV_MUL_FWD, RM_T, PRUNE_IN | PRUNE_OUT);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_PHASE && fl == 0) MSL_ROW(ROW_INV_MUL_FWD, RM_PHASE, 0);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_PHASE && fl == (PRUNE_IN | PRUNE_OUT))
    MSL_ROW(ROW_INV_MUL_FWD, RM_PHASE, PRUNE_IN | PRUNE_OUT);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_PHASE && fl == STORE_AUX) MSL_ROW(ROW_INV_MUL_FWD, RM_PHASE, STORE_AUX);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_PHASE && fl == (PRUNE_IN | PRUNE_OUT | STORE_AUX))
    MSL_ROW(ROW_INV_MUL_FWD, RM_PHASE, PRUNE_IN | PRUNE_OUT | STORE_AUX);
  else return hipErrorInvalidValue;
#undef MSL_ROW
  return hipGetLastError();
}

template <int N>
static hipError_t row_n(int mode, int mul, const RowArgs& a, hipStream_t s, int var) {
  if (var == 2) return row_nv<N, 2>(mode, mul, a, s);
  if (var == 1) return row_nv<N, 1>(mode, mul, a, s);
  return row_nv<N, 0>(mode, mul, a, s);
}

template <int N, int CV>
static hipError_t col_nv(int mode, int mul, const ColArgs& a, hipStream_t s) {
  using Cfg = ColCfg<N, CV>;
  constexpr int C = Cfg::C;
  if (a.nbatch <= 0) return hipSuccess;
  if (a.nx % C != 0) return hipErrorInvalidValue;
  if (a.tiles_lo < 0 || 2 * a.tiles_lo > a.nx / ColTile0<N>::C) return hipErrorInvalidValue;
  // tiles_lo counts tiles of the default width; convert to this variant's
  ColArgs b = a;
  if (a.tiles_lo > 0) {
    if ((a.nx / 4) % C != 0) return hipErrorInvalidValue;
    b.tiles_lo = a.nx / 4 / C;
  }
  dim3 grid(b.tiles_lo > 0 ? 2 * b.tiles_lo : a.nx / C, a.nbatch), block(Cfg::THREADS);
  {
#define MSL_COL(M, U) hipLaunchKernelGGL((k_col<N, M, U, CV>), grid, block, 0, s, b)
    if (mode == COL_FWD) MSL_COL(COL_FWD, CM_SCALE);
    else if (mode == COL_INV) MSL_COL(COL_INV, CM_SCALE);
    else if (mode == COL_FWD_MUL_INV && mul == CM_PROP) MSL_COL(COL_FWD_MUL_INV, CM_PROP);
    else if (mode == COL_FWD_MUL_INV && mul == CM_PROPCONJ) MSL_COL(COL_FWD_MUL_INV, CM_PROPCONJ);
    else if (mode == COL_FWD_MUL_INV && mul == CM_BAND) MSL_COL(COL_FWD_MUL_INV, CM_BAND);
    else if (mode == COL_FWD_MUL_INV && mul == CM_GRAD) MSL_COL(COL_FWD_MUL_INV, CM_GRAD);
    else if (mode == COL_FWD_MUL_INV && mul == CM_SCALE) MSL_COL(COL_FWD_MUL_INV, CM_SCALE);
    else if (mode == COL_FWD_MUL && mul == CM_PROP) MSL_COL(COL_FWD_MUL, CM_PROP);
    else if (mode == COL_FWD_MUL && mul == CM_SCALE) MSL_COL(COL_FWD_MUL, CM_SCALE);
    else if (mode == COL_FWD_DETECT && mul == CM_PROP) MSL_COL(COL_FWD_DETECT, CM_PROP);
    else if (mode == COL_FWD_DETECT && mul == CM_SCALE) MSL_COL(COL_FWD_DETECT, CM_SCALE);
    else if (mode == COL_MUL_INV && mul == CM_SCALE) MSL_COL(COL_MUL_INV, CM_SCALE);
    else return hipErrorInvalidValue;
#undef MSL_COL
  }
  return hipGetLastError();
}

template <int N>
static hipError_t col_n(int mode, int mul, const ColArgs& a, hipStream_t s, int var) {
  // the detector pass reduces over the workgroup in an interleaved layout
  if (mode == COL_FWD_DETECT) {
    if (var == 4) return col_nv<N, 4>(mode, mul, a, s);
    return var == 2 ? col_nv<N, 2>(mode, mul, a, s) : col_nv<N, 0>(mode, mul, a, s);
  }
  if (var == 4) return col_nv<N, 4>(mode, mul, a, s);
  if (var == 3) return col_nv<N, 3>(mode, mul, a, s);
  if (var == 2) return col_nv<N, 2>(mode, mul, a, s);
  if (var == 1) return col_nv<N, 1>(mode, mul, a, s);
  return col_nv<N, 0>(mode, mul, a, s);
}

hipError_t launch_row_v(int mode, int mul, const RowArgs& a, hipStream_t s, int var) {
  switch (a.nx) {
    case 16: return row_n<16>(mode, mul, a, s, var);
    case 32: return row_n<32>(mode, mul, a, s, var);
    case 64: return row_n<64>(mode, mul, a, s, var);
    case 128: return row_n<128>(mode, mul, a, s, var);
    case 256: return row_n<256>(mode, mul, a, s, var);
    case 512: return row_n<512>(mode, mul, a, s, var);
    case 1024: return row_n<1024>(mode, mul, a, s, var);
    case 2048: return row_n<2048>(mode, mul, a, s, var);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_col_v(int mode, int mul, const ColArgs& a, hipStream_t s, int var) {
  switch (a.ny) {
    case 16: return col_n<16>(mode, mul, a, s, var);
    case 32: return col_n<32>(mode, mul, a, s, var);
    case 64: return col_n<64>(mode, mul, a, s, var);
    case 128: return col_n<128>(mode, mul, a, s, var);
    case 256: return col_n<256>(mode, mul, a, s, var);
    case 512: return col_n<512>(mode, mul, a, s, var);
    case 1024: return col_n<1024>(mode, mul, a, s, var);
    case 2048: return col_n<2048>(mode, mul, a, s, var);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_row(int mode, int mul, const RowArgs& a, hipStream_t s) {
  return launch_row_v(mode, mul, a, s, g_row_var[lg2(a.nx)]);
}
hipError_t launch_col(int mode, int mul, const ColArgs& a, hipStream_t s) {
  return launch_col_v(mode, mul, a, s, g_col_var[lg2(a.ny)]);
}

}  // namespace msl
