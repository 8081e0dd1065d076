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
//         2 wave-local exchanges, LDS pad 1/8
// column: 0 [row][col] interleaved LDS, workgroup barriers
//         1 [col][row] LDS with pad 1/16, workgroup barriers
//         2 staged: coalesced tile load into [col][row] LDS, then one
//           wavefront per column with wave-local exchanges (16 columns)
//         3 staged with 8 columns per workgroup
static int g_row_var[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static int g_col_var[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

int pass_variant(int kind, int n) { return kind == 0 ? g_row_var[lg2(n)] : g_col_var[lg2(n)]; }
void set_pass_variant(int kind, int n, int v) {
  if (kind == 0) g_row_var[lg2(n)] = v;
  else g_col_var[lg2(n)] = v;
}
int num_pass_variants(int kind) { return kind == 0 ? 3 : 4; }

template <int N> struct ColTile0 { static constexpr int C = (N >= 2048) ? 4 : (N >= 1024 ? 8 : 16); };

template <int N, int CV> struct ColCfg {
  static constexpr int T = Plan<N>::T;
  static constexpr bool STAGED = (CV >= 2) && (T == 64);
  static constexpr int C = STAGED ? (CV == 2 ? (N >= 1024 ? 8 : 16) : (N >= 1024 ? 4 : 8)) : ColTile0<N>::C;
  static constexpr int SH = (CV == 0) ? 0 : 4;
  // [col][row] line stride in float2: span + padding so consecutive columns
  // start two banks apart (stride = 2 mod 32 float2)
  static constexpr int SPAN = LdsLine<SH>::span(N);
  static constexpr int LSC = SPAN + ((2 - SPAN % 32) + 32) % 32;
  static constexpr int LDS = (CV == 0) ? N * C : C * LSC;
  static constexpr int THREADS = C * T;
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

// ------------------------------------------------------------------ rows
template <int N, int RV> struct RowCfg {
  static constexpr int T = Plan<N>::T;
  static constexpr bool WAVE = RV >= 1 && T <= 64;
  static constexpr int SH = RV == 1 ? 4 : (RV == 2 ? 3 : 0);
  static constexpr int LINES = (256 / T) > 0 ? 256 / T : 1;
  static constexpr int LS = LdsLine<SH>::span(N) + 1;
};

template <int N, int MODE, int MUL, bool SHIFT, int RV>
__global__ void __launch_bounds__((RowCfg<N, RV>::LINES * Plan<N>::T)) k_row(RowArgs a) {
  using Cfg = RowCfg<N, RV>;
  using S = typename std::conditional<Cfg::WAVE, WaveSync, BlockSync>::type;
  using LL = LdsLine<Cfg::SH>;
  constexpr int T = Plan<N>::T, E = N / T, LINES = Cfg::LINES, LS = Cfg::LS;
  __shared__ float2 lds[LINES * LS];
  const int tid = threadIdx.x, t = tid % T, ln = tid / T;
  const long long line = (long long)blockIdx.x * LINES + ln;
  const bool ok = line < (long long)a.nbatch * a.ny;
  const int b = ok ? (int)(line / a.ny) : 0;
  const int y = ok ? (int)(line - (long long)b * a.ny) : 0;
  float2 v[E];
  if (ok) {
    const float2* src = a.src + (long long)b * a.src_bstride;
    if constexpr (SHIFT) {
      const int2 sh = a.shift[b];
      const float2* row = src + (long long)((y + sh.y) & (a.ny - 1)) * N;
#pragma unroll
      for (int c = 0; c < E; ++c) v[c] = row[(t + c * T + sh.x) & (N - 1)];
    } else {
      const float2* row = src + (long long)y * N;
#pragma unroll
      for (int c = 0; c < E; ++c) v[c] = row[t + c * T];
    }
  } else {
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = make_float2(0.f, 0.f);
  }
  const LL L{lds + ln * LS, 1};
  if constexpr (MODE == ROW_INV || MODE == ROW_INV_MUL_FWD) fft_line<Plan<N>, N, +1, S>(v, t, a.tw, L);
  if constexpr (MUL == RM_T) {
    if (ok) {
      const float2* tr = a.t + (long long)(b / a.t_group) * a.t_bstride + (long long)y * N;
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
  if constexpr (MODE == ROW_FWD || MODE == ROW_INV_MUL_FWD) fft_line<Plan<N>, N, -1, S>(v, t, a.tw, L);
  if (ok) {
    float2* row = a.dst + (long long)b * a.dst_bstride + (long long)y * N;
#pragma unroll
    for (int c = 0; c < E; ++c) row[t + c * T] = v[c];
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
template <int N, int MODE, int MUL, class S, class LL>
__device__ __forceinline__ void col_body(float2 (&v)[N / Plan<N>::T], int t, int x, const ColArgs& a, const LL& L) {
  using Tr = ColModeTraits<MODE>;
  constexpr int T = Plan<N>::T, E = N / T;
  if constexpr (Tr::FWD) fft_line<Plan<N>, N, -1, S>(v, t, a.tw, L);
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
  if constexpr (Tr::INV) fft_line<Plan<N>, N, +1, S>(v, t, a.tw, L);
}

// variants 0, 1: the C columns of a tile spread over the workgroup, thread
// (col, t) holds rows t + c T of column col; workgroup barriers
template <int N, int MODE, int MUL, int CV>
__global__ void __launch_bounds__((ColCfg<N, CV>::THREADS)) k_col(ColArgs a) {
  using Cfg = ColCfg<N, CV>;
  constexpr int T = Plan<N>::T, E = N / T, C = Cfg::C;
  __shared__ float2 lds[Cfg::LDS];
  const int col = threadIdx.x % C, t = threadIdx.x / C;
  const int tile = blockIdx.x, b = blockIdx.y;
  const int x = tile * C + col;
  const long long nx = a.nx;
  float2 v[E];
  {
    const float2* src = a.src + (long long)b * a.src_bstride + x;
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = src[(long long)(t + c * T) * nx];
  }
  using LL = LdsLine<Cfg::SH>;
  const LL L = (CV == 0) ? LL{lds + col, C} : LL{lds + col * Cfg::LSC, 1};
  col_body<N, MODE, MUL, BlockSync>(v, t, x, a, L);
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
    if (threadIdx.x == 0) a.partial[(long long)b * (a.nx / C) + tile] = red[0] * a.det_scale;
    if (a.dst == nullptr) return;
  }
  {
    float2* dst = a.dst + (long long)b * a.dst_bstride + x;
#pragma unroll
    for (int c = 0; c < E; ++c) dst[(long long)(t + c * T) * nx] = v[c];
  }
}

// variants 2, 3 (T == 64): the tile moves between HBM and LDS with
// coalesced C-column rows; in between, wavefront w transforms column w on
// its own with wave-local exchanges, so a pass has two workgroup barriers.
template <int N, int MODE, int MUL, int CV>
__global__ void __launch_bounds__((ColCfg<N, CV>::THREADS)) k_col_staged(ColArgs a) {
  using Cfg = ColCfg<N, CV>;
  static_assert(Cfg::STAGED, "staged layout needs one wavefront per column");
  constexpr int T = 64, E = N / T, C = Cfg::C, LSC = Cfg::LSC, RPT = (C * T) / C;  // rows per load step
  using LL = LdsLine<Cfg::SH>;
  __shared__ float2 lds[Cfg::LDS];
  const int tid = threadIdx.x;
  const int tile = blockIdx.x, b = blockIdx.y;
  const long long nx = a.nx;
  // coalesced load: thread (lc, r0) moves rows r0 + c RPT of column lc
  {
    const int lc = tid % C, r0 = tid / C;
    const float2* src = a.src + (long long)b * a.src_bstride + tile * C + lc;
    float2 tmp[N / RPT];
#pragma unroll
    for (int c = 0; c < N / RPT; ++c) tmp[c] = src[(long long)(r0 + c * RPT) * nx];
#pragma unroll
    for (int c = 0; c < N / RPT; ++c) lds[lc * LSC + LL::pad(r0 + c * RPT)] = tmp[c];
  }
  __syncthreads();
  const int w = tid / T, t = tid % T;
  const LL L{lds + w * LSC, 1};
  float2 v[E];
#pragma unroll
  for (int c = 0; c < E; ++c) v[c] = L.at(t + c * T);
  WaveSync::sync();
  col_body<N, MODE, MUL, WaveSync>(v, t, tile * C + w, a, L);
  WaveSync::sync();
#pragma unroll
  for (int c = 0; c < E; ++c) L.at(t + c * T) = v[c];
  __syncthreads();
  {
    const int lc = tid % C, r0 = tid / C;
    float2* dst = a.dst + (long long)b * a.dst_bstride + tile * C + lc;
#pragma unroll
    for (int c = 0; c < N / RPT; ++c) dst[(long long)(r0 + c * RPT) * nx] = lds[lc * LSC + LL::pad(r0 + c * RPT)];
  }
}

// ------------------------------------------------------------------ dispatch
template <int N, int RV>
static hipError_t row_nv(int mode, int mul, const RowArgs& a, hipStream_t s) {
  constexpr int LINES = RowCfg<N, RV>::LINES, T = Plan<N>::T;
  const long long nlines = (long long)a.nbatch * a.ny;
  if (nlines <= 0) return hipSuccess;
  dim3 grid((unsigned)((nlines + LINES - 1) / LINES)), block(LINES * T);
  const bool sh = a.shift != nullptr;
#define MSL_ROW(M, U, SH) hipLaunchKernelGGL((k_row<N, M, U, SH, RV>), grid, block, 0, s, a)
  if (mode == ROW_FWD && mul == RM_NONE) { if (sh) MSL_ROW(ROW_FWD, RM_NONE, true); else MSL_ROW(ROW_FWD, RM_NONE, false); }
  else if (mode == ROW_FWD && mul == RM_T) { if (sh) MSL_ROW(ROW_FWD, RM_T, true); else MSL_ROW(ROW_FWD, RM_T, false); }
  else if (mode == ROW_FWD && mul == RM_PHASE) { if (sh) MSL_ROW(ROW_FWD, RM_PHASE, true); else MSL_ROW(ROW_FWD, RM_PHASE, false); }
  else if (mode == ROW_INV && mul == RM_NONE && !sh) MSL_ROW(ROW_INV, RM_NONE, false);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && !sh) MSL_ROW(ROW_INV_MUL_FWD, RM_T, false);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_PHASE && !sh) MSL_ROW(ROW_INV_MUL_FWD, RM_PHASE, false);
  else return hipErrorInvalidValue;
#undef MSL_ROW
  return hipGetLastError();
}

template <int N>
static hipError_t row_n(int mode, int mul, const RowArgs& a, hipStream_t s, int var) {
  switch (var) {
    case 1: return row_nv<N, 1>(mode, mul, a, s);
    case 2: return row_nv<N, 2>(mode, mul, a, s);
    default: return row_nv<N, 0>(mode, mul, a, s);
  }
}

template <int N, int CV>
static hipError_t col_nv(int mode, int mul, const ColArgs& a, hipStream_t s) {
  using Cfg = ColCfg<N, CV>;
  constexpr int C = Cfg::C;
  if (a.nbatch <= 0) return hipSuccess;
  if (a.nx % C != 0) return hipErrorInvalidValue;
  dim3 grid(a.nx / C, a.nbatch), block(Cfg::THREADS);
  if constexpr (Cfg::STAGED) {
#define MSL_COLS(M, U) hipLaunchKernelGGL((k_col_staged<N, M, U, CV>), grid, block, 0, s, a)
    if (mode == COL_FWD) MSL_COLS(COL_FWD, CM_SCALE);
    else if (mode == COL_INV) MSL_COLS(COL_INV, CM_SCALE);
    else if (mode == COL_FWD_MUL_INV && mul == CM_PROP) MSL_COLS(COL_FWD_MUL_INV, CM_PROP);
    else if (mode == COL_FWD_MUL_INV && mul == CM_PROPCONJ) MSL_COLS(COL_FWD_MUL_INV, CM_PROPCONJ);
    else if (mode == COL_FWD_MUL_INV && mul == CM_BAND) MSL_COLS(COL_FWD_MUL_INV, CM_BAND);
    else if (mode == COL_FWD_MUL_INV && mul == CM_GRAD) MSL_COLS(COL_FWD_MUL_INV, CM_GRAD);
    else if (mode == COL_FWD_MUL_INV && mul == CM_SCALE) MSL_COLS(COL_FWD_MUL_INV, CM_SCALE);
    else if (mode == COL_FWD_MUL && mul == CM_PROP) MSL_COLS(COL_FWD_MUL, CM_PROP);
    else if (mode == COL_FWD_MUL && mul == CM_SCALE) MSL_COLS(COL_FWD_MUL, CM_SCALE);
    else if (mode == COL_MUL_INV && mul == CM_SCALE) MSL_COLS(COL_MUL_INV, CM_SCALE);
    else return hipErrorInvalidValue;
#undef MSL_COLS
  } else {
#define MSL_COL(M, U) hipLaunchKernelGGL((k_col<N, M, U, CV>), grid, block, 0, s, a)
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
  // the detector pass reduces over the workgroup in the interleaved layout
  if (mode == COL_FWD_DETECT) return col_nv<N, 0>(mode, mul, a, s);
  if constexpr (Plan<N>::T == 64) {
    if (var == 2) return col_nv<N, 2>(mode, mul, a, s);
    if (var == 3) return col_nv<N, 3>(mode, mul, a, s);
  }
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
