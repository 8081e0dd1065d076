// k_pass.hip -- row and column FFT passes with fused pointwise hooks.
//
// One multislice step over a batch of waves is two passes over HBM
// (transmission multiply of U.cu:2105-2109, FFT/propagator/IFFT of
// U.cu:2117-2122, restated):
//   row pass    IFFT_x -> * t(x, y) -> FFT_x               (ROW_INV_MUL_FWD)
//   column pass FFT_y  -> * P(kx, ky) / (nx ny) -> IFFT_y  (COL_FWD_MUL_INV)
// Between slices a wave stays in the (kx, y) domain; the leading IFFT_x of
// the next row pass returns it to real space in registers.
#include "msl_fft.hpp"
#include "msl_internal.hpp"

namespace msl {

template <int N> struct ColTile { static constexpr int C = (N >= 2048) ? 4 : (N >= 1024 ? 8 : 16); };

int col_tile(int ny) {
  switch (ny) {
    case 16: return ColTile<16>::C;
    case 32: return ColTile<32>::C;
    case 64: return ColTile<64>::C;
    case 128: return ColTile<128>::C;
    case 256: return ColTile<256>::C;
    case 512: return ColTile<512>::C;
    case 1024: return ColTile<1024>::C;
    case 2048: return ColTile<2048>::C;
  }
  return 0;
}

bool fft_size_supported(int n) { return col_tile(n) > 0; }

// ------------------------------------------------------------------ rows
template <int N> struct RowLines { static constexpr int L = (256 / Plan<N>::T) > 0 ? 256 / Plan<N>::T : 1; };

template <int N, int MODE, int MUL, bool SHIFT>
__global__ void __launch_bounds__(RowLines<N>::L * Plan<N>::T) k_row(RowArgs a) {
  constexpr int T = Plan<N>::T, E = N / T, LINES = RowLines<N>::L, LS = N + 1;
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
  const LdsLine L{lds + ln * LS, 1};
  if constexpr (MODE == ROW_INV || MODE == ROW_INV_MUL_FWD) fft_line<N, +1>(v, t, a.tw, L);
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
  if constexpr (MODE == ROW_FWD || MODE == ROW_INV_MUL_FWD) fft_line<N, -1>(v, t, a.tw, L);
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

template <int N, int MODE, int MUL>
__global__ void __launch_bounds__(ColTile<N>::C * Plan<N>::T) k_col(ColArgs a) {
  constexpr int T = Plan<N>::T, E = N / T, C = ColTile<N>::C;
  __shared__ float2 lds[N * C];
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
  const LdsLine L{lds + col, C};
  constexpr bool FWD = MODE == COL_FWD || MODE == COL_FWD_MUL_INV || MODE == COL_FWD_MUL || MODE == COL_FWD_DETECT;
  constexpr bool MULT = MODE == COL_FWD_MUL_INV || MODE == COL_FWD_MUL || MODE == COL_FWD_DETECT || MODE == COL_MUL_INV;
  constexpr bool INV = MODE == COL_INV || MODE == COL_FWD_MUL_INV || MODE == COL_MUL_INV;
  if constexpr (FWD) fft_line<N, -1>(v, t, a.tw, L);
  // frequencies are read only by modes that need them (scale-only passes may
  // run without k-grids)
  constexpr bool NEEDK = (MULT && MUL != CM_SCALE) || MODE == COL_FWD_DETECT;
  float kxv = 0.f;
  if constexpr (NEEDK) kxv = a.kx[x];
  if constexpr (MULT) {
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
  if constexpr (INV) fft_line<N, +1>(v, t, a.tw, L);
  if constexpr (MODE == COL_FWD_DETECT) {
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

// ------------------------------------------------------------------ dispatch
template <int N>
static hipError_t row_n(int mode, int mul, const RowArgs& a, hipStream_t s) {
  constexpr int LINES = RowLines<N>::L, T = Plan<N>::T;
  const long long nlines = (long long)a.nbatch * a.ny;
  if (nlines <= 0) return hipSuccess;
  dim3 grid((unsigned)((nlines + LINES - 1) / LINES)), block(LINES * T);
  const bool sh = a.shift != nullptr;
#define MSL_ROW(M, U, S) hipLaunchKernelGGL((k_row<N, M, U, S>), grid, block, 0, s, a)
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
static hipError_t col_n(int mode, int mul, const ColArgs& a, hipStream_t s) {
  constexpr int C = ColTile<N>::C, T = Plan<N>::T;
  if (a.nbatch <= 0) return hipSuccess;
  if (a.nx % C != 0) return hipErrorInvalidValue;
  dim3 grid(a.nx / C, a.nbatch), block(C * T);
#define MSL_COL(M, U) hipLaunchKernelGGL((k_col<N, M, U>), grid, block, 0, s, a)
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
  return hipGetLastError();
}

hipError_t launch_row(int mode, int mul, const RowArgs& a, hipStream_t s) {
  switch (a.nx) {
    case 16: return row_n<16>(mode, mul, a, s);
    case 32: return row_n<32>(mode, mul, a, s);
    case 64: return row_n<64>(mode, mul, a, s);
    case 128: return row_n<128>(mode, mul, a, s);
    case 256: return row_n<256>(mode, mul, a, s);
    case 512: return row_n<512>(mode, mul, a, s);
    case 1024: return row_n<1024>(mode, mul, a, s);
    case 2048: return row_n<2048>(mode, mul, a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_col(int mode, int mul, const ColArgs& a, hipStream_t s) {
  switch (a.ny) {
    case 16: return col_n<16>(mode, mul, a, s);
    case 32: return col_n<32>(mode, mul, a, s);
    case 64: return col_n<64>(mode, mul, a, s);
    case 128: return col_n<128>(mode, mul, a, s);
    case 256: return col_n<256>(mode, mul, a, s);
    case 512: return col_n<512>(mode, mul, a, s);
    case 1024: return col_n<1024>(mode, mul, a, s);
    case 2048: return col_n<2048>(mode, mul, a, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace msl
