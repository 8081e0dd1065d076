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

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "msl_fft.hpp"
#include "msl_internal.hpp"

namespace msl {

// the launch structures as device code reads them from tables in memory
// (pointers in the global address space, same layout)
struct RowArgsG {
  MSL_ROW_ARGS(MSL_GLOBAL)
};
struct ColArgsG {
  MSL_COL_ARGS(MSL_GLOBAL)
};
static_assert(sizeof(RowArgsG) == sizeof(RowArgs) && sizeof(ColArgsG) == sizeof(ColArgs), "layouts differ");
template <class A> struct GlobalView;
template <> struct GlobalView<RowArgs> { using T = RowArgsG; };
template <> struct GlobalView<ColArgs> { using T = ColArgsG; };

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
//         3 as 1 with 512-thread workgroups (twice the lines)
//         4 a workgroup per row looping over waves, transmission in registers
//           (k_row_pl; IFFT -> t -> FFT only, the best of the others
//           elsewhere)
// column: 0 [row][col] interleaved LDS, workgroup barriers
//         1 [col][row] LDS with pad 1/16, workgroup barriers
//         2 as 0 with the alternative plan
//         3 as 0 with half as many columns per workgroup (more workgroups
//           per CU to overlap the barriers)
//         4 as 0 with two LDS copies used by alternate exchanges (one
//           barrier per exchange instead of two)
//         5 a workgroup per tile looping over waves, propagator factors in
//           registers (k_col_pl; the propagation pass only, the length's
//           default elsewhere)
//         6 as 7 with each column held by consecutive threads of one wave
//           (wave-local exchanges, no workgroup barriers; narrower row
//           segments per load instruction).  It replaced a layout that
//           staged the tile through LDS for the same purpose (slower).
//         7 as 2 with half as many columns per workgroup
// (a staged column layout with one wavefront per column and an LDS pad of
// 1/8 were measured slower or equal on MI355X and were removed)
// defaults from msl_autotune_passes on MI355X (band-pruned passes, 1-4 GiB
// of waves): rows of 512 take the alternative plan (0.67 vs 0.71 ms per 1024
// waves for 1), rows of 1024 the 512-thread workgroups (0.81 vs 0.91 ms per
// 256 waves); rows of 512 and 1024 then went to variant 4 (one workgroup per
// row looping over waves; 512: 0.58 vs 0.69 ms per 1024 waves for 2, 1024:
// 0.78 vs 0.85 ms per 256 waves for 3; 1.131 vs 1.080e6 probe-slice/s on the
// 512^2 STEM workload).  Columns: 512 half-width tiles on the alternative plan (0.48
// vs 0.51 ms per 1024 waves for half-width tiles on the default plan, 1 %
// more probe-slices per second on the 512^2 STEM workload); 256, 1024 and
// 2048 variant 0 (1024: 0.54 vs 0.62 ms per 256 waves for the half-width
// tiles once the inverse skipped the propagator's zero rows).  At 2048 the
// loops over waves won once they existed: rows variant 4 (0.447-0.469 vs
// 0.525-0.531 ms per 32 waves for variant 1), columns variant 5 (0.376-0.378
// vs 0.440-0.447 ms for variant 0); at 256 variants 1 and 0 stay fastest.
static int g_row_var[12] = {0, 0, 0, 0, 1, 1, 1, 1, 1, 4, 4, 4};
static int g_col_var[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 7, 0, 5};

int pass_variant(int kind, int n) { return kind == 0 ? g_row_var[lg2(n)] : g_col_var[lg2(n)]; }
void set_pass_variant(int kind, int n, int v) {
  if (kind == 0) g_row_var[lg2(n)] = v;
  else g_col_var[lg2(n)] = v;
}
int num_pass_variants(int kind) { return kind == 0 ? 5 : 8; }

template <int N> struct ColTile0 { static constexpr int C = (N >= 2048) ? 4 : (N >= 1024 ? 8 : 16); };

template <int N, int CV> struct ColCfg {
  using P = typename std::conditional<CV == 2 || CV == 6 || CV == 7, AltPlan<N>, Plan<N>>::type;
  // variant 6: consecutive threads of a wave hold one column (T <= 64), so the
  // exchanges synchronise the wave only; loads take C / (64 / T) columns'
  // worth of each row per instruction instead of the whole tile row
  static constexpr bool WAVECOL = CV == 6 && P::T <= 64 && 64 % P::T == 0;
  static constexpr bool INTERLEAVED = CV != 1 && !WAVECOL;
  static constexpr int T = P::T;
  static constexpr int C = (CV == 3 || CV == 6 || CV == 7) ? ColTile0<N>::C / 2 : ColTile0<N>::C;
  static constexpr int SH = INTERLEAVED ? 0 : 4;
  // [col][row] line stride in float2: span + padding so consecutive columns
  // start two banks apart (stride = 2 mod 32 float2)
  static constexpr int SPAN = LdsLine<SH>::span(N);
  static constexpr int LSC = SPAN + ((2 - SPAN % 32) + 32) % 32;
  static constexpr bool DUAL = CV == 4;
  static constexpr int LDS = INTERLEAVED ? N * C * (DUAL ? 2 : 1) : C * LSC;
  static constexpr int THREADS = C * T;
  // occupancy floor only for half-width tiles of short columns (E <= 8): six
  // waves per SIMD (80 VGPRs, three workgroups per CU instead of two) ran 5 %
  // faster at 512; the long transforms (E = 16) spill under any floor, and
  // the full-width tiles ran slower with it
  static constexpr int MIN_WAVES = (CV == 3 && N / T <= 8) ? 6 : 1;
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

// Workgroups are dispatched round-robin over the XCDs, each with its own L2.
// A column tile narrower than 16 columns reads part of each 128-byte line;
// the neighbouring tile reads the rest.  Remapping the linear workgroup id so
// that every XCD walks a contiguous range of logical ids puts neighbouring
// tiles in one L2 at about the same time, so each line is fetched once.
constexpr int kXcds = 8;
__device__ __forceinline__ long long xcd_remap(long long id, long long total) {
  const long long whole = total - total % kXcds;
  if (id >= whole) return id;
  return (id % kXcds) * (whole / kXcds) + id / kXcds;
}

// ------------------------------------------------------------------ rows
template <int N, int RV> struct RowCfg {
  using P = typename std::conditional<RV == 2, AltPlan<N>, Plan<N>>::type;
  static constexpr int T = P::T;
  static constexpr bool WAVE = RV >= 1 && T <= 64;
  static constexpr int SH = RV >= 1 ? 4 : 0;
  // variant 3: 512-thread workgroups (eight one-wave lines at 512 and 1024:
  // two workgroups, sixteen waves per CU at 1024, where four-line
  // workgroups of 43 KB of LDS fit three)
  static constexpr int LINES = ((RV == 3 ? 512 : 256) / T) > 0 ? (RV == 3 ? 512 : 256) / T : 1;
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
// T_SLOTS (row loop only): each wave reads the transmission of its own slot
// (RowArgs::t_index or b / t_group) instead of one shared by the launch
enum { PRUNE_IN = 1, PRUNE_OUT = 2, STORE_AUX = 4, T_SLOTS = 8 };
template <int E> struct BandBlocks {
  static constexpr int LO = E / 4 > 0 ? E / 4 : 1;
  static __host__ __device__ constexpr bool active(int c) { return c < LO || c >= E - LO; }
};

// The row pass body: the LINES lines of workgroup-sized block blk; lds holds
// LINES * LS elements, tw the N twiddles staged in LDS.  k_row (one length
// per launch, arguments in kernel registers) and k_row_group (several
// lengths per launch, arguments read from a table in device memory) both
// expand this one body.
template <int N, int MODE, int MUL, int FL, int RV, class A>
__device__ __forceinline__ void row_block(const A& a, long long blk, float2* lds, const float2* stw) {
  using Cfg = RowCfg<N, RV>;
  using P = typename Cfg::P;
  using S = typename std::conditional<Cfg::WAVE, WaveSync, BlockSync>::type;
  using LL = LdsLine<Cfg::SH>;
  constexpr int T = Cfg::T, E = N / T, LINES = Cfg::LINES, LS = Cfg::LS;
  const int tid = threadIdx.x, t = tid % T, ln = tid / T;
  // The block's first line splits into wave b0 and row on the scalar unit
  // (the block index is uniform; the grid holds no block past the last
  // line, so b0 < nbatch); a thread's own line is db waves further (db = 0
  // unless LINES straddles a wave), with no per-thread 64-bit division.
  const long long first = blk * LINES;
  const long long b0 = first / a.ny;
  int y = (int)(first - b0 * a.ny) + ln;
  int db = 0;
  while (y >= a.ny) {  // at most LINES / ny + 1 steps; none when ny % LINES == 0
    y -= a.ny;
    ++db;
  }
  const bool ok = b0 + db < a.nbatch;
  const int b = ok ? (int)b0 + db : 0;
  // element offset of this thread's row in an array of waves of stride bs
  auto roff = [&](long long bs) { return b0 * bs + ((long long)db * bs + (long long)y * N); };
  float2 v[E];
#pragma unroll
  for (int c = 0; c < E; ++c) v[c] = make_float2(0.f, 0.f);
  if (ok) {
    const float2* row = a.src + roff(a.src_bstride);
#pragma unroll
    for (int c = 0; c < E; ++c)
      if (!(FL & PRUNE_IN) || BandBlocks<E>::active(c)) v[c] = row[t + c * T];
  }
  const LL L{lds + ln * LS, 1};
  constexpr bool ZIN = (FL & PRUNE_IN) != 0 && E % 4 == 0 && P::R1 % 4 == 0;  // first transform's input
  if constexpr (MODE == ROW_INV || MODE == ROW_INV_MUL_FWD || MODE == ROW_INV_FD_FWD)
    fft_line<P, N, +1, S, LL, ZIN>(v, t, stw, L);
  if constexpr ((FL & STORE_AUX) != 0) {
    // read back only much later (by the adjoint pass): streamed past the caches
    if (ok) {
      float2* row = a.aux + roff(a.dst_bstride);
#pragma unroll
      for (int c = 0; c < E; ++c) nt_store(v[c], &row[t + c * T]);
    }
  }
  if constexpr (MODE == ROW_INV_FD_FWD) {
    // finiteDifferenceKernel, U.cu:995-1023: v holds the Laplacian G of psi
    if (ok) {
      const long long o = roff(a.dst_bstride);
      const long long slot = a.t_index ? (long long)a.t_index[b] : (long long)(b / a.t_group);
      const float2* vr = a.t + slot * a.t_bstride + (long long)y * N;
      const float dz = a.fd_dz, wavel = a.fd_wavel;
      const float al = 2.f * 3.14159f * dz / wavel;
      const float den = 1.f + al * al;
      const float2 rcp = make_float2(1.f / den, -al / den);                    // 1 / (1 + i a)
      const float2 ratio = make_float2((1.f - al * al) / den, -2.f * al / den);  // (1 - i a) / (1 + i a)
      // in groups of Q elements: the loads of one group are issued together,
      // and fewer of them are live at once
      constexpr int Q = E > 8 ? 8 : E;
#pragma unroll
      for (int c0 = 0; c0 < E; c0 += Q) {
        float2 pm[Q], p0[Q];
        float vv[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          pm[q] = a.fd_pm[o + t + (c0 + q) * T];
          p0[q] = a.fd_p0[o + t + (c0 + q) * T];
          vv[q] = vr[t + (c0 + q) * T].x;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const float2 p = p0[q], g = v[c0 + q];
          const float w = dz * dz * 4.f * 3.14159f * a.sigma * vv[q] / wavel;
          const float2 u = make_float2(2.f * p.x - dz * dz * g.x - w * p.x, 2.f * p.y - dz * dz * g.y - w * p.y);
          v[c0 + q] = csub(cmul(rcp, u), cmul(ratio, pm[q]));
        }
        asm volatile("" ::: "memory");
      }
    }
  } else if constexpr (MUL == RM_SEED) {
    // G = 2 dL/dI psi of an intensity loss (dL/dRe + i dL/dIm), in the row's
    // own position of the seed array
    if (ok) {
      const float* sr = a.seed + roff(a.src_bstride);
#pragma unroll
      for (int c = 0; c < E; ++c) v[c] = cscale(v[c], 2.f * sr[t + c * T] * a.scale);
    }
  } else if constexpr (MUL == RM_T) {
    if (ok) {
      // slot of wave b0 on the scalar unit; per thread only past a wave edge
      long long slot = a.t_index ? (long long)a.t_index[b0] : b0 / a.t_group;
      if (db != 0) slot = a.t_index ? (long long)a.t_index[b] : (long long)(b / a.t_group);
      const float2* tr = a.t + slot * a.t_bstride + (long long)y * N;
#pragma unroll
      for (int c = 0; c < E; ++c) v[c] = cscale(cmul(v[c], tr[t + c * T]), a.scale);
    }
  } else {
    if (a.scale != 1.0f) {
#pragma unroll
      for (int c = 0; c < E; ++c) v[c] = cscale(v[c], a.scale);
    }
  }
  if constexpr (MODE == ROW_FWD) fft_line<P, N, -1, S, LL, ZIN>(v, t, stw, L);
  if constexpr (MODE == ROW_INV_MUL_FWD || MODE == ROW_INV_FD_FWD) fft_line<P, N, -1, S>(v, t, stw, L);
  if (ok) {
    float2* row = a.dst + roff(a.dst_bstride);
#pragma unroll
    for (int c = 0; c < E; ++c)
      if (!(FL & PRUNE_OUT) || BandBlocks<E>::active(c)) row[t + c * T] = v[c];
  }
}

// Logical block (wave b, row block yb) = b ybs + yb of hardware workgroup h.
// Each XCD takes a contiguous range of logical blocks (xcd_remap), and inside
// it the same row block of kRowWaves consecutive waves runs back to back:
// those waves read the same transmission rows, which then come from the L2
// once instead of once per wave (a 2048^2 slice, 32 MiB, does not stay in an
// XCD's L2 between one wave and the next).
constexpr int kRowWaves = 8;
__device__ __forceinline__ long long row_interleave(long long r, int nbatch, int ybs) {
  const long long span = (long long)kRowWaves * ybs;
  const long long full = (long long)(nbatch / kRowWaves) * span;  // blocks of complete wave groups
  if (r >= full) return r;
  const long long grp = r / span, rem = r - grp * span;
  const int yb = (int)(rem / kRowWaves), bi = (int)(rem - (long long)yb * kRowWaves);
  return (grp * kRowWaves + bi) * ybs + yb;
}
__device__ __forceinline__ long long row_order(long long h, long long total, int nbatch, int ybs) {
  return row_interleave(xcd_remap(h, total), nbatch, ybs);
}

template <int N, int MODE, int MUL, int FL, int RV>
__global__ void __launch_bounds__((RowCfg<N, RV>::LINES * RowCfg<N, RV>::T)) k_row(RowArgs a) {
  __shared__ float2 lds[RowCfg<N, RV>::LINES * RowCfg<N, RV>::LS];
  __shared__ float2 stw[N];  // twiddles staged in LDS: no global loads inside the transforms
  for (int i = threadIdx.x; i < N; i += RowCfg<N, RV>::LINES * RowCfg<N, RV>::T) stw[i] = a.tw[i];
  __syncthreads();
  long long blk = blockIdx.x;
  // wave-interleaved, XCD-aware order where a shared slice outgrows the L2
  // (measured slower than the plain order at 1024 and below, where the slice
  // stays in the L2 anyway)
  if constexpr (N >= 2048 && MUL != RM_NONE)
    if (N % RowCfg<N, RV>::LINES == 0 && a.ny == N)
      blk = row_order(blockIdx.x, gridDim.x, a.nbatch, N / RowCfg<N, RV>::LINES);
  row_block<N, MODE, MUL, FL, RV>(a, blk, lds, stw);
}

// Variant 4, for IFFT_x -> t -> FFT_x: a workgroup owns row y of LINES x
// per_line waves.  Each thread keeps its E transmission values in registers
// for all of them (the other layouts load them once per wave), and requests
// the next wave's row before transforming the current one.  With one
// transmission shared by the launch (a batched STEM chunk, the
// differentiable forward) they are loaded once; with per-wave slots
// (T_SLOTS: the serving engine's batches of several jobs, from 256 up) a
// line reloads them when its next wave reads another slot.
template <int N> struct RowPL {
  using P = typename std::conditional<(N <= 512), AltPlan<N>, Plan<N>>::type;
  static constexpr int T = P::T;
  static constexpr bool WAVE = T <= 64;
  static constexpr int SH = 4;
  static constexpr int LINES = (256 / T) > 0 ? 256 / T : 1;
  static constexpr int LS = LdsLine<SH>::span(N) + 1;
};

template <int N, int FL>
__global__ void __launch_bounds__((RowPL<N>::LINES * RowPL<N>::T)) k_row_pl(RowArgs a, int per_line) {
  using Cfg = RowPL<N>;
  using P = typename Cfg::P;
  using S = typename std::conditional<Cfg::WAVE, WaveSync, BlockSync>::type;
  using LL = LdsLine<Cfg::SH>;
  constexpr int T = Cfg::T, E = N / T, LINES = Cfg::LINES, LS = Cfg::LS, NT = LINES * T;
  __shared__ float2 lds[LINES * LS];
  __shared__ float2 stw[N];
  const int tid = threadIdx.x, t = tid % T, ln = tid / T;
  for (int i = tid; i < N; i += NT) stw[i] = a.tw[i];
  const int y = blockIdx.x, g = blockIdx.y;
  // lines past the batch read wave 0 and store nothing
  auto wave_of = [&](int q) { return (g * per_line + q) * LINES + ln; };
  // transmission slot of wave b (T_SLOTS; otherwise the launch's one row)
  auto slot_of = [&](int b) -> long long {
    const int bb = b < a.nbatch ? b : 0;
    return a.t_index ? (long long)a.t_index[bb] : (long long)(bb / a.t_group);
  };
  float2 tr[E];
  auto load_t = [&](long long slot) {
    const float2* trow = a.t + slot * a.t_bstride + (long long)y * N;
#pragma unroll
    for (int c = 0; c < E; ++c) tr[c] = trow[t + c * T];
    // the scale goes into the factors once (every caller passes 1: the
    // products are those of the unscaled form)
    if (a.scale != 1.f) {
#pragma unroll
      for (int c = 0; c < E; ++c) tr[c] = cscale(tr[c], a.scale);
    }
  };
  long long cur = (FL & T_SLOTS) ? slot_of(wave_of(0)) : 0;
  load_t(cur);
  __syncthreads();
  const LL L{lds + ln * LS, 1};
  const float2 zero = make_float2(0.f, 0.f);
  auto load = [&](float2 (&v)[E], int b) {
    const float2* row = a.src + (long long)(b < a.nbatch ? b : 0) * a.src_bstride + (long long)y * N;
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = (!(FL & PRUNE_IN) || BandBlocks<E>::active(c)) ? row[t + c * T] : zero;
  };
  constexpr bool ZIN = (FL & PRUNE_IN) != 0 && E % 4 == 0 && P::R1 % 4 == 0;
  float2 vn[E];
  load(vn, wave_of(0));
  for (int q = 0; q < per_line; ++q) {
    const int b = wave_of(q);
    float2 v[E];
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = vn[c];
    if (q + 1 < per_line) load(vn, wave_of(q + 1));
    fft_line<P, N, +1, S, LL, ZIN>(v, t, stw, L);
    if constexpr ((FL & STORE_AUX) != 0) {
      // the real-space wave for the adjoint pass, streamed past the caches
      if (b < a.nbatch) {
        float2* row = a.aux + (long long)b * a.dst_bstride + (long long)y * N;
#pragma unroll
        for (int c = 0; c < E; ++c) nt_store(v[c], &row[t + c * T]);
      }
    }
    if constexpr ((FL & T_SLOTS) != 0) {
      // a line's waves lie LINES apart: a new job's wave brings its row
      const long long sb = slot_of(b);
      if (sb != cur) {
        cur = sb;
        load_t(sb);
      }
    }
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = cmul(v[c], tr[c]);
    fft_line<P, N, -1, S>(v, t, stw, L);
    if (b < a.nbatch) {
      float2* row = a.dst + (long long)b * a.dst_bstride + (long long)y * N;
#pragma unroll
      for (int c = 0; c < E; ++c)
        if (!(FL & PRUNE_OUT) || BandBlocks<E>::active(c)) row[t + c * T] = v[c];
    }
  }
}

// ------------------------------------------------------------------ columns
// CreatePropsKernel, U.cu:419: the propagator is non-zero when
// sqrt(kx^2 + ky^2) < kmax
__device__ __forceinline__ bool in_prop_band(float kxv, float kyv, float kmax) {
  return sqrtf(kxv * kxv + kyv * kyv) < kmax;
}

// Per column x the rows inside the propagator band are r <= rows[x].x and
// r >= rows[x].y: |ky| grows from row 0 up to N/2 and from row N-1 down to
// N/2+1, and the FP32 test above is monotone in |ky|.  bad[x] != 0 flags a
// column where that shape does not hold.
__global__ void k_prop_rows(const float* kx, const float* ky, int nx, int ny, float kmax, int2* rows, int* bad) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const float kxv = kx[x];
  int hi = -1;
  while (hi + 1 < ny && in_prop_band(kxv, ky[hi + 1], kmax)) ++hi;
  int lo = ny;
  while (lo - 1 > hi && in_prop_band(kxv, ky[lo - 1], kmax)) --lo;
  int stray = 0;
  for (int r = hi + 1; r < lo; ++r) stray |= in_prop_band(kxv, ky[r], kmax) ? 1 : 0;
  rows[x] = make_int2(hi, lo);
  bad[x] = stray;
}

hipError_t make_prop_rows(const float* kx, const float* ky, int nx, int ny, float kmax, int2* rows, hipStream_t s) {
  if (nx <= 0 || ny <= 0 || !kx || !ky || !rows) return hipErrorInvalidValue;
  int* bad = nullptr;
  hipError_t e = hipMalloc(&bad, nx * sizeof(int));
  if (e != hipSuccess) return e;
  std::vector<int> h(nx, 1);
  hipLaunchKernelGGL(k_prop_rows, dim3((nx + 63) / 64), dim3(64), 0, s, kx, ky, nx, ny, kmax, rows, bad);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), bad, nx * sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(bad);
  if (e != hipSuccess) return e;
  for (int x = 0; x < nx; ++x)
    if (h[x]) return hipErrorInvalidValue;
  return hipSuccess;
}

template <int MUL, class A>
__device__ __forceinline__ float2 kmul(const A& a, int x, int r, float kxv, float kyv, int2 pr) {
  if constexpr (MUL == CM_SCALE) {
    return make_float2(a.scale, 0.f);
  } else if constexpr (MUL == CM_PROP || MUL == CM_PROPCONJ || MUL == CM_PROP_ZB || MUL == CM_PROPCONJ_ZB) {
    // not reached: col_body applies the propagator itself
    return make_float2(0.f, 0.f);
  } else if constexpr (MUL == CM_BAND) {
    // BandLimitKernel, U.cu:1037-1040: keeps |k| <= kmax
    const float k = sqrtf(kxv * kxv + kyv * kyv);
    return make_float2(k <= a.kmax ? a.scale : 0.f, 0.f);
  } else if constexpr (MUL == CM_BANDGRAD) {
    // BandLimitKernel then gradKernel (U.cu:1037-1040, 959-960) as one factor
    const float k = sqrtf(kxv * kxv + kyv * kyv);
    const float gr = -4.f * 3.14159f * 3.14159f * (kxv * kxv + kyv * kyv);
    return make_float2(k <= a.kmax ? a.scale * gr : 0.f, 0.f);
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
template <class P, int N, int MODE, int MUL, class S, class LL, class A>
__device__ __forceinline__ void col_body(float2 (&v)[N / P::T], int t, int x, const A& a, const LL& L,
                                         const float2* tw) {
  using Tr = ColModeTraits<MODE>;
  constexpr int T = P::T, E = N / T;
  if constexpr (Tr::FWD) fft_line<P, N, -1, S>(v, t, tw, L);
  // frequencies are read only by modes that need them (scale-only passes may
  // run without k-grids)
  constexpr bool NEEDK = (Tr::MULT && (MUL == CM_BAND || MUL == CM_GRAD || MUL == CM_BANDGRAD));
  constexpr bool NEEDROWS =
      (Tr::MULT && (MUL == CM_PROP || MUL == CM_PROPCONJ || MUL == CM_PROP_ZB || MUL == CM_PROPCONJ_ZB));
  constexpr bool CONJ = MUL == CM_PROPCONJ || MUL == CM_PROPCONJ_ZB;
  // the propagator zeroes rows N/4 .. 3N/4 - 1 (canonical blocks E/4 .. 3E/4 - 1)
  constexpr bool ZB = (MUL == CM_PROP_ZB || MUL == CM_PROPCONJ_ZB) && E % 4 == 0 && P::R1 % 4 == 0;
  float kxv = 0.f;
  if constexpr (NEEDK) kxv = a.kx[x];
  int2 pr = make_int2(0, 0);
  if constexpr (NEEDROWS) pr = a.prop_rows[x];
  if constexpr (NEEDROWS) {
    // propagator inside |k| < kmax (rows per column from k_prop_rows): the
    // factor is loaded for every row and zeroed by a multiplier, so the
    // compiler issues the loads together instead of one branch (and one
    // memory round trip) per element
    const float2 pxv = a.px[x];
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int r = t + c * T;
      float2 p = cmul(pxv, a.py[r]);
      if constexpr (CONJ) p.y = -p.y;
      const float m = (r <= pr.x || r >= pr.y) ? a.scale : 0.f;
      v[c] = cmul(v[c], cscale(p, m));
    }
  } else if constexpr (Tr::MULT) {
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int r = t + c * T;
      float kyv = 0.f;
      if constexpr (NEEDK) kyv = a.ky[r];
      v[c] = cmul(v[c], kmul<MUL>(a, x, r, kxv, kyv, pr));
    }
  } else if (a.scale != 1.0f) {
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = cscale(v[c], a.scale);
  }
  if constexpr (Tr::INV) fft_line<P, N, +1, S, LL, ZB>(v, t, tw, L);
}

// variants 0, 1: the C columns of a tile spread over the workgroup, thread
// (col, t) holds rows t + c T of column col; workgroup barriers
// launched tile bx (of nbx per wave) of wave b; lds holds Cfg::LDS elements,
// stw the twiddles of length N
template <int N, int MODE, int MUL, class Cfg, class A>
__device__ __forceinline__ void col_block(const A& a, int bx, int b, int nbx, float2* lds, const float2* stw) {
  constexpr int T = Cfg::T, E = N / T, C = Cfg::C;
  const int col = Cfg::WAVECOL ? threadIdx.x / T : threadIdx.x % C;
  const int t = Cfg::WAVECOL ? threadIdx.x % T : threadIdx.x / C;
  using S = typename std::conditional<Cfg::WAVECOL, WaveSync, BlockSync>::type;
  // band-limited runs launch only the tiles at both ends of kx
  const int tile = (a.tiles_lo > 0 && bx >= a.tiles_lo) ? bx + a.nx / C - 2 * a.tiles_lo : bx;
  const int x = tile * C + col;
  const long long nx = a.nx;
  float2 v[E];
  {
    // the wave's base is uniform (a scalar register) and the offsets within
    // one wave fit 32 bits: one vector add per access instead of a 64-bit
    // multiply-add
    const float2* src = a.src + (long long)b * a.src_bstride;
    const unsigned o0 = (unsigned)t * (unsigned)nx + (unsigned)x, step = (unsigned)T * (unsigned)nx;
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = src[o0 + (unsigned)c * step];
  }
  using LL = LdsLine<Cfg::SH>;
  const LL L = Cfg::INTERLEAVED ? LL{lds + col, C} : LL{lds + col * Cfg::LSC, 1};
  if constexpr (Cfg::DUAL) {
    const DualLine<LL> D{L, LL{lds + N * C + col, C}};
    col_body<typename Cfg::P, N, MODE, MUL, S>(v, t, x, a, D, stw);
  } else {
    col_body<typename Cfg::P, N, MODE, MUL, S>(v, t, x, a, L, stw);
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
    if (threadIdx.x == 0) a.partial[(long long)b * nbx + bx] = red[0] * a.det_scale;
    if (a.dst == nullptr) return;
  }
  {
    float2* dst = a.dst + (long long)b * a.dst_bstride;
    const unsigned o0 = (unsigned)t * (unsigned)nx + (unsigned)x, step = (unsigned)T * (unsigned)nx;
#pragma unroll
    for (int c = 0; c < E; ++c) dst[o0 + (unsigned)c * step] = v[c];
  }
}

template <int N, int MODE, int MUL, int CV>
__global__ void __launch_bounds__((ColCfg<N, CV>::THREADS), (ColCfg<N, CV>::MIN_WAVES)) k_col(ColArgs a) {
  using Cfg = ColCfg<N, CV>;
  __shared__ float2 lds[Cfg::LDS];
  __shared__ float2 stw[N];  // twiddles staged in LDS
  for (int i = threadIdx.x; i < N; i += Cfg::THREADS) stw[i] = a.tw[i];
  __syncthreads();
  // logical (tile, wave) of this workgroup, XCD-aware
  const long long gid = xcd_remap((long long)blockIdx.y * gridDim.x + blockIdx.x, (long long)gridDim.x * gridDim.y);
  col_block<N, MODE, MUL, Cfg>(a, (int)(gid % gridDim.x), (int)(gid / gridDim.x), (int)gridDim.x, lds, stw);
}

// ------------------------------------------------------------------ columns looping over waves
// Variant 5, for the propagation pass (FFT_y -> P -> IFFT_y): a workgroup
// owns one column tile and loops over per_tile waves of the batch; each
// thread forms its propagator factors once (the other layouts load and form
// them again for every wave).  Tile, plan and LDS layout of the length's
// default variant (7 at 512, 0 otherwise), held to four waves per SIMD (112
// VGPRs at 512).  Requesting the next wave's tile during the transforms as
// well took 174 VGPRs, two waves per SIMD, and ran slower than variant 7.
template <int N> struct ColPL { static constexpr int CV = N == 512 ? 7 : 0; };

template <int N, int MUL>
__global__ void __launch_bounds__((ColCfg<N, ColPL<N>::CV>::THREADS), 4) k_col_pl(ColArgs a, int per_tile) {
  using Cfg = ColCfg<N, ColPL<N>::CV>;
  using P = typename Cfg::P;
  using LL = LdsLine<Cfg::SH>;
  constexpr int T = Cfg::T, E = N / T, C = Cfg::C;
  __shared__ float2 lds[Cfg::LDS];
  __shared__ float2 stw[N];
  for (int i = threadIdx.x; i < N; i += Cfg::THREADS) stw[i] = a.tw[i];
  const long long gid = xcd_remap((long long)blockIdx.y * gridDim.x + blockIdx.x, (long long)gridDim.x * gridDim.y);
  const int bx = (int)(gid % gridDim.x), g = (int)(gid / gridDim.x);
  const int col = threadIdx.x % C, t = threadIdx.x / C;
  const int tile = (a.tiles_lo > 0 && bx >= a.tiles_lo) ? bx + a.nx / C - 2 * a.tiles_lo : bx;
  const int x = tile * C + col;
  const long long nx = a.nx;
  constexpr bool CONJ = MUL == CM_PROPCONJ || MUL == CM_PROPCONJ_ZB;
  // the propagator rows N/4 .. 3N/4 - 1 are zero (canonical blocks E/4 .. 3E/4 - 1)
  constexpr bool ZB = (MUL == CM_PROP_ZB || MUL == CM_PROPCONJ_ZB) && E % 4 == 0 && P::R1 % 4 == 0;
  // P / N^2 inside |k| < kmax (the per-column row runs), as col_body forms it
  float2 f[E];
  {
    const int2 pr = a.prop_rows[x];
    const float2 pxv = a.px[x];
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int r = t + c * T;
      float2 p = cmul(pxv, a.py[r]);
      if constexpr (CONJ) p.y = -p.y;
      const float m = (r <= pr.x || r >= pr.y) ? a.scale : 0.f;
      f[c] = cscale(p, m);
    }
  }
  __syncthreads();
  const LL L = Cfg::INTERLEAVED ? LL{lds + col, C} : LL{lds + col * Cfg::LSC, 1};
  // (64-bit addressing here: the 32-bit offsets of col_block, held across
  // the loop, spill under the four-wave register budget)
  auto load = [&](float2 (&v)[E], int b) {
    const float2* src = a.src + (long long)b * a.src_bstride + x;
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = src[(long long)(t + c * T) * nx];
  };
  for (int q = 0; q < per_tile; ++q) {
    const int b = g * per_tile + q;
    if (b >= a.nbatch) break;  // uniform over the workgroup
    float2 v[E];
    load(v, b);
    fft_line<P, N, -1, BlockSync>(v, t, stw, L);
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = cmul(v[c], f[c]);
    fft_line<P, N, +1, BlockSync, LL, ZB>(v, t, stw, L);
    float2* dst = a.dst + (long long)b * a.dst_bstride + x;
#pragma unroll
    for (int c = 0; c < E; ++c) dst[(long long)(t + c * T) * nx] = v[c];
  }
}

// ------------------------------------------------------------------ grouped (uneven sizes)
// One launch covers the waves of several grid sizes, one segment per size
// (256 .. 2048): the workgroups of all segments are concatenated and each runs
// the body of its block's length.  Rows use 256 threads per workgroup at every
// length, columns 512; per length the plan and tile are those the tuned
// per-size passes use where the shape allows.
template <int N> struct GroupRow { static constexpr int RV = N == 512 ? 2 : 1; };
template <int N> struct GroupCol {
  using P = Plan<N>;
  static constexpr bool INTERLEAVED = true, DUAL = false, WAVECOL = false;
  static constexpr int T = P::T, THREADS = 512, C = THREADS / T, SH = 0, LSC = 0, LDS = N * C;
  static_assert(C * T == THREADS && C <= N / 4, "tile does not fit the band");
};
constexpr int cmax(int a, int b) { return a > b ? a : b; }
template <int N> constexpr int group_row_lds() { return RowCfg<N, GroupRow<N>::RV>::LINES * RowCfg<N, GroupRow<N>::RV>::LS; }
constexpr int kGroupRowLds = cmax(cmax(group_row_lds<256>(), group_row_lds<512>()),
                                  cmax(group_row_lds<1024>(), group_row_lds<2048>()));
constexpr int kGroupColLds = cmax(cmax(GroupCol<256>::LDS, GroupCol<512>::LDS), cmax(GroupCol<1024>::LDS, GroupCol<2048>::LDS));
constexpr int kGroupTw = 2048;

// Segment boundaries and lengths as plain kernel arguments: the segment of a
// workgroup and its length come from scalar compares and selects, and only
// the segment's pass arguments are read from the table in device memory
// (one level of dependent loads before the body, as for a per-size pass).
// (A persistent loop over the blocks of all segments was measured by the
// compiler at 256 VGPRs with the four bodies in one loop: one workgroup per
// block keeps the register count of the largest body.)
struct GroupHdr {
  long long b1, b2, b3, total;  // first block of segments 1..3 (= total when absent)
  int n0, n1, n2, n3;           // lengths
  int x0, x1, x2, x3;           // columns: launched tiles per wave
};
template <class T> __device__ __forceinline__ T pick4(int s, T a, T b, T c, T d) {
  return s == 0 ? a : s == 1 ? b : s == 2 ? c : d;
}

template <int MODE, int MUL, int FL>
__global__ void __launch_bounds__(256) k_row_group(const PassGroup<RowArgs>* g, GroupHdr h) {
  __shared__ float2 lds[kGroupRowLds];
  __shared__ float2 stw[kGroupTw];
  // XCD-aware, and at 2048 wave-interleaved inside the segment (as k_row)
  const long long it = xcd_remap(blockIdx.x, gridDim.x);
  if (it >= h.total) return;
  const int s = (it >= h.b1) + (it >= h.b2) + (it >= h.b3);
  const long long blk = it - pick4<long long>(s, 0, h.b1, h.b2, h.b3);
  const int n = pick4(s, h.n0, h.n1, h.n2, h.n3);
  const RowArgsG& a = reinterpret_cast<const RowArgsG*>(g->a)[s];
  for (int i = threadIdx.x; i < n; i += 256) stw[i] = a.tw[i];
  __syncthreads();
  switch (n) {
    case 256: row_block<256, MODE, MUL, FL, GroupRow<256>::RV>(a, blk, lds, stw); break;
    case 512: row_block<512, MODE, MUL, FL, GroupRow<512>::RV>(a, blk, lds, stw); break;
    case 1024: row_block<1024, MODE, MUL, FL, GroupRow<1024>::RV>(a, blk, lds, stw); break;
    default: {
      constexpr int ybs = 2048 / RowCfg<2048, GroupRow<2048>::RV>::LINES;
      row_block<2048, MODE, MUL, FL, GroupRow<2048>::RV>(a, row_interleave(blk, a.nbatch, ybs), lds, stw);
      break;
    }
  }
}

template <int MODE, int MUL>
__global__ void __launch_bounds__(512) k_col_group(const PassGroup<ColArgs>* g, GroupHdr h) {
  __shared__ float2 lds[kGroupColLds];
  __shared__ float2 stw[kGroupTw];
  // XCD-aware: each XCD takes a contiguous range of blocks, so neighbouring
  // tiles share an L2
  const long long it = xcd_remap(blockIdx.x, gridDim.x);
  if (it >= h.total) return;
  const int s = (it >= h.b1) + (it >= h.b2) + (it >= h.b3);
  const long long blk = it - pick4<long long>(s, 0, h.b1, h.b2, h.b3);
  const int n = pick4(s, h.n0, h.n1, h.n2, h.n3), nbx = pick4(s, h.x0, h.x1, h.x2, h.x3);
  const int b = (int)(blk / nbx), bx = (int)(blk - (long long)b * nbx);
  const ColArgsG& a = reinterpret_cast<const ColArgsG*>(g->a)[s];
  for (int i = threadIdx.x; i < n; i += 512) stw[i] = a.tw[i];
  __syncthreads();
  switch (n) {
    case 256: col_block<256, MODE, MUL, GroupCol<256>>(a, bx, b, nbx, lds, stw); break;
    case 512: col_block<512, MODE, MUL, GroupCol<512>>(a, bx, b, nbx, lds, stw); break;
    case 1024: col_block<1024, MODE, MUL, GroupCol<1024>>(a, bx, b, nbx, lds, stw); break;
    default: col_block<2048, MODE, MUL, GroupCol<2048>>(a, bx, b, nbx, lds, stw); break;
  }
}

template <class A> static GroupHdr group_hdr(const PassGroup<A>& h) {
  GroupHdr r{};
  const long long total = h.blk0[h.nseg];
  r.total = total;
  r.b1 = h.nseg > 1 ? h.blk0[1] : total;
  r.b2 = h.nseg > 2 ? h.blk0[2] : total;
  r.b3 = h.nseg > 3 ? h.blk0[3] : total;
  int n[4] = {0, 0, 0, 0}, x[4] = {1, 1, 1, 1};
  for (int k = 0; k < h.nseg; ++k) {
    n[k] = h.a[k].nx;
    x[k] = h.nbx[k] > 0 ? h.nbx[k] : 1;
  }
  r.n0 = n[0]; r.n1 = n[1]; r.n2 = n[2]; r.n3 = n[3];
  r.x0 = x[0]; r.x1 = x[1]; r.x2 = x[2]; r.x3 = x[3];
  return r;
}

static bool group_len(int n) { return n == 256 || n == 512 || n == 1024 || n == 2048; }

template <int N> static int group_row_lines() { return RowCfg<N, GroupRow<N>::RV>::LINES; }
static int group_row_lines(int n) {
  return n == 256 ? group_row_lines<256>() : n == 512 ? group_row_lines<512>() : n == 1024 ? group_row_lines<1024>()
                                                                                           : group_row_lines<2048>();
}
static int group_col_tile(int n) {
  return n == 256 ? GroupCol<256>::C : n == 512 ? GroupCol<512>::C : n == 1024 ? GroupCol<1024>::C : GroupCol<2048>::C;
}

hipError_t build_row_group(const RowArgs* segs, int nseg, PassGroup<RowArgs>& g) {
  if (nseg < 1 || nseg > kGroupMaxSeg) return hipErrorInvalidValue;
  g = PassGroup<RowArgs>{};
  int k = 0;
  long long total = 0;
  for (int i = 0; i < nseg; ++i) {
    const RowArgs& a = segs[i];
    if (!group_len(a.nx) || a.ny != a.nx || a.aux || a.prune_in != segs[0].prune_in ||
        a.prune_out != segs[0].prune_out)
      return hipErrorInvalidValue;
    if (a.nbatch <= 0) continue;
    g.a[k] = a;
    g.blk0[k] = total;
    const int lines = group_row_lines(a.nx);
    total += ((long long)a.nbatch * a.ny + lines - 1) / lines;
    ++k;
  }
  g.nseg = k;
  g.blk0[k] = total;
  return hipSuccess;
}

hipError_t build_col_group(const ColArgs* segs, int nseg, PassGroup<ColArgs>& g) {
  if (nseg < 1 || nseg > kGroupMaxSeg) return hipErrorInvalidValue;
  g = PassGroup<ColArgs>{};
  int k = 0;
  long long total = 0;
  for (int i = 0; i < nseg; ++i) {
    const ColArgs& a = segs[i];
    if (!group_len(a.nx) || a.ny != a.nx || !a.prop_rows) return hipErrorInvalidValue;
    if (a.nbatch <= 0) continue;
    const int C = group_col_tile(a.nx);
    g.a[k] = a;
    // tiles_lo counts tiles of the default width; convert to this tile's
    if (a.tiles_lo > 0) {
      if ((a.nx / 4) % C != 0) return hipErrorInvalidValue;
      g.a[k].tiles_lo = a.nx / 4 / C;
    }
    g.nbx[k] = g.a[k].tiles_lo > 0 ? 2 * g.a[k].tiles_lo : a.nx / C;
    g.blk0[k] = total;
    total += (long long)g.nbx[k] * a.nbatch;
    ++k;
  }
  g.nseg = k;
  g.blk0[k] = total;
  return hipSuccess;
}

hipError_t launch_row_group(int mode, int mul, const PassGroup<RowArgs>& h, const PassGroup<RowArgs>* d,
                            hipStream_t s) {
  const long long total = h.blk0[h.nseg];
  if (h.nseg == 0 || total <= 0) return hipSuccess;
  if (!d) return hipErrorInvalidValue;
  const int fl = (h.a[0].prune_in ? PRUNE_IN : 0) | (h.a[0].prune_out ? PRUNE_OUT : 0);
  if (total > 0x7fffffffLL) return hipErrorInvalidValue;
  for (int k = 0; k < h.nseg; ++k)
    if (h.a[k].ny <= 0) return hipErrorInvalidValue;  // the kernels walk rows in steps of ny
  const GroupHdr hd = group_hdr(h);
#define MSL_RG(M, U, F) hipLaunchKernelGGL((k_row_group<M, U, F>), dim3((unsigned)total), dim3(256), 0, s, d, hd)
  if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == PRUNE_OUT) MSL_RG(ROW_INV_MUL_FWD, RM_T, PRUNE_OUT);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == (PRUNE_IN | PRUNE_OUT))
    MSL_RG(ROW_INV_MUL_FWD, RM_T, PRUNE_IN | PRUNE_OUT);
  else return hipErrorInvalidValue;
#undef MSL_RG
  return hipGetLastError();
}

hipError_t launch_col_group(int mode, int mul, const PassGroup<ColArgs>& h, const PassGroup<ColArgs>* d,
                            hipStream_t s) {
  const long long total = h.blk0[h.nseg];
  if (h.nseg == 0 || total <= 0) return hipSuccess;
  if (!d) return hipErrorInvalidValue;
  if (total > 0x7fffffffLL) return hipErrorInvalidValue;
  const GroupHdr hd = group_hdr(h);
#define MSL_CG(M, U) hipLaunchKernelGGL((k_col_group<M, U>), dim3((unsigned)total), dim3(512), 0, s, d, hd)
  if (mode == COL_FWD_MUL_INV && mul == CM_PROP) MSL_CG(COL_FWD_MUL_INV, CM_PROP);
  else return hipErrorInvalidValue;
#undef MSL_CG
  return hipGetLastError();
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
  if (mode == ROW_INV_FD_FWD && (!a.fd_pm || !a.fd_p0 || !a.t)) return hipErrorInvalidValue;
  if (mul == RM_SEED && (mode != ROW_FWD || !a.seed)) return hipErrorInvalidValue;
#define MSL_ROW(M, U, F) hipLaunchKernelGGL((k_row<N, M, U, F, RV>), grid, block, 0, s, a)
  if (mode == ROW_FWD && mul == RM_NONE && fl == 0) MSL_ROW(ROW_FWD, RM_NONE, 0);
  else if (mode == ROW_FWD && mul == RM_NONE && fl == PRUNE_OUT) MSL_ROW(ROW_FWD, RM_NONE, PRUNE_OUT);
  else if (mode == ROW_FWD && mul == RM_T && fl == 0) MSL_ROW(ROW_FWD, RM_T, 0);
  else if (mode == ROW_FWD && mul == RM_T && fl == PRUNE_OUT) MSL_ROW(ROW_FWD, RM_T, PRUNE_OUT);
  else if (mode == ROW_FWD && mul == RM_SEED && fl == 0) MSL_ROW(ROW_FWD, RM_SEED, 0);
  else if (mode == ROW_FWD && mul == RM_SEED && fl == PRUNE_OUT) MSL_ROW(ROW_FWD, RM_SEED, PRUNE_OUT);
  else if (mode == ROW_INV && mul == RM_NONE && fl == 0) MSL_ROW(ROW_INV, RM_NONE, 0);
  else if (mode == ROW_INV && mul == RM_NONE && fl == PRUNE_IN) MSL_ROW(ROW_INV, RM_NONE, PRUNE_IN);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == 0) MSL_ROW(ROW_INV_MUL_FWD, RM_T, 0);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == PRUNE_OUT) MSL_ROW(ROW_INV_MUL_FWD, RM_T, PRUNE_OUT);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == (PRUNE_IN | PRUNE_OUT))
    MSL_ROW(ROW_INV_MUL_FWD, RM_T, PRUNE_IN | PRUNE_OUT);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == STORE_AUX) MSL_ROW(ROW_INV_MUL_FWD, RM_T, STORE_AUX);
  else if (mode == ROW_INV_MUL_FWD && mul == RM_T && fl == (PRUNE_IN | PRUNE_OUT | STORE_AUX))
    MSL_ROW(ROW_INV_MUL_FWD, RM_T, PRUNE_IN | PRUNE_OUT | STORE_AUX);
  else if (mode == ROW_INV_FD_FWD && mul == RM_T && fl == (PRUNE_IN | PRUNE_OUT))
    MSL_ROW(ROW_INV_FD_FWD, RM_T, PRUNE_IN | PRUNE_OUT);
  else if (mode == ROW_INV_FD_FWD && mul == RM_T && fl == 0) MSL_ROW(ROW_INV_FD_FWD, RM_T, 0);
  else return hipErrorInvalidValue;
#undef MSL_ROW
  return hipGetLastError();
}

// variant 4 where its layout applies (see k_row_pl); elsewhere the best of
// the other layouts for the length (2 at 512, 3 at 1024, 1 otherwise)
template <int N>
static hipError_t row_pl(int mode, int mul, const RowArgs& a, hipStream_t s) {
  const bool shared_t = a.t && !a.t_index && a.t_group >= a.nbatch;
  // per-wave transmission slots (the serving engine's batches of several
  // jobs) take the loop's slot-following form, from 256 up
  const bool slots = a.t && !shared_t && N >= 256;
  const int fl = (a.prune_in ? PRUNE_IN : 0) | (a.prune_out ? PRUNE_OUT : 0) | (a.aux ? STORE_AUX : 0) |
                 (slots ? T_SLOTS : 0);
  // a batch narrower than two waves per line would leave most lines idle
  if (mode != ROW_INV_MUL_FWD || mul != RM_T || !(shared_t || slots) || a.ny > 65535 ||
      a.nbatch < 2 * RowPL<N>::LINES) {
    if constexpr (N == 512) return row_nv<N, 2>(mode, mul, a, s);
    else if constexpr (N == 1024) return row_nv<N, 3>(mode, mul, a, s);
    else return row_nv<N, 1>(mode, mul, a, s);
  }
  if (a.nbatch <= 0) return hipSuccess;
  constexpr int LINES = RowPL<N>::LINES, T = RowPL<N>::T;
  // four waves per line: 1.129e6 probe-slice/s on the 512^2 STEM workload
  // against 1.122e6 with eight and 1.111e6 with two
  const int per_line = std::max(1, std::min(4, (a.nbatch + LINES - 1) / LINES));
  const int groups = (a.nbatch + LINES * per_line - 1) / (LINES * per_line);
  if (groups > 65535) return row_nv<N, 1>(mode, mul, a, s);
  dim3 grid(a.ny, groups), block(LINES * T);
  if (fl == (PRUNE_IN | PRUNE_OUT)) hipLaunchKernelGGL((k_row_pl<N, PRUNE_IN | PRUNE_OUT>), grid, block, 0, s, a, per_line);
  else if (fl == (PRUNE_IN | PRUNE_OUT | STORE_AUX))
    hipLaunchKernelGGL((k_row_pl<N, PRUNE_IN | PRUNE_OUT | STORE_AUX>), grid, block, 0, s, a, per_line);
  else if (fl == STORE_AUX) hipLaunchKernelGGL((k_row_pl<N, STORE_AUX>), grid, block, 0, s, a, per_line);
  else if (fl == PRUNE_OUT) hipLaunchKernelGGL((k_row_pl<N, PRUNE_OUT>), grid, block, 0, s, a, per_line);
  else if (fl == PRUNE_IN) hipLaunchKernelGGL((k_row_pl<N, PRUNE_IN>), grid, block, 0, s, a, per_line);
  else if (fl == 0) hipLaunchKernelGGL((k_row_pl<N, 0>), grid, block, 0, s, a, per_line);
  else if constexpr (N >= 256) {
    if (fl == (PRUNE_IN | PRUNE_OUT | T_SLOTS))
      hipLaunchKernelGGL((k_row_pl<N, PRUNE_IN | PRUNE_OUT | T_SLOTS>), grid, block, 0, s, a, per_line);
    else if (fl == (PRUNE_OUT | T_SLOTS))
      hipLaunchKernelGGL((k_row_pl<N, PRUNE_OUT | T_SLOTS>), grid, block, 0, s, a, per_line);
    else if (fl == T_SLOTS) hipLaunchKernelGGL((k_row_pl<N, T_SLOTS>), grid, block, 0, s, a, per_line);
    else return row_nv<N, 1>(mode, mul, a, s);
  } else {
    return row_nv<N, 1>(mode, mul, a, s);
  }
  return hipGetLastError();
}

template <int N>
static hipError_t row_n(int mode, int mul, const RowArgs& a, hipStream_t s, int var) {
  if (var == 4) return row_pl<N>(mode, mul, a, s);
  if (var == 3) return row_nv<N, 3>(mode, mul, a, s);
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
    else if (mode == COL_FWD_MUL_INV && mul == CM_PROP_ZB) MSL_COL(COL_FWD_MUL_INV, CM_PROP_ZB);
    else if (mode == COL_FWD_MUL_INV && mul == CM_PROPCONJ_ZB) MSL_COL(COL_FWD_MUL_INV, CM_PROPCONJ_ZB);
    else if (mode == COL_FWD_MUL_INV && mul == CM_BAND) MSL_COL(COL_FWD_MUL_INV, CM_BAND);
    else if (mode == COL_FWD_MUL_INV && mul == CM_GRAD) MSL_COL(COL_FWD_MUL_INV, CM_GRAD);
    else if (mode == COL_FWD_MUL_INV && mul == CM_BANDGRAD) MSL_COL(COL_FWD_MUL_INV, CM_BANDGRAD);
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

// variant 5 for the propagation pass of a batch (k_col_pl); the length's
// default layout otherwise
template <int N>
static hipError_t col_pl(int mode, int mul, const ColArgs& a, hipStream_t s) {
  constexpr int CV = ColPL<N>::CV, C = ColCfg<N, CV>::C;
  const bool prop = mode == COL_FWD_MUL_INV &&
                    (mul == CM_PROP || mul == CM_PROPCONJ || mul == CM_PROP_ZB || mul == CM_PROPCONJ_ZB);
  if (!prop || a.nbatch < 2) return col_nv<N, CV>(mode, mul, a, s);
  if (a.nx % C != 0 || a.tiles_lo < 0 || 2 * a.tiles_lo > a.nx / ColTile0<N>::C) return hipErrorInvalidValue;
  // tiles_lo counts tiles of the default width; convert to this variant's
  ColArgs b = a;
  if (a.tiles_lo > 0) {
    if ((a.nx / 4) % C != 0) return hipErrorInvalidValue;
    b.tiles_lo = a.nx / 4 / C;
  }
  const int per_tile = std::min(4, a.nbatch);
  const int groups = (a.nbatch + per_tile - 1) / per_tile;
  if (groups > 65535) return col_nv<N, CV>(mode, mul, a, s);
  dim3 grid(b.tiles_lo > 0 ? 2 * b.tiles_lo : a.nx / C, groups), block(ColCfg<N, CV>::THREADS);
  if (mul == CM_PROP) hipLaunchKernelGGL((k_col_pl<N, CM_PROP>), grid, block, 0, s, b, per_tile);
  else if (mul == CM_PROPCONJ) hipLaunchKernelGGL((k_col_pl<N, CM_PROPCONJ>), grid, block, 0, s, b, per_tile);
  else if (mul == CM_PROP_ZB) hipLaunchKernelGGL((k_col_pl<N, CM_PROP_ZB>), grid, block, 0, s, b, per_tile);
  else hipLaunchKernelGGL((k_col_pl<N, CM_PROPCONJ_ZB>), grid, block, 0, s, b, per_tile);
  return hipGetLastError();
}

template <int N>
static hipError_t col_n(int mode, int mul, const ColArgs& a, hipStream_t s, int var) {
  // the detector pass reduces over the workgroup in an interleaved layout
  if (mode == COL_FWD_DETECT) {
    if (var == 4) return col_nv<N, 4>(mode, mul, a, s);
    return var == 2 ? col_nv<N, 2>(mode, mul, a, s) : col_nv<N, 0>(mode, mul, a, s);
  }
  if (var == 6) {
    if constexpr (ColCfg<N, 6>::WAVECOL) return col_nv<N, 6>(mode, mul, a, s);
    else return col_nv<N, 0>(mode, mul, a, s);
  }
  if (var == 7) return col_nv<N, 7>(mode, mul, a, s);
  if (var == 5) return col_pl<N>(mode, mul, a, s);
  if (var == 4) return col_nv<N, 4>(mode, mul, a, s);
  if (var == 3) return col_nv<N, 3>(mode, mul, a, s);
  if (var == 2) return col_nv<N, 2>(mode, mul, a, s);
  if (var == 1) return col_nv<N, 1>(mode, mul, a, s);
  return col_nv<N, 0>(mode, mul, a, s);
}

hipError_t launch_row_v(int mode, int mul, const RowArgs& a, hipStream_t s, int var) {
  // the kernels walk rows in steps of ny: a grid needs at least one row
  if (a.ny <= 0 || a.nbatch < 0) return hipErrorInvalidValue;
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

static bool needs_prop_rows(int mode, int mul) {
  return (mul == CM_PROP || mul == CM_PROPCONJ || mul == CM_PROP_ZB || mul == CM_PROPCONJ_ZB) && mode != COL_FWD &&
         mode != COL_INV;
}

bool prop_rows_quarter(const float* hky, int ny, float kmax) {
  if (ny < 4) return false;
  const float kyv = hky[ny / 4], kxv = 0.0f;
  return !(std::sqrt(kxv * kxv + kyv * kyv) < kmax);
}

// Column launches put the waves on grid.y, at most 65535 per launch: longer
// batches run as consecutive launches over slices of the batch (the detector
// pass, whose partial sums are indexed by wave, is refused instead).
constexpr int kMaxGridY = 65535;

hipError_t launch_col_v(int mode, int mul, const ColArgs& a, hipStream_t s, int var) {
  if (needs_prop_rows(mode, mul) && !a.prop_rows) return hipErrorInvalidValue;
  if (a.nbatch > kMaxGridY) {
    if (mode == COL_FWD_DETECT) return hipErrorInvalidValue;
    for (long long b0 = 0; b0 < a.nbatch; b0 += kMaxGridY) {
      ColArgs c = a;
      c.nbatch = (int)std::min<long long>(kMaxGridY, a.nbatch - b0);
      c.src = a.src + b0 * a.src_bstride;
      c.dst = a.dst + b0 * a.dst_bstride;
      const hipError_t e = launch_col_v(mode, mul, c, s, var);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  // the inverse transform after the propagator skips rows known to be zero
  if (mode == COL_FWD_MUL_INV && a.zb_rows) {
    if (mul == CM_PROP) mul = CM_PROP_ZB;
    else if (mul == CM_PROPCONJ) mul = CM_PROPCONJ_ZB;
  }
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
