// k_diff.hip -- adjoint row pass of the differentiable multislice.
//
// Forward slice i (row pass, then column pass):
//   a_i = F_x( t_i . w_i ),  w_i = IF_x(z_{i-1}) (= psi_{i-1}),  t_i = exp(i sigma V_i)
//   z_i = IF_y( P F_y(a_i) ) / (nx ny)
// With G the gradient dL/dRe + i dL/dIm carried backwards, the adjoint of the
// row pass is
//   G_u = IF_x(G_a),  G_w = conj(t_i) G_u,  G_{z_{i-1}} = F_x(G_w)
//   dL/dV_i = -sigma Im(conj(G_w) w_i)        (summed over probes)
// computed here in one pass: per row, every probe's G_a and w_i (or z_{i-1})
// are read once, G is written back in place, and the gradient row is reduced
// over the probes of the workgroup in a fixed order (no atomics).
#include "msl_diff.hpp"
#include "msl_fft.hpp"

#include <type_traits>

namespace msl {

// One line per T threads; lines never straddle a wavefront when T <= 64, so
// the FFT exchanges synchronise the wave only (padded LDS, as the row pass).
// At 1024 the stored-wave pass holds 235 VGPRs and the one that transforms z
// 247 (two waves per SIMD): two four-line workgroups per CU, and a 1024-row
// grid fills the 512 slots twice.  Both variants use the same line count
// (adj_lines takes the variant so the host need not assume it).
template <int N, bool REAL_Z> struct AdjCfg {
  static constexpr int T = Plan<N>::T;
  static constexpr bool WAVE = T <= 64;
  static constexpr int L = (256 / T) > 0 ? 256 / T : 1;
  static constexpr int SH = WAVE ? 4 : 0;
  static constexpr int LS = LdsLine<SH>::span(N) + 1;
};

// active canonical blocks of a band-limited line (kx < N/4 or >= 3N/4)
template <int E> __device__ __forceinline__ constexpr bool band_block(int c) {
  return c < (E / 4 > 0 ? E / 4 : 1) || c >= E - (E / 4 > 0 ? E / 4 : 1);
}

// REAL_Z: Z holds w_i = psi_{i-1} in real space (stored by the forward row
// pass, or psi0), so only G is transformed; otherwise Z holds z_{i-1} in the
// (kx, y) domain and w_i = IF_x(z_{i-1}) is recomputed here.
// The row's transmission exp(i sigma V) comes from the forward pass (one
// array per slice): every line of the workgroup processes row y, so each
// thread keeps its E values in registers for all the probes it processes.
template <int N, bool REAL_Z, bool PRUNE>
__global__ void __launch_bounds__((AdjCfg<N, REAL_Z>::L * Plan<N>::T)) k_row_adj(AdjArgs a) {
  using Cfg = AdjCfg<N, REAL_Z>;
  using S = typename std::conditional<Cfg::WAVE, WaveSync, BlockSync>::type;
  using LL = LdsLine<Cfg::SH>;
  constexpr int T = Plan<N>::T, E = N / T, LINES = Cfg::L, LS = Cfg::LS, NT = LINES * T;
  static_assert(LS * 2 >= N, "the reduction reuses the line buffers");
  __shared__ float2 lds[LINES * LS];
  __shared__ float2 stw[N];  // twiddles staged in LDS
  const int tid = threadIdx.x, t = tid % T, ln = tid / T;
  const int y = blockIdx.x, g = blockIdx.y;
  for (int i = tid; i < N; i += NT) stw[i] = a.tw[i];
  float2 tr[E];
  {
    const float2* trow = a.t + (long long)y * N;
#pragma unroll
    for (int c = 0; c < E; ++c) tr[c] = trow[t + c * T];
  }
  __syncthreads();
  const LL L{lds + ln * LS, 1};
  float acc[E];
#pragma unroll
  for (int c = 0; c < E; ++c) acc[c] = 0.f;
  // Software pipeline (PIPE): with two waves per SIMD little latency is
  // hidden by other waves, so each probe's stored wave and the next probe's
  // G row are requested before this probe's transforms (236 VGPRs at 1024).
  // The variants that also transform z, or read unpruned rows, would need
  // more than 256 VGPRs that way: they load G at the top of the iteration
  // and w after the first transform.  Lines past the batch read wave 0 and
  // are zeroed after the loads: unconditional loads issue together instead
  // of one branch each.
  constexpr bool PIPE = REAL_Z && PRUNE;
  const float2 zero = make_float2(0.f, 0.f);
  auto row_of = [&](int q) {
    const int p = (g * a.per_line + q) * LINES + ln;
    return (long long)(p < a.nbatch ? p : 0) * a.bstride + (long long)y * N;
  };
  auto load_g = [&](float2 (&v)[E], long long r) {
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = (!PRUNE || band_block<E>(c)) ? a.G[r + t + c * T] : zero;
  };
  // the stored real-space waves are read once: past the caches, so the
  // gradient rows stay resident for the next column pass
  auto load_w = [&](float2 (&v)[E], long long r) {
#pragma unroll
    for (int c = 0; c < E; ++c) {
      if constexpr (REAL_Z) v[c] = nt_load(&a.Z[r + t + c * T]);
      else v[c] = (!PRUNE || band_block<E>(c)) ? a.Z[r + t + c * T] : zero;
    }
  };
  float2 gn[E];
  if constexpr (PIPE) load_g(gn, row_of(0));
  // band-limited inputs (PRUNE): the first butterflies skip the zero blocks
  constexpr bool ZIN = PRUNE && E % 4 == 0 && Plan<N>::R1 % 4 == 0;
  for (int q = 0; q < a.per_line; ++q) {
    const int p = (g * a.per_line + q) * LINES + ln;
    const bool ok = p < a.nbatch;
    const long long ro = row_of(q);
    float2 gu[E], w[E];
    if constexpr (PIPE) {
#pragma unroll
      for (int c = 0; c < E; ++c) gu[c] = gn[c];
      load_w(w, ro);
      if (q + 1 < a.per_line) load_g(gn, row_of(q + 1));
    } else {
      load_g(gu, ro);
    }
#pragma unroll
    for (int c = 0; c < E; ++c) gu[c] = ok ? gu[c] : zero;
    fft_line<Plan<N>, N, +1, S, LL, ZIN>(gu, t, stw, L);  // G_u = IF_x(G_a)
    if constexpr (!PIPE) load_w(w, ro);
#pragma unroll
    for (int c = 0; c < E; ++c) w[c] = ok ? w[c] : zero;
    if constexpr (!REAL_Z) fft_line<Plan<N>, N, +1, S, LL, ZIN>(w, t, stw, L);  // w = IF_x(z_{i-1})
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const float2 gw = cmulc(gu[c], tr[c]);  // conj(t) G_u
      acc[c] += -a.sigma * (gw.x * w[c].y - gw.y * w[c].x);
      gu[c] = gw;
    }
    fft_line<Plan<N>, N, -1, S>(gu, t, stw, L);  // G_{z_{i-1}} = F_x(G_w)
    if (ok) {
#pragma unroll
      for (int c = 0; c < E; ++c)
        if (!PRUNE || band_block<E>(c)) a.Gout[ro + t + c * T] = gu[c];
    }
  }
  // reduce the LINES partial gradient rows in a fixed order
  __syncthreads();
  float* red = reinterpret_cast<float*>(lds);
#pragma unroll
  for (int c = 0; c < E; ++c) red[ln * N + t + c * T] = acc[c];
  __syncthreads();
  if (ln == 0) {
    const bool direct = a.ngroups == 1 && a.grad != nullptr;
    float* out = direct ? a.grad + (long long)y * N : a.partial + (long long)g * a.ny * N + (long long)y * N;
#pragma unroll
    for (int c = 0; c < E; ++c) {
      float s = 0.f;
      for (int l = 0; l < LINES; ++l) s += red[l * N + t + c * T];
      out[t + c * T] = (direct && a.accumulate) ? out[t + c * T] + s : s;
    }
  }
}


template <int N> static hipError_t adj_n(const AdjArgs& a, bool real_z, hipStream_t s) {
  constexpr int T = Plan<N>::T;
  dim3 grid(a.ny, a.ngroups);
  if (real_z) {
    dim3 block(AdjCfg<N, true>::L * T);
    if (a.prune) hipLaunchKernelGGL((k_row_adj<N, true, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_row_adj<N, true, false>), grid, block, 0, s, a);
  } else {
    dim3 block(AdjCfg<N, false>::L * T);
    if (a.prune) hipLaunchKernelGGL((k_row_adj<N, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_row_adj<N, false, false>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

int adj_lines(int nx, bool real_z) {
  switch (nx) {
#define MSL_ADJ_LINES(n) \
  case n: return real_z ? AdjCfg<n, true>::L : AdjCfg<n, false>::L;
    MSL_ADJ_LINES(16) MSL_ADJ_LINES(32) MSL_ADJ_LINES(64) MSL_ADJ_LINES(128)
    MSL_ADJ_LINES(256) MSL_ADJ_LINES(512) MSL_ADJ_LINES(1024) MSL_ADJ_LINES(2048)
#undef MSL_ADJ_LINES
  }
  return 0;
}

hipError_t launch_row_adjoint(const AdjArgs& a, bool real_z, hipStream_t s) {
  switch (a.nx) {
    case 16: return adj_n<16>(a, real_z, s);
    case 32: return adj_n<32>(a, real_z, s);
    case 64: return adj_n<64>(a, real_z, s);
    case 128: return adj_n<128>(a, real_z, s);
    case 256: return adj_n<256>(a, real_z, s);
    case 512: return adj_n<512>(a, real_z, s);
    case 1024: return adj_n<1024>(a, real_z, s);
    case 2048: return adj_n<2048>(a, real_z, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace msl
