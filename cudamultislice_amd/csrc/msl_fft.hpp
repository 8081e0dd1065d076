// msl_fft.hpp -- device-side Stockham FFT for one line (row or column) held
// in registers, with LDS exchanges between radix stages.  gfx950 / wave64.
//
// A line of N complex values is owned by T threads; thread t holds the E =
// N/T values at positions t + c*T, c = 0..E-1 ("canonical order").  Every
// Stockham stage reads canonical positions and, except the last, scatters
// its outputs through LDS; the last stage leaves its outputs in canonical
// order again.  Hence an inverse transform, a pointwise multiply and a
// forward transform chain without any exchange in between, and the first
// load / last store of a line are coalesced (stride-T across the threads).
#pragma once
#include <hip/hip_runtime.h>

namespace msl {

__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a * W4 with W4 = exp(DIR * 2 pi i / 4): forward (DIR=-1) is -i, inverse +i
// streaming (non-temporal) access to complex values that are not reused soon
typedef float msl_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 nt_load(const float2* p) {
  const msl_f2v v = __builtin_nontemporal_load(reinterpret_cast<const msl_f2v*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void nt_store(float2 v, float2* p) {
  msl_f2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<msl_f2v*>(p));
}

template <int DIR> __host__ __device__ __forceinline__ float2 rot4(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int DIR> __host__ __device__ __forceinline__ void dft2(float2& a0, float2& a1) {
  float2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

// Z: a1 = a2 = 0 on entry (the middle of a band-limited line).  Adding an
// exact zero changes nothing but the sign of a zero result, so the shortened
// butterfly gives the values of the full one.
template <int DIR, bool Z = false>
__host__ __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  if constexpr (Z) {
    const float2 x = a0, r = rot4<DIR>(a3);
    a0 = cadd(x, a3);
    a2 = csub(x, a3);
    a1 = csub(x, r);
    a3 = cadd(x, r);
    return;
  }
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = rot4<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// Z: a[2..5] = 0 on entry
template <int DIR, bool Z = false> __host__ __device__ __forceinline__ void dft8(float2 (&a)[8]) {
  constexpr float h = 0.70710678118654752f;
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<DIR, Z>(e0, e1, e2, e3);
  dft4<DIR, Z>(o0, o1, o2, o3);
  // o1 *= W8^1, o2 *= W8^2, o3 *= W8^3 with W8 = exp(DIR 2 pi i / 8)
  if (DIR < 0) {
    o1 = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
    o3 = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
  } else {
    o1 = make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y));
    o3 = make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y));
  }
  o2 = rot4<DIR>(o2);
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}

// Z: a[4..11] = 0 on entry
template <int DIR, bool Z = false> __host__ __device__ __forceinline__ void dft16(float2 (&a)[16]) {
  // 16 = 4 x 4: DFT4 over n1 (stride 4), twiddle W16^(n2 k1), DFT4 over n2
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<DIR, Z>(a[n2], a[n2 + 4], a[n2 + 8], a[n2 + 12]);
  // now a[n2 + 4 k1] = B[n2][k1]; apply W16^(n2*k1), W16 = exp(DIR 2 pi i/16)
  const float sg = DIR < 0 ? -1.f : 1.f;
  const float2 w1 = make_float2(c1, sg * s1), w2 = make_float2(h, sg * h), w3 = make_float2(s1, sg * c1);
  const float2 w6 = make_float2(-h, sg * h), w9 = make_float2(-c1, -sg * s1);
  a[1 + 4] = cmul(a[1 + 4], w1);
  a[1 + 8] = cmul(a[1 + 8], w2);
  a[1 + 12] = cmul(a[1 + 12], w3);
  a[2 + 4] = cmul(a[2 + 4], w2);
  a[2 + 8] = rot4<DIR>(a[2 + 8]);
  a[2 + 12] = cmul(a[2 + 12], w6);
  a[3 + 4] = cmul(a[3 + 4], w3);
  a[3 + 8] = cmul(a[3 + 8], w6);
  a[3 + 12] = cmul(a[3 + 12], w9);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<DIR>(a[4 * k1 + 0], a[4 * k1 + 1], a[4 * k1 + 2], a[4 * k1 + 3]);
  // a[4 k1 + k2] holds X[k1 + 4 k2]; transpose into natural order
  float2 o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = a[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = o[i];
}

// Z: a[R/4 .. 3R/4 - 1] = 0 on entry (R >= 4)
template <int R, int DIR, bool Z = false> __host__ __device__ __forceinline__ void dftR(float2 (&a)[R]) {
  static_assert(!Z || R >= 4, "zero-middle butterflies need R >= 4");
  if constexpr (R == 2) dft2<DIR>(a[0], a[1]);
  else if constexpr (R == 4) dft4<DIR, Z>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 8) dft8<DIR, Z>(a);
  else if constexpr (R == 16) dft16<DIR, Z>(a);
}

// Per-length plan: T threads per line, radices of up to three stages.
template <int N> struct Plan;
template <> struct Plan<16> { static constexpr int T = 4, R1 = 4, R2 = 4, R3 = 1; };
template <> struct Plan<32> { static constexpr int T = 4, R1 = 8, R2 = 4, R3 = 1; };
template <> struct Plan<64> { static constexpr int T = 8, R1 = 8, R2 = 8, R3 = 1; };
template <> struct Plan<128> { static constexpr int T = 8, R1 = 16, R2 = 8, R3 = 1; };
template <> struct Plan<256> { static constexpr int T = 16, R1 = 16, R2 = 16, R3 = 1; };
template <> struct Plan<512> { static constexpr int T = 64, R1 = 8, R2 = 8, R3 = 8; };
template <> struct Plan<1024> { static constexpr int T = 64, R1 = 16, R2 = 8, R3 = 8; };
template <> struct Plan<2048> { static constexpr int T = 128, R1 = 16, R2 = 16, R3 = 8; };

// Alternative plans: a different thread count per line and radix order (more
// butterflies per thread and stage, or fewer threads), selectable per length
// as pass variants.
template <int N> struct AltPlan;
template <> struct AltPlan<16> { static constexpr int T = 2, R1 = 8, R2 = 2, R3 = 1; };
template <> struct AltPlan<32> { static constexpr int T = 8, R1 = 4, R2 = 4, R3 = 2; };
template <> struct AltPlan<64> { static constexpr int T = 4, R1 = 16, R2 = 4, R3 = 1; };
template <> struct AltPlan<128> { static constexpr int T = 16, R1 = 8, R2 = 8, R3 = 2; };
template <> struct AltPlan<256> { static constexpr int T = 32, R1 = 8, R2 = 8, R3 = 4; };
template <> struct AltPlan<512> { static constexpr int T = 32, R1 = 16, R2 = 8, R3 = 4; };
template <> struct AltPlan<1024> { static constexpr int T = 32, R1 = 16, R2 = 8, R3 = 8; };
template <> struct AltPlan<2048> { static constexpr int T = 64, R1 = 16, R2 = 16, R3 = 8; };

// Exchange synchronisation: a whole workgroup, or one wavefront when every
// line lives inside a single wave (LDS traffic of one wave is processed in
// order; the fences keep the compiler from moving accesses across).
struct BlockSync {
  __device__ static __forceinline__ void sync() { __syncthreads(); }
};
struct WaveSync {
  __device__ static __forceinline__ void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// LDS addressing of one line: element p lives at base[pad(p) * stride] with
// pad(p) = p + (p >> SH) (SH = 0: unpadded), which spreads the power-of-two
// strides of the Stockham scatter over the LDS banks.
template <int SH> struct LdsLine {
  static constexpr bool kDual = false;
  float2* base;
  int stride;
  static __host__ __device__ constexpr int pad(int p) { return SH > 0 ? p + (p >> SH) : p; }
  static __host__ __device__ constexpr int span(int n) { return pad(n - 1) + 1; }
  __device__ __forceinline__ float2& at(int p) const { return base[pad(p) * stride]; }
  // element p + off where pad(p + off) = pad(p) + pad(off): off a multiple of
  // Q, or p a multiple of Q.  With off a compile-time constant the address is
  // one register plus an immediate offset.
  static constexpr int Q = SH > 0 ? (1 << SH) : 1;
  __device__ __forceinline__ float2& at(int p, int off) const { return base[(pad(p) + pad(off)) * stride]; }
  template <int B> __device__ __forceinline__ const LdsLine& get() const { return *this; }
};

// Two copies of a line's LDS storage used by alternate exchanges: a stage
// then needs one barrier (before its reads) instead of two, because a write
// to one copy can only meet reads of the exchange two back, which every
// thread finished before the previous exchange's barrier.  fft_line picks the
// copy per stage so that consecutive exchanges alternate, also across the
// two transforms of a pass (which run in opposite directions).
template <class LL> struct DualLine {
  static constexpr bool kDual = true;
  static constexpr int Q = LL::Q;
  LL a, b;
  template <int B> __device__ __forceinline__ const LL& get() const { return B ? b : a; }
};

// One Stockham stage of radix R after L points have been combined.  Input:
// canonical v[c] = x[t + c T].  Butterfly b (j = t + b T) takes
// x[j + m N/R] = v[b + m NB]; outputs go to y[(j/L) L R + (j%L) + m L].
// If LAST, outputs stay in registers in canonical order; otherwise they are
// written to LDS, and read back in canonical order after a barrier.
// ZIN (first stage only): v[c] = 0 for E/4 <= c < 3E/4, so every butterfly
// gets zeros at m = R/4 .. 3R/4 - 1 (c = b + m NB with b < NB).
template <int N, int T, int R, int L, int DIR, bool LAST, class S, int BUF, class LL, bool ZIN = false>
__device__ __forceinline__ void stockham_stage(float2 (&v)[N / T], int t, const float2* __restrict__ tw,
                                               const LL& lds) {
  constexpr int E = N / T, NB = E / R;
  static_assert(E % R == 0, "radix must divide the per-thread count");
  static_assert(!ZIN || (L == 1 && R % 4 == 0), "zero-middle input only in a first stage of radix 4k");
  float2 out[E];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const int k = j & (L - 1);
    float2 a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = v[b + m * NB];
    if constexpr (L > 1) {
      constexpr int step = N / (L * R);
#pragma unroll
      for (int m = 1; m < R; ++m) {
        float2 w = tw[m * k * step];
        if (DIR > 0) w.y = -w.y;
        a[m] = cmul(a[m], w);
      }
    }
    dftR<R, DIR, ZIN>(a);
    if constexpr (LAST) {
#pragma unroll
      for (int m = 0; m < R; ++m) out[b + m * NB] = a[m];
    } else {
      const int base = (j / L) * L * R + k;
      constexpr int Q = LL::Q;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        if constexpr (L % Q == 0 || (L == 1 && R % Q == 0)) lds.template get<BUF>().at(base, m * L) = a[m];
        else lds.template get<BUF>().at(base + m * L) = a[m];
      }
    }
  }
  if constexpr (LAST) {
#pragma unroll
    for (int c = 0; c < E; ++c) v[c] = out[c];
  } else {
    S::sync();
#pragma unroll
    for (int c = 0; c < E; ++c) {
      if constexpr (T % LL::Q == 0) v[c] = lds.template get<BUF>().at(t, c * T);
      else v[c] = lds.template get<BUF>().at(t + c * T);
    }
    if constexpr (!LL::kDual) S::sync();
  }
}

// Full transform of one line with plan P (T threads, radices R1 R2 R3).
// tw points at the N-entry table exp(-2 pi i k / N).  Every thread that
// shares the synchronisation scope must call this together.
// ZIN: the line is band-limited, v[c] = 0 for E/4 <= c < 3E/4 (E = N / T).
template <class P, int N, int DIR, class S = BlockSync, class LL = LdsLine<0>, bool ZIN = false>
__device__ __forceinline__ void fft_line(float2 (&v)[N / P::T], int t, const float2* __restrict__ tw, const LL& lds) {
  constexpr int T = P::T;
  if constexpr (P::R3 > 1) {  // two exchanges: copies 0, 1 in every transform
    stockham_stage<N, T, P::R1, 1, DIR, false, S, 0, LL, ZIN>(v, t, tw, lds);
    stockham_stage<N, T, P::R2, P::R1, DIR, false, S, 1>(v, t, tw, lds);
    stockham_stage<N, T, P::R3, P::R1 * P::R2, DIR, true, S, 0>(v, t, tw, lds);
  } else {  // one exchange: the copy follows the direction
    stockham_stage<N, T, P::R1, 1, DIR, false, S, (DIR < 0 ? 1 : 0), LL, ZIN>(v, t, tw, lds);
    stockham_stage<N, T, P::R2, P::R1, DIR, true, S, 0>(v, t, tw, lds);
  }
}

// back-compatible form: the default plan of length N
template <int N, int DIR> __device__ __forceinline__ void fft_line(float2 (&v)[N / Plan<N>::T], int t,
                                                                   const float2* __restrict__ tw,
                                                                   const LdsLine<0>& lds) {
  fft_line<Plan<N>, N, DIR, BlockSync, LdsLine<0>>(v, t, tw, lds);
}

// pointer qualifier of the global address space (see MSL_ROW_ARGS); the
// host pass of a .hip file sees plain pointers (same layout)
#if defined(__HIP_DEVICE_COMPILE__)
#define MSL_GLOBAL __attribute__((address_space(1)))
#else
#define MSL_GLOBAL
#endif

}  // namespace msl
