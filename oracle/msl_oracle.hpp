// msl_oracle.hpp -- CPU restatement of the CUDAMultislice algorithm.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
// engine in cudamultislice_amd/.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, and only as the checker or as the
// timed CPU baseline; the product never links, loads or calls it.
//
// Every routine cites the reference lines it restates ("U.cu" is
// "GUI Simulation/UnmanagedGUISimulation.cu").  Continuous quantities are
// evaluated in Real (double for the oracle, float for the CPU baseline).
// Discrete decisions whose outcome the reference takes in FP32 (bin ids,
// slice counts, k-grid values, aperture / band-limit masks) are made in FP32
// exactly as the reference makes them, so both sides agree on every mask.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- threads
inline int& num_threads() {
  static int n = (int)std::max(1u, std::thread::hardware_concurrency());
  return n;
}

inline void parallel_for(int n, const std::function<void(int, int)>& body) {
  int nt = std::min(num_threads(), std::max(n, 1));
  if (nt <= 1 || n < 2) { body(0, n); return; }
  std::vector<std::thread> th;
  int chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back(body, b, e);
  }
  for (auto& t : th) t.join();
}

// ---------------------------------------------------------------- Bessel
// Abramowitz & Stegun 9.8.1 / 9.8.2 (I0) and 9.8.5 / 9.8.6 (K0), the
// polynomial approximations the reference evaluates in U.cu:41-111.
template <class Real> Real bessi0(Real x) {
  static const double a[7] = {1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.0360768, 0.0045813};
  static const double b[9] = {0.39894228, 0.01328592, 0.00225319, -0.00157565, 0.00916281,
                              -0.02057706, 0.02635537, -0.01647633, 0.00392377};
  Real ax = std::fabs(x), sum, t;
  if (ax <= Real(3.75)) {
    t = x / Real(3.75);
    t = t * t;
    sum = Real(a[6]);
    for (int i = 5; i >= 0; --i) sum = sum * t + Real(a[i]);
  } else {
    t = Real(3.75) / ax;
    sum = Real(b[8]);
    for (int i = 7; i >= 0; --i) sum = sum * t + Real(b[i]);
    sum = std::exp(ax) * sum / std::sqrt(ax);
  }
  return sum;
}

template <class Real> Real bessk0(Real x) {
  static const double a[7] = {-0.57721566, 0.42278420, 0.23069756, 0.03488590, 0.00262698, 0.00010750, 0.00000740};
  static const double b[7] = {1.25331414, -0.07832358, 0.02189568, -0.01062446, 0.00587872, -0.00251540, 0.00053208};
  Real ax = std::fabs(x), x2, sum;
  if (ax > 0 && ax <= Real(2)) {
    x2 = ax / Real(2);
    x2 = x2 * x2;
    sum = Real(a[6]);
    for (int i = 5; i >= 0; --i) sum = sum * x2 + Real(a[i]);
    sum = -std::log(ax / Real(2)) * bessi0<Real>(x) + sum;
  } else if (ax > Real(2)) {
    x2 = Real(2) / ax;
    sum = Real(b[6]);
    for (int i = 5; i >= 0; --i) sum = sum * x2 + Real(b[i]);
    sum = std::exp(-ax) * sum / std::sqrt(ax);
  } else {
    sum = Real(1.0e20);
  }
  return sum;
}

// ---------------------------------------------------------------- Kirkland
// vatom: 3D atomic potential in volts, U.cu:160-193.  r2 is the squared
// distance; r is clamped to 0.25 A.
template <class Real> Real vatom(int Z, Real r2, const float* fp) {
  const float* p = fp + (Z - 1) * 12;
  Real r = std::sqrt(std::fabs(r2));
  if (r < Real(0.25)) r = Real(0.25);
  Real suml = 0, sumg = 0;
  Real x = Real(2) * Real(3.141592654) * r;
  for (int i = 0; i < 6; i += 2) suml += (Real(1) / r) * Real(p[i]) * std::exp(-x * std::sqrt(Real(p[i + 1])));
  x = Real(3.141592654) * r;
  x = x * x;
  for (int i = 6; i < 12; i += 2)
    sumg += Real(p[i]) * std::exp(-x / Real(p[i + 1])) * std::pow(Real(p[i + 1]), Real(-1.5));
  return Real(150.4121417) * suml + Real(266.5157269) * sumg;
}

// vzatom2: projected potential in V*A, U.cu:119-152.
template <class Real> Real vzatom(int Z, Real r2, const float* fp) {
  const float* p = fp + (Z - 1) * 12;
  Real r = std::sqrt(std::fabs(r2));
  if (r < Real(0.25)) r = Real(0.25);
  Real suml = 0, sumg = 0;
  Real x = Real(2) * Real(3.141592654) * r;
  for (int i = 0; i < 6; i += 2) suml += Real(p[i]) * bessk0<Real>(x * std::sqrt(Real(p[i + 1])));
  x = Real(3.141592654) * r;
  x = x * x;
  for (int i = 6; i < 12; i += 2) sumg += Real(p[i]) * std::exp(-x / Real(p[i + 1])) / Real(p[i + 1]);
  return Real(300.8242834) * suml + Real(150.4121417) * sumg;
}

// ---------------------------------------------------------------- FFT
// Unnormalised complex DFT, forward exp(-2 pi i), inverse exp(+2 pi i): the
// cuFFT C2C convention the reference relies on (U.cu:1961, 2119-2121).
// Radix-2 for powers of two, a direct O(n^2) sum otherwise.
template <class Real> struct FFT1D {
  int n = 0;
  bool pow2 = false;
  std::vector<std::complex<Real>> tw;  // exp(-2 pi i k / n), k < n
  std::vector<int> rev;
  explicit FFT1D(int n_ = 0) { init(n_); }
  void init(int n_) {
    n = n_;
    if (n <= 0) return;
    pow2 = (n & (n - 1)) == 0;
    tw.resize(n);
    for (int k = 0; k < n; ++k) {
      double a = -2.0 * M_PI * (double)k / (double)n;
      tw[k] = std::complex<Real>((Real)std::cos(a), (Real)std::sin(a));
    }
    if (pow2) {
      rev.assign(n, 0);
      int lg = 0;
      while ((1 << lg) < n) ++lg;
      for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b) if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        rev[i] = r;
      }
    }
  }
  // in-place transform of n elements with stride s; scratch holds >= n values
  void run(std::complex<Real>* d, int s, bool inverse, std::complex<Real>* scratch) const {
    if (pow2) {
      for (int i = 0; i < n; ++i) scratch[rev[i]] = d[(size_t)i * s];
      for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len)
          for (int j = 0; j < half; ++j) {
            std::complex<Real> w = tw[j * step];
            if (inverse) w = std::conj(w);
            std::complex<Real> u = scratch[i + j], v = scratch[i + j + half] * w;
            scratch[i + j] = u + v;
            scratch[i + j + half] = u - v;
          }
      }
      for (int i = 0; i < n; ++i) d[(size_t)i * s] = scratch[i];
    } else {
      for (int k = 0; k < n; ++k) {
        std::complex<Real> acc = 0;
        for (int j = 0; j < n; ++j) {
          std::complex<Real> w = tw[(int)(((long long)j * k) % n)];
          if (inverse) w = std::conj(w);
          acc += d[(size_t)j * s] * w;
        }
        scratch[k] = acc;
      }
      for (int k = 0; k < n; ++k) d[(size_t)k * s] = scratch[k];
    }
  }
};

template <class Real> struct FFT2D {
  int nx = 0, ny = 0;
  FFT1D<Real> fx, fy;
  void init(int nx_, int ny_) { nx = nx_; ny = ny_; fx.init(nx); fy.init(ny); }
  // data is [ny][nx] row-major
  void run(std::complex<Real>* d, bool inverse, bool threaded = true) const {
    auto rows = [&](int b, int e) {
      std::vector<std::complex<Real>> s(std::max(nx, ny));
      for (int y = b; y < e; ++y) fx.run(d + (size_t)y * nx, 1, inverse, s.data());
    };
    auto cols = [&](int b, int e) {
      std::vector<std::complex<Real>> s(std::max(nx, ny));
      for (int x = b; x < e; ++x) fy.run(d + x, nx, inverse, s.data());
    };
    if (threaded) { parallel_for(ny, rows); parallel_for(nx, cols); }
    else { rows(0, ny); cols(0, nx); }
  }
};

// ---------------------------------------------------------------- Philox
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based
// generator used for the frozen-phonon displacements.  It replaces the
// reference's unseeded rand() stream (U.cu:29-37), which cannot be replayed.
inline void philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    unsigned __int128 p0 = (unsigned __int128)M0 * c0;
    unsigned __int128 p1 = (unsigned __int128)M1 * c2;
    uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Gaussian displacements for frozen-phonon configuration `config`: atom i,
// axis a uses Philox block (i, config, a, 0) under key (seed, 0x4D534C); the
// first two 64-bit outputs feed a Box-Muller transform.  Displacement =
// float(sigma * N(0,1)), added to the float coordinate as U.cu:1549-1551 adds
// TDSRand().
inline double philox_normal(uint64_t seed, uint64_t config, uint64_t atom, uint64_t axis) {
  uint64_t ctr[4] = {atom, config, axis, 0}, key[2] = {seed, 0x4D534CULL}, o[4];
  philox4x64(ctr, key, o);
  double u1 = ((double)(o[0] >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0,1]
  double u2 = (double)(o[1] >> 11) * (1.0 / 9007199254740992.0);          // [0,1)
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// ---------------------------------------------------------------- detector
// Radial table lookup of DQEFunc (U.cu:1120-1132): signed frequency indices
// (x - N for x >= N/2), r = sqrtf(x'^2 + y'^2) / binning in FP32, entry
// floor(r) up to 724, else 0.0001.
inline float detector_table(const float* tab, int x, int y, int nx, int ny, int binning) {
  const int xd = x < nx / 2 ? x : x - nx, yd = y < ny / 2 ? y : y - ny;
  volatile float r = std::sqrt((float)(xd * xd + yd * yd));
  r = r / (float)binning;
  const int i = (int)std::floor((float)r);
  return i <= 724 ? tab[i] : 0.0001f;
}

// Standard normal of the detector's shot noise at pixel i (Philox block
// (i, 0xDE7EC7, 3, 0) under key (seed, 0x4D534C)), clamped to +-3 as the
// reference clamps its rand() Box-Muller draw (U.cu:2385-2392).
inline double detector_normal(uint64_t seed, uint64_t i, uint64_t axis = 3) {
  const double g = philox_normal(seed, 0xDE7EC7ULL, i, axis);
  return g < -3 ? -3 : (g > 3 ? 3 : g);
}

// Readout noise of GetImValue (U.cu:2497-2508): 4 + 1.33 g, draw `counter`
// (Philox block (counter, 0xDE7EC7, 4, 0)).
inline float readout_noise(uint64_t seed, uint64_t counter) {
  return (float)(4.0 + detector_normal(seed, counter, 4) * 1.33);
}

// Stochastic detector model of SimulateImage (U.cu:2345-2428): the
// intensity image I -> FFT nf -> * sqrt(DQE(r)) -> IFFT nf -> modulus a ->
// counts round(dose a + sqrt(|dose a|) g) -> FFT nf -> * NTF(r) ->
// IFFT (64 nf) (8 CCD counts per electron, squared).  out: the complex
// image; stats: |out| at its first maximum and minimum, and the standard
// deviation of |out| over the top-left min(N, 512)^2 pixels (the
// reference's contrast region, U.cu:2454-2491).
template <class Real>
void detector_image(int nx, int ny, const std::vector<Real>& I, float dose, uint64_t seed, const float* dqe,
                    const float* ntf, int binning, std::vector<std::complex<Real>>& out, double stats[3]) {
  using C = std::complex<Real>;
  const size_t n = (size_t)nx * ny;
  const Real nf = Real(1) / std::sqrt(Real(n));
  FFT2D<Real> f;
  f.init(nx, ny);
  std::vector<C> F(n);
  for (size_t i = 0; i < n; ++i) F[i] = C(I[i], 0);
  f.run(F.data(), false);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x)
      F[(size_t)y * nx + x] *= nf * std::sqrt(Real(detector_table(dqe, x, y, nx, ny, binning)));
  f.run(F.data(), true);
  for (size_t i = 0; i < n; ++i) {
    const Real a = std::abs(F[i] * nf);
    const Real m = Real(dose) * a;
    F[i] = C(std::round(m + std::sqrt(std::fabs(m)) * Real(detector_normal(seed, i))), 0);
  }
  f.run(F.data(), false);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) F[(size_t)y * nx + x] *= nf * Real(detector_table(ntf, x, y, nx, ny, binning));
  f.run(F.data(), true);
  out.resize(n);
  for (size_t i = 0; i < n; ++i) out[i] = F[i] * (Real(64) * nf);
  size_t imx = 0, imn = 0;
  for (size_t i = 1; i < n; ++i) {
    if (std::abs(out[i]) > std::abs(out[imx])) imx = i;
    if (std::abs(out[i]) < std::abs(out[imn])) imn = i;
  }
  stats[0] = (double)std::abs(out[imx]);
  stats[1] = (double)std::abs(out[imn]);
  const int sw = std::min(nx, 512), sh = std::min(ny, 512);
  double mean = 0, var = 0;
  for (int y = 0; y < sh; ++y)
    for (int x = 0; x < sw; ++x) mean += (double)std::abs(out[(size_t)y * nx + x]);
  mean /= (double)sw * sh;
  for (int y = 0; y < sh; ++y)
    for (int x = 0; x < sw; ++x) {
      const double d = (double)std::abs(out[(size_t)y * nx + x]) - mean;
      var += d * d;
    }
  stats[2] = std::sqrt(var / ((double)sw * sh));
}

// ---------------------------------------------------------------- element map
// GetZNum, U.cu:1188-1243: the reference's 25-symbol table; any other symbol
// (including Ti) maps to hydrogen.
inline int ref_znum(const std::string& s) {
  static const char* sym[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Na", "Mg", "Si", "P",
                              "S", "Cl", "Ca", "Cr", "Br", "Fe", "Sr", "Ru", "La", "Sm", "Ta", "W"};
  static const int z[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 14, 15, 16, 17, 20, 24, 35, 26, 38, 44, 57, 62, 73, 74};
  for (int i = 0; i < 25; ++i) if (s == sym[i]) return z[i];
  return 1;
}

// ---------------------------------------------------------------- simulation
struct Atoms {
  std::vector<int> Z;
  std::vector<float> x, y, z;
};

template <class Real> struct Sim {
  using C = std::complex<Real>;
  // structure (U.cuh:127-133) and binned copy (U.cu:1358-1536)
  Atoms atoms;
  float MaxX = 0, MinX = 0, MaxY = 0, MinY = 0, MaxZ = 0, MinZ = 0;
  int nSlab = 0;                        // 1 A binning slabs
  std::vector<int> bids, zids;          // per input atom
  std::vector<int> start;               // CSR, nSlab*2500+1
  std::vector<float> bx, by, bz;        // binned positions relative to Minimum*
  std::vector<int> bZ;
  std::vector<float> fparams;           // 103*12
  // microscope (U.cu:1749-1761)
  float kV = 0, df = 0, A1 = 0, A1phi = 0, A2 = 0, A2phi = 0, Cs = 0, beta = 0, delta = 0, obj = 0;
  // calculation variables (U.cu:1763-1875)
  int resX = 0, resY = 0;
  float ps = 0, offx = 0, offy = 0, ksizex = 0, ksizey = 0;
  Real wavel = 0, sigma2 = 0;           // evaluated in Real
  float wavel_f = 0;                    // FP32 value, for FP32-decided masks
  float dzFD = 0;                       // FP32 as in the reference
  int numberOfSlices = 0;
  int loadblocksx = 0, loadblocksy = 0, loadslicesz = 0;
  float kmax = 0;
  std::vector<float> kx, ky;            // FP32 frequencies, U.cu:1877-1903
  Real nf = 0;                          // 1/sqrt(resX*resY)
  float conventionaldz = 1;
  FFT2D<Real> fft;
  std::vector<C> PsiMinus, Psi, PsiPlus, prop;
  std::vector<Real> TDS;
  std::vector<Real> image;
  // compat=physical: the reference quirks below fixed instead of reproduced
  //   Q2 slice i integrates [H - i dz, H - (i-1) dz) (the reference starts one
  //      slice lower, U.cu:2097, and never integrates the top slab)
  //   Q3 the slab window reaches every atom within sqrt(5) A of the slice's
  //      samples (the reference's fixed +-3 one-angstrom slabs, U.cu:602-605,
  //      truncate slices thicker than ~0.76 A)
  //   Q5 band limit at 2/3 of the Nyquist frequency (U.cu:1929-1953 uses about
  //      half of it)
  //   Q6 probe normalised to unit intensity sum |psi|^2 = 1 (U.cu:2698-2705
  //      divides by sum |psi|)
  //   Q7 CBED accumulates the unitary diffracted intensity |FFT psi|^2 / (nx ny)
  //      (U.cu:2741-2755 accumulates the unnormalised amplitude)
  bool physical = false;

  size_t npix() const { return (size_t)resX * resY; }

  // LoadAtomFile extents, U.cu:1323-1355: first arg-max/arg-min per axis, +-2.
  void set_atoms(const Atoms& a) {
    atoms = a;
    size_t n = a.x.size();
    size_t mx = 0, nx = 0, my = 0, ny = 0, mz = 0, nz = 0;
    for (size_t i = 1; i < n; ++i) {
      if (a.x[i] > a.x[mx]) mx = i;
      if (a.y[i] > a.y[my]) my = i;
      if (a.z[i] > a.z[mz]) mz = i;
      if (a.x[i] < a.x[nx]) nx = i;
      if (a.y[i] < a.y[ny]) ny = i;
      if (a.z[i] < a.z[nz]) nz = i;
    }
    MaxX = a.x[mx] + 2; MinX = a.x[nx] - 2;
    MaxY = a.y[my] + 2; MinY = a.y[ny] - 2;
    MaxZ = a.z[mz] + 2; MinZ = a.z[nz] - 2;
  }

  // atombin2 (U.cu:277-303) + host CSR (U.cu:1405-1518).  dx/dy/dz are
  // optional frozen-phonon displacements (U.cu:1549-1551).  Returns false if
  // an atom falls outside the 50x50 x nSlab bin grid (undefined behaviour in
  // the reference).
  bool bin(const float* dx, const float* dy, const float* dzp) {
    const int xB = 50, yB = 50;
    float h = MaxZ - MinZ;
    nSlab = (int)std::ceil(h / 1.0f);
    nSlab += (nSlab == 0);
    size_t n = atoms.x.size();
    std::vector<float> X(n), Y(n), Zp(n);
    bids.assign(n, 0); zids.assign(n, 0);
    volatile float fxB = (float)xB, fyB = (float)yB;
    for (size_t i = 0; i < n; ++i) {
      X[i] = atoms.x[i] + (dx ? dx[i] : 0.0f);
      Y[i] = atoms.y[i] + (dy ? dy[i] : 0.0f);
      Zp[i] = atoms.z[i] + (dzp ? dzp[i] : 0.0f);
      volatile float t1 = X[i] - MinX, t2 = MaxX - MinX;
      volatile float t3 = t1 / t2;
      volatile float t4 = t3 * fxB;
      int bidx = (int)std::floor((float)t4);
      volatile float u1 = Y[i] - MinY, u2 = MaxY - MinY;
      volatile float u3 = u1 / u2;
      volatile float u4 = u3 * fyB;
      int bidy = (int)std::floor((float)u4);
      volatile float w1 = MaxZ - Zp[i];
      volatile float w2 = w1 / 1.0f;
      int zid = (int)std::floor((float)w2);
      if (bidx < 0 || bidx >= xB || bidy < 0 || bidy >= yB || zid < 0 || zid >= nSlab) return false;
      bids[i] = bidx + xB * bidy;
      zids[i] = zid;
    }
    // counting sort on key zid*2500 + bid, stable in input order
    size_t nb = (size_t)nSlab * xB * yB;
    start.assign(nb + 1, 0);
    for (size_t i = 0; i < n; ++i) start[(size_t)zids[i] * xB * yB + bids[i] + 1]++;
    for (size_t b = 0; b < nb; ++b) start[b + 1] += start[b];
    std::vector<int> fill(start.begin(), start.end() - 1);
    bx.resize(n); by.resize(n); bz.resize(n); bZ.resize(n);
    for (size_t i = 0; i < n; ++i) {
      int d = fill[(size_t)zids[i] * xB * yB + bids[i]]++;
      bx[d] = X[i] - MinX;
      by[d] = Y[i] - MinY;
      bz[d] = Zp[i] - MinZ;
      bZ[d] = atoms.Z[i];
    }
    return true;
  }

  void apply_microscope(float kV_, float df_, float a1, float a1p, float cs, float b, float d, float o) {
    kV = kV_; df = df_; A1 = a1; A1phi = a1p; Cs = cs; beta = b; delta = d; obj = o;
  }

  // SetCalculationVariables / STEMSetCalculationVariables, U.cu:1763-1875 and
  // 2567-2670.  The FD step size, slice count and bin window are FP32 host
  // quantities in the reference and are reproduced in FP32; lambda and sigma
  // are also kept in Real.
  void set_calc(float ps_, float df_, int rx, int ry, float sizeX, float sizeY, float sizeZ, float ox, float oy,
                bool stem) {
    (void)sizeX; (void)sizeY;
    resX = rx; resY = ry; df = df_; offx = stem ? 0.f : ox; offy = stem ? 0.f : oy; ps = ps_;
    ksizex = ps * (float)resX;
    ksizey = ps * (float)resY;
    // FP32 path, same operation order as U.cu:1791-1799
    float Pi = 3.1415926f, V = kV, echarge = 1.6e-019f;
    float wl = 6.63e-034f * 3e+008f / std::sqrt((echarge * V * 1000.f * (2.f * 9.11e-031f * 9e+016f + echarge * V * 1000.f))) * 1e+010f;
    float s2 = (2.f * Pi / (wl * V * 1000.f)) * ((9.11e-031f * 9e+016f + echarge * V * 1000.f) / (2.f * 9.11e-031f * 9e+016f + echarge * V * 1000.f));
    wavel_f = wl;
    {
      Real Pd = Real(3.1415926), Vd = Real(kV), e = Real(1.6e-19);
      Real m = Real(9.11e-31), c2 = Real(9e16);
      wavel = Real(6.63e-34) * Real(3e8) / std::sqrt(e * Vd * Real(1000) * (Real(2) * m * c2 + e * Vd * Real(1000))) * Real(1e10);
      sigma2 = (Real(2) * Pd / (wavel * Vd * Real(1000))) * ((m * c2 + e * Vd * Real(1000)) / (Real(2) * m * c2 + e * Vd * Real(1000)));
    }
    float V2 = V * 1000.f;
    float p1 = (float)resX / (2.f * ksizex), p2 = (float)resY / (2.f * ksizey);
    float ke2 = 0.666666f * (p1 * p1 + p2 * p2);
    float qA = (ke2 * ke2 * 16.f * Pi * Pi * Pi * Pi) - (32.f * Pi * Pi * Pi * ke2 * s2 * V2 / wl) + (16.f * Pi * Pi * s2 * s2 * V2 * V2 / (wl * wl));
    float qB = 16.f * Pi * Pi * (ke2 - (s2 * V2 / (Pi * wl)) - (1.f / (4.f * wl * wl)));
    float qC = 3.f;
    float disc = qB * qB - 4.f * qA * qC;
    float b24ac = std::sqrt(disc);
    float msd = (-qB + b24ac) / (2.f * qA);
    msd = (float)(0.99 * (double)std::sqrt(msd));
    if (stem && msd > 0.06) msd = 0.06f;
    int nFD = (int)std::ceil(sizeZ / msd);
    nFD += (nFD == 0);
    numberOfSlices = nFD;
    dzFD = msd;
    if (MaxX > MinX && MaxY > MinY) {
      loadblocksx = (int)std::ceil(std::sqrt(5.0f) / ((MaxX - MinX) / 50.f));
      loadblocksy = (int)std::ceil(std::sqrt(5.0f) / ((MaxY - MinY) / 50.f));
    }
    loadslicesz = (int)std::ceil(std::sqrt(5.0f) / 1.0f);
  }

  // PreCalculateFrequencies, U.cu:1877-1964.  Nyquist bin positive; kmax is
  // half of the (N/2-1) frequency, decided in FP32.
  void precalc() {
    int imidx = (int)std::floor(resX / 2 + 0.5), imidy = (int)std::floor(resY / 2 + 0.5);
    kx.resize(resX); ky.resize(resY);
    for (int i = 1; i <= resX; ++i) kx[i - 1] = ((i - 1) > imidx) ? (float)((i - 1) - resX) / ksizex : (float)(i - 1) / ksizex;
    for (int i = 1; i <= resY; ++i) ky[i - 1] = ((i - 1) > imidy) ? (float)((i - 1) - resY) / ksizey : (float)(i - 1) / ksizey;
    float hx = kx[imidx - 1] * 1 / 2, hy = ky[imidy - 1] * 1 / 2;
    float kmx = (float)std::pow(hx, 2), kmy = (float)std::pow(hy, 2);
    kmax = (kmy <= kmx) ? kmy : kmx;
    kmax = std::sqrt(kmax);
    if (physical) kmax = (2.0f / 3.0f) * std::min(kx[imidx], ky[imidy]);
    fft.init(resX, resY);
    nf = Real(1) / std::sqrt(Real(resX) * Real(resY));
  }

  // CreatePropsKernel, U.cu:399-428: exp(-i pi dz lambda k^2) inside |k|<kmax.
  void make_prop(std::vector<C>& P, Real dz) const {
    P.assign(npix(), C(0, 0));
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        float k2f = kx[x] * kx[x] + ky[y] * ky[y];
        if (std::sqrt(k2f) < kmax) {
          Real k2 = Real(kx[x]) * Real(kx[x]) + Real(ky[y]) * Real(ky[y]);
          Real a = Real(3.14159265) * dz * wavel * k2;
          P[(size_t)y * resX + x] = C(std::cos(a), -std::sin(a));
        }
      }
  }

  // Bin-window iteration shared by the three potential modes: every atom in
  // slabs [k0,k1] whose xy bin lies within reach (sqrt(5) A) of the pixel.
  // The reference's per-thread-block window (U.cu:620-631) is a superset of
  // this set; the rad2 < 5 test then selects exactly the same atoms.
  template <class F> void for_atoms(Real px, Real py, int k0, int k1, F&& f) const {
    Real wx = (Real(MaxX) - Real(MinX)) / 50, wy = (Real(MaxY) - Real(MinY)) / 50;
    Real reach = std::sqrt(Real(5));
    int i0 = std::max(0, (int)std::floor((px - reach) / wx) - 1), i1 = std::min(49, (int)std::floor((px + reach) / wx) + 1);
    int j0 = std::max(0, (int)std::floor((py - reach) / wy) - 1), j1 = std::min(49, (int)std::floor((py + reach) / wy) + 1);
    for (int k = k0; k <= k1; ++k)
      for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
          size_t b = (size_t)k * 2500 + (size_t)j * 50 + i;
          for (int l = start[b]; l < start[b + 1]; ++l) f(l);
        }
  }

  // FP32 cutoff decisions, as the reference kernels take them in FP32
  // (U.cu:637-640, 784, 931): pixel position fma(x, ps, off), squared
  // distances fma(dy, dy, dx*dx) and fma(dz, dz, rxy).  The potential value
  // itself is evaluated in Real.
  static float pix_f(int i, float ps, float off) { return std::fma((float)i, ps, off); }
  static float d2xy_f(float px, float py, float ax, float ay) {
    float dx = px - ax, dy = py - ay;
    return std::fma(dy, dy, dx * dx);
  }
  static float d2z_f(float rxy, float z, float az) {
    float dd = z - az;
    return std::fma(dd, dd, rxy);
  }
  static float sample_z_f(float zf, int m, float dz) { return zf - ((float)m * dz) / 20.0f; }

  // binnedAtomPKernel3, U.cu:582-653: 20-point z integration of vatom over
  // the slice below z, slabs floor(H-z) +- 3, transmission exp(i sigma V).
  // mode 0 returns the transmission, mode 3 returns the integrated potential
  // (sumz) as the real part.
  // FP32 slice position of conventional slice `it`: MaxZ - MinZ - dz*it as
  // the reference computes it in float (U.cu:2097, 2052)
  float slice_z(float dz, int it) const {
    float h = MaxZ - MinZ;
    float d = dz * (float)it;
    return h - d;
  }
  // top of conventional slice `it`: slice_z, or one slice higher under
  // compat=physical (Q2)
  float conv_z(float dz, int it) const { return slice_z(dz, physical ? it - 1 : it); }
  // slab index floor((MaxZ - MinZ - z) / 1), decided in FP32 as in U.cu:602
  int slab_of(float z) const {
    float h = MaxZ - MinZ;
    float d = h - z;
    return (int)std::floor(d / 1.0f);
  }

  void potential_conv(Real z, Real convdz, std::vector<C>& out, int mode = 0) const {
    out.assign(npix(), C(0, 0));
    int kc = slab_of((float)z);
    // Q3 (compat=physical): down to the slab of the lowest sample, +-3
    int kb = physical ? slab_of((float)z - (float)convdz) : kc;
    int k0 = std::max(0, kc - loadslicesz), k1 = std::min(nSlab - 1, kb + loadslicesz);
    parallel_for(resY, [&](int yb, int ye) {
      for (int y = yb; y < ye; ++y)
        for (int x = 0; x < resX; ++x) {
          Real px = Real(offx) + Real(x) * Real(ps), py = Real(offy) + Real(y) * Real(ps);
          float pxf = pix_f(x, ps, offx), pyf = pix_f(y, ps, offy);
          Real sumz = 0;
          for_atoms(px, py, k0, k1, [&](int l) {
            float rxyf = d2xy_f(pxf, pyf, bx[l], by[l]);
            if (rxyf >= 5.0f) return;
            Real ddx = px - Real(bx[l]), ddy = py - Real(by[l]);
            Real rxy = ddx * ddx + ddy * ddy;
            for (int m = 1; m <= 20; ++m) {
              if (!(d2z_f(rxyf, sample_z_f((float)z, m, (float)convdz), bz[l]) < 5.0f)) continue;
              Real z2 = z - Real(m) * convdz / Real(20);
              Real ddz = z2 - Real(bz[l]);
              Real r2 = rxy + ddz * ddz;
              sumz += convdz / Real(20) * vatom<Real>(bZ[l], r2, fparams.data());
            }
          });
          size_t id = (size_t)y * resX + x;
          if (mode == 3) out[id] = C(sumz, 0);
          else out[id] = C(std::cos(sigma2 * sumz), std::sin(sigma2 * sumz));
        }
    });
  }

  // Projected (Bessel/Gaussian) mode, the unlaunched
  // binnedAtomPKernelConventional3 (U.cu:730-795) generalised to a slice of
  // thickness convdz: slabs floor(H-z) .. floor(H-z+convdz)-1, xy distance
  // only, sum of vzatom2, transmission exp(i sigma V).
  void potential_proj(Real z, Real convdz, std::vector<C>& out, int mode = 0) const {
    out.assign(npix(), C(0, 0));
    float hd = (MaxZ - MinZ) - (float)z;
    float he = hd + (float)convdz;
    int k0 = (int)std::floor(hd), k1 = (int)std::floor(he) - 1;
    k0 = std::max(0, k0); k1 = std::min(nSlab - 1, k1);
    parallel_for(resY, [&](int yb, int ye) {
      for (int y = yb; y < ye; ++y)
        for (int x = 0; x < resX; ++x) {
          Real px = Real(offx) + Real(x) * Real(ps), py = Real(offy) + Real(y) * Real(ps);
          float pxf = pix_f(x, ps, offx), pyf = pix_f(y, ps, offy);
          Real sumz = 0;
          if (k0 <= k1)
            for_atoms(px, py, k0, k1, [&](int l) {
              if (!(d2xy_f(pxf, pyf, bx[l], by[l]) < 5.0f)) return;
              Real ddx = px - Real(bx[l]), ddy = py - Real(by[l]);
              Real r2 = ddx * ddx + ddy * ddy;
              sumz += vzatom<Real>(bZ[l], r2, fparams.data());
            });
          size_t id = (size_t)y * resX + x;
          if (mode == 3) out[id] = C(sumz, 0);
          else out[id] = C(std::cos(sigma2 * sumz), std::sin(sigma2 * sumz));
        }
    });
  }

  // binnedAtomPKernelFD, U.cu:883-946: point-sampled vatom on the plane z,
  // slabs +-3, no xy offset; real potential.
  void potential_fd(Real z, std::vector<C>& out) const {
    out.assign(npix(), C(0, 0));
    int kc = slab_of((float)z);
    int k0 = std::max(0, kc - loadslicesz), k1 = std::min(nSlab - 1, kc + loadslicesz);
    parallel_for(resY, [&](int yb, int ye) {
      for (int y = yb; y < ye; ++y)
        for (int x = 0; x < resX; ++x) {
          Real px = Real(x) * Real(ps), py = Real(y) * Real(ps);
          float pxf = pix_f(x, ps, 0.0f), pyf = pix_f(y, ps, 0.0f);
          Real sumz = 0;
          for_atoms(px, py, k0, k1, [&](int l) {
            if (!(d2z_f(d2xy_f(pxf, pyf, bx[l], by[l]), (float)z, bz[l]) < 5.0f)) return;
            Real ddx = px - Real(bx[l]), ddy = py - Real(by[l]), ddz = z - Real(bz[l]);
            Real r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            sumz += vatom<Real>(bZ[l], r2, fparams.data());
          });
          out[(size_t)y * resX + x] = C(sumz, 0);
        }
    });
  }

  // binnedAtomPKernelRS, U.cu:656-727: 10-point integration of vatom over the
  // 1 A below z (samples z - m/10, weight 1/10), slabs floor(H-z) +- 3, the
  // block offset in xy; real potential (no sigma).
  void potential_rs(Real z, std::vector<C>& out) const {
    out.assign(npix(), C(0, 0));
    int kc = slab_of((float)z);
    int k0 = std::max(0, kc - loadslicesz), k1 = std::min(nSlab - 1, kc + loadslicesz);
    parallel_for(resY, [&](int yb, int ye) {
      for (int y = yb; y < ye; ++y)
        for (int x = 0; x < resX; ++x) {
          Real px = Real(offx) + Real(x) * Real(ps), py = Real(offy) + Real(y) * Real(ps);
          float pxf = pix_f(x, ps, offx), pyf = pix_f(y, ps, offy);
          Real sumz = 0;
          for_atoms(px, py, k0, k1, [&](int l) {
            float rxyf = d2xy_f(pxf, pyf, bx[l], by[l]);
            if (rxyf >= 5.0f) return;
            Real ddx = px - Real(bx[l]), ddy = py - Real(by[l]);
            Real rxy = ddx * ddx + ddy * ddy;
            for (int m = 1; m <= 10; ++m) {
              volatile float z2f = (float)z - (float)m / 10.0f;
              if (!(d2z_f(rxyf, z2f, bz[l]) < 5.0f)) continue;
              Real ddz = (z - Real(m) / Real(10)) - Real(bz[l]);
              sumz += vatom<Real>(bZ[l], rxy + ddz * ddz, fparams.data()) / Real(10);
            }
          });
          out[(size_t)y * resX + x] = C(sumz, 0);
        }
    });
  }

  // DifferentialXKernel + DifferentialYKernel (U.cu:798-844): per axis
  // (4/3 d1 - d2/12) / ps^2 with d1 = f[+1] + f[-1] - f[0] and
  // d2 = f[+2] + f[-2] - f[0] (f[0] subtracted once, as written there);
  // neighbours outside the grid read element 0 of the row / column.
  void rs_laplacian(const std::vector<C>& w, std::vector<C>& out) const {
    out.assign(npix(), C(0, 0));
    const Real s = Real(1) / (Real(ps) * Real(ps));
    auto nb = [](int i, int d, int n) {
      const int j = i + d;
      return (j >= 0 && j < n) ? j : 0;
    };
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        const C f = w[(size_t)y * resX + x];
        auto at = [&](int xx, int yy) { return w[(size_t)yy * resX + xx]; };
        const C d1x = at(nb(x, 1, resX), y) + at(nb(x, -1, resX), y) - f;
        const C d2x = at(nb(x, 2, resX), y) + at(nb(x, -2, resX), y) - f;
        const C d1y = at(x, nb(y, 1, resY)) + at(x, nb(y, -1, resY)) - f;
        const C d2y = at(x, nb(y, 2, resY)) + at(x, nb(y, -2, resY)) - f;
        out[(size_t)y * resX + x] = s * (Real(4) / Real(3) * (d1x + d1y) - (d2x + d2y) / Real(12));
      }
  }

  // MultisliceStepRS, U.cu:2126-2178 with RSMS (U.cu:846-877): the series
  // w1 = 1 + i p1 V, w_m = w_{m-1} + i p_m (L w_{m-1} + V w_{m-1}) for m = 2..10,
  // p_m = (1/m) lambda dz / (4 pi) with dz = 1 A and z = H - it.  The series
  // restarts from 1 every step; the wave becomes PsiPlus = w10 (step 1 keeps
  // PsiMinus, later steps set PsiMinus = w10 * Psi before Psi = w10).
  void step_rs(int it) {
    const float zf = (MaxZ - MinZ) - 1.0f * (float)it;
    std::vector<C> V, L;
    potential_rs(Real(zf), V);
    std::vector<C> w(npix());
    const Real base = Real(wavel_f) * Real(1) / (Real(4) * Real(3.1415927f));
    for (size_t i = 0; i < npix(); ++i) w[i] = C(1, base * V[i].real());
    for (int m = 2; m <= 10; ++m) {
      rs_laplacian(w, L);
      const Real pre = base / Real(m);
      for (size_t i = 0; i < npix(); ++i) w[i] += C(0, pre) * (L[i] + V[i].real() * w[i]);
    }
    if (it != 1)
      for (size_t i = 0; i < npix(); ++i) PsiMinus[i] = w[i] * Psi[i];
    Psi = w;
    PsiPlus = w;
  }

  // InitialiseWavefunctions(dz), U.cu:2003-2038 (conventional path: psi = 1).
  void init_plane(float dz) {
    conventionaldz = dz;
    prop.clear();
    PsiMinus.assign(npix(), C(1, 0));
    Psi.assign(npix(), C(1, 0));
    PsiPlus.assign(npix(), C(0, 0));
  }

  // InitialiseWavefunctions() for FD, U.cu:1966-2001: psi_- = 1 and psi =
  // propagate(psi_-, dzFD).
  void init_plane_fd() {
    conventionaldz = 1;
    PsiMinus.assign(npix(), C(1, 0));
    Psi = PsiMinus;
    propagate(Psi, Real(dzFD));
    PsiPlus.assign(npix(), C(0, 0));
  }

  // psi <- IFFT(P(dz) FFT(psi)) * nf^2 (U.cu:1986-1993)
  void propagate(std::vector<C>& w, Real dz) const {
    std::vector<C> P;
    make_prop(P, dz);
    fft.run(w.data(), false);
    for (size_t i = 0; i < npix(); ++i) w[i] *= P[i] * nf;
    fft.run(w.data(), true);
    for (auto& v : w) v *= nf;
  }

  // STEMinitializingKernel + STEMInitialiseWavefunctions, U.cu:354-383 and
  // 2672-2724: probe exp(i(theta - chi)) inside k < alpha/lambda, unnormalised
  // inverse FFT, divided by sum |psi| (L1 of the amplitude).
  void probe(int posx, int posy, std::vector<C>& w) const {
    w.assign(npix(), C(0, 0));
    float apert = (float)((double)obj * 0.001 / (double)wavel_f);
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        float kf = std::sqrt(kx[x] * kx[x] + ky[y] * ky[y]);
        if (kf < apert) {
          Real k0x = kx[x], k0y = ky[y];
          Real k2 = k0x * k0x + k0y * k0y;
          Real Pi = Real(3.14159265);
          Real chi = Pi * wavel * k2 * (Real(Cs) * wavel * wavel * k2 * Real(0.5) + Real(df));
          Real th = Real(2) * Pi * (k0x * Real(posx) * Real(ps) + k0y * Real(posy) * Real(ps));
          w[(size_t)y * resX + x] = C(std::cos(th - chi), std::sin(th - chi));
        }
      }
    fft.run(w.data(), true);
    Real s = 0;
    if (physical) {  // Q6: unit intensity
      for (auto& v : w) s += std::norm(v);
      s = std::sqrt(s);
    } else {
      for (auto& v : w) s += std::abs(v);
    }
    for (auto& v : w) v /= s;
  }

  void init_stem(int posx, int posy) {
    prop.clear();
    probe(posx, posy, PsiMinus);
    Psi = PsiMinus;
    propagate(Psi, Real(dzFD));
    PsiPlus.assign(npix(), C(0, 0));
  }

  // MultisliceStepConv, U.cu:2085-2124.
  void step_conv(int it, int potmode = 0) {
    Real z = Real(conv_z(conventionaldz, it));
    std::vector<C> t;
    if (potmode == 1) potential_proj(z, Real(conventionaldz), t);
    else potential_conv(z, Real(conventionaldz), t);
    std::vector<C>& src = (it == 1) ? PsiMinus : Psi;
    for (size_t i = 0; i < npix(); ++i) PsiPlus[i] = t[i] * src[i];
    if (it == 1 || prop.size() != npix()) make_prop(prop, Real(conventionaldz));
    fft.run(PsiPlus.data(), false);
    for (size_t i = 0; i < npix(); ++i) PsiPlus[i] *= prop[i] * nf;
    fft.run(PsiPlus.data(), true);
    for (size_t i = 0; i < npix(); ++i) Psi[i] = PsiPlus[i] * nf;
  }

  void bandlimit(std::vector<C>& w) const {
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        float k = std::sqrt(kx[x] * kx[x] + ky[y] * ky[y]);
        if (!(k <= kmax)) w[(size_t)y * resX + x] = C(0, 0);
      }
  }

  // MultisliceStep (finite difference), U.cu:2040-2083 with the kernels
  // BandLimitKernel (1025-1042), gradKernel (948-962) and
  // finiteDifferenceKernel (995-1023).
  // Test hook: a caller-given real potential replaces the atoms' one in
  // step_fd (empty: the potential of U.cu:883-946).
  std::vector<Real> fd_potential_override;

  void step_fd(int it) {
    std::vector<C> V;
    fd_potential(it, V);
    fd_update(V, PsiMinus, Psi, PsiPlus);
  }

  // The band-limited real potential of finite-difference step `it`
  // (U.cu:2050-2063): the atoms' point-sampled potential at z = H - dzFD it
  // (or the test hook's fixed one), FFT, band limit, IFFT / (nx ny).
  void fd_potential(int it, std::vector<C>& V) const {
    Real z = Real(slice_z(dzFD, it));
    if (fd_potential_override.size() == npix()) {
      V.resize(npix());
      for (size_t i = 0; i < npix(); ++i) V[i] = C(fd_potential_override[i], 0);
    } else {
      potential_fd(z, V);
    }
    fft.run(V.data(), false);
    bandlimit(V);
    fft.run(V.data(), true);
    for (auto& v : V) v *= nf * nf;
  }

  // One finite-difference update (U.cu:2064-2081) from the band-limited
  // potential V: Laplacian G of psi (gradKernel), the update of
  // finiteDifferenceKernel into psi_+, band limit of psi_+; then psi_- <- psi,
  // psi <- psi_+.
  void fd_update(const std::vector<C>& V, std::vector<C>& Pm, std::vector<C>& Ps, std::vector<C>& Pp) const {
    std::vector<C> G = Ps;
    fft.run(G.data(), false);
    Real p9 = Real(3.14159);
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        Real k2 = Real(kx[x]) * Real(kx[x]) + Real(ky[y]) * Real(ky[y]);
        G[(size_t)y * resX + x] *= nf * Real(-4) * p9 * p9 * k2;
      }
    fft.run(G.data(), true);
    for (auto& v : G) v *= nf;
    Real dz = Real(dzFD);
    C cPlus(1, Real(2) * p9 * dz / wavel), cMinus(1, -Real(2) * p9 * dz / wavel);
    C rcp = Real(1) / cPlus, ratio = cMinus / cPlus;
    Pp.resize(npix());
    for (size_t i = 0; i < npix(); ++i) {
      C a = Real(2) * Ps[i] - dz * dz * G[i] - dz * dz * Real(4) * p9 * sigma2 * V[i].real() * Ps[i] / wavel;
      Pp[i] = rcp * a - ratio * Pm[i];
    }
    fft.run(Pp.data(), false);
    bandlimit(Pp);
    fft.run(Pp.data(), true);
    for (auto& v : Pp) v *= nf * nf;
    Pm = Ps;
    Ps = Pp;
  }

  // AddTDSWaves, U.cu:2734-2766: TDS += |fftshift(FFT_unnormalised(psi))|.
  void add_tds() {
    if (TDS.size() != npix()) TDS.assign(npix(), 0);
    std::vector<C> F = Psi;
    fft.run(F.data(), false);
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        int x2 = (x >= resX / 2) ? x - resX / 2 : x + resX / 2;
        int y2 = (y >= resY / 2) ? y - resY / 2 : y + resY / 2;
        const C f = F[(size_t)y2 * resX + x2];
        TDS[(size_t)y * resX + x] += physical ? std::norm(f) * nf * nf : std::abs(f);  // Q7
      }
  }

  // Annular dark-field signal (absent from the reference): sum of
  // |FFT_unitary(psi)|^2 over lambda*|k| in [inner, outer) mrad.
  Real adf(const std::vector<C>& w, float inner_mrad, float outer_mrad) const {
    std::vector<C> F = w;
    fft.run(F.data(), false);
    float kin = inner_mrad * 1e-3f / wavel_f, kout = outer_mrad * 1e-3f / wavel_f;
    Real s = 0;
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        float k = std::sqrt(kx[x] * kx[x] + ky[y] * ky[y]);
        if (k >= kin && k < kout) s += std::norm(F[(size_t)y * resX + x]);
      }
    return s * nf * nf;
  }

  // Deterministic CTF image: SimulateImage steps U.cu:2322-2343 with
  // imagingKernel2 (U.cu:1044-1080).  I = |IFFT(CTF . FFT(psi) nf) nf|^2.
  // df_override < -1e29 keeps the stored defocus.
  void ctf_image(std::vector<Real>& I, const std::vector<C>& w, double df_use) const {
    std::vector<C> F = w;
    fft.run(F.data(), false);
    for (auto& v : F) v *= nf;
    float objap2 = ((obj * 0.001f) / wavel_f) * ((obj * 0.001f) / wavel_f);
    Real lam = wavel, b = Real(beta) / Real(1000), d = Real(delta), dfv = Real(df_use), cs = Real(Cs);
    Real factor = b * b / (Real(4) * lam * lam);
    for (int y = 0; y < resY; ++y)
      for (int x = 0; x < resX; ++x) {
        size_t id = (size_t)y * resX + x;
        float k2f = kx[x] * kx[x] + ky[y] * ky[y];
        if (!(k2f < objap2)) { F[id] = C(0, 0); continue; }
        Real k0x = kx[x], k0y = ky[y];
        Real k2 = k0x * k0x + k0y * k0y, k = std::sqrt(k2);
        Real pi1 = Real(3.141593);
        Real es_arg = pi1 * k * lam * Real(2) * dfv + Real(2) * pi1 * lam * lam * lam * cs * k2 * k;
        Real es = std::exp(-factor * es_arg * es_arg);
        Real ed = std::exp(Real(-0.25) * d * d * pi1 * pi1 * k2 * k2 * lam * lam);
        Real phi = std::atan2(k0y, k0x);
        Real chi = Real(3.1415926) * lam * k2 *
                   (Real(0.5) * cs * lam * lam * k2 + dfv + Real(A1) * std::sin(Real(2) * (phi - Real(A1phi))) +
                    Real(2) * Real(A2) * lam * k * std::sin(Real(3) * (phi - Real(A2phi))) / Real(3));
        F[id] = es * ed * F[id] * C(std::cos(chi), -std::sin(chi));
      }
    fft.run(F.data(), true);
    I.resize(npix());
    for (size_t i = 0; i < npix(); ++i) I[i] = std::norm(F[i] * nf);
  }
};

}  // namespace orc
