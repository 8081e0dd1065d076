// msl_oracle.cpp -- C-ABI over the CPU restatement in msl_oracle.hpp.
//
// TEST INFRASTRUCTURE ONLY (see msl_oracle.hpp).  Loaded by tests/ through
// oracle/oracle.py; bench.py times the float instantiation as its CPU
// baseline.  Waves cross this boundary as interleaved (re, im) doubles.
#include "msl_oracle.hpp"

#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>

using namespace orc;

namespace {

struct Handle {
  bool is_float = false;
  std::unique_ptr<Sim<double>> d;
  std::unique_ptr<Sim<float>> f;
};

template <class F> auto dispatch(Handle* h, F&& fn) {
  if (h->is_float) return fn(*h->f);
  return fn(*h->d);
}

template <class Real>
void to_complex(const double* in, size_t n, std::vector<std::complex<Real>>& out) {
  out.resize(n);
  for (size_t i = 0; i < n; ++i) out[i] = std::complex<Real>((Real)in[2 * i], (Real)in[2 * i + 1]);
}
template <class Real>
void from_complex(const std::vector<std::complex<Real>>& in, double* out) {
  for (size_t i = 0; i < in.size(); ++i) { out[2 * i] = (double)in[i].real(); out[2 * i + 1] = (double)in[i].imag(); }
}

// Batched STEM over one frozen-phonon configuration (the binned atoms
// currently held).  Transmission functions for every slice are built once
// and shared by all probes; each probe then runs the conventional step of
// U.cu:2085-2124 from its own probe (U.cu:2672-2724).  Threads split probes.
template <class Real>
int stem_batch(Sim<Real>& s, int nprobe, const int* posx, const int* posy, float dz, int nslices, int potmode,
               float adf_in, float adf_out, double* adf, double* exitw, double* cbed, int nthreads) {
  using C = std::complex<Real>;
  size_t np = s.npix();
  s.conventionaldz = dz;
  std::vector<std::vector<C>> t(nslices);
  for (int i = 1; i <= nslices; ++i) {
    Real z = Real(s.conv_z(dz, i));
    if (potmode == 1) s.potential_proj(z, Real(dz), t[i - 1]);
    else s.potential_conv(z, Real(dz), t[i - 1]);
  }
  std::vector<C> P;
  s.make_prop(P, Real(dz));
  int saved = num_threads();
  int nt = nthreads > 0 ? nthreads : saved;
  num_threads() = 1;  // inner FFTs single-threaded; parallel over probes
  std::vector<std::vector<Real>> cb(nt);
  auto worker = [&](int tid) {
    if (cbed) cb[tid].assign(np, 0);
    std::vector<C> w, F;
    for (int p = tid; p < nprobe; p += nt) {
      s.probe(posx[p], posy[p], w);
      for (int i = 0; i < nslices; ++i) {
        for (size_t k = 0; k < np; ++k) w[k] *= t[i][k];
        s.fft.run(w.data(), false, false);
        for (size_t k = 0; k < np; ++k) w[k] *= P[k] * s.nf;
        s.fft.run(w.data(), true, false);
        for (size_t k = 0; k < np; ++k) w[k] *= s.nf;
      }
      if (exitw) from_complex(w, exitw + (size_t)p * 2 * np);
      F = w;
      s.fft.run(F.data(), false, false);
      float kin = adf_in * 1e-3f / s.wavel_f, kout = adf_out * 1e-3f / s.wavel_f;
      Real acc = 0;
      for (int y = 0; y < s.resY; ++y)
        for (int x = 0; x < s.resX; ++x) {
          float k = std::sqrt(s.kx[x] * s.kx[x] + s.ky[y] * s.ky[y]);
          size_t id = (size_t)y * s.resX + x;
          if (k >= kin && k < kout) acc += std::norm(F[id]);
          if (cbed) {
            int x2 = (x >= s.resX / 2) ? x - s.resX / 2 : x + s.resX / 2;
            int y2 = (y >= s.resY / 2) ? y - s.resY / 2 : y + s.resY / 2;
            cb[tid][(size_t)y * s.resX + x] += std::abs(F[(size_t)y2 * s.resX + x2]);
          }
        }
      if (adf) adf[p] = (double)(acc * s.nf * s.nf);
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nt; ++i) th.emplace_back(worker, i);
  for (auto& x : th) x.join();
  num_threads() = saved;
  if (cbed) {
    for (size_t k = 0; k < np; ++k) {
      double acc = 0;
      for (int i = 0; i < nt; ++i) acc += (double)cb[i][k];
      cbed[k] = acc;
    }
  }
  return 0;
}

// Batched finite-difference STEM over the binned atoms held: every probe
// starts as STEMInitialiseWavefunctions does (probe, then one propagator
// step of dzFD for the second wave, U.cu:2672-2724) and runs nsteps
// MultisliceStep (U.cu:2040-2083); the band-limited potential of each step
// is computed once and shared.  Threads split probes.  adf[p] as in
// stem_batch.
template <class Real>
int fd_batch(Sim<Real>& s, int nprobe, const int* posx, const int* posy, int nsteps, float adf_in, float adf_out,
             double* adf, int nthreads) {
  using C = std::complex<Real>;
  std::vector<std::vector<C>> V(nsteps);
  for (int it = 1; it <= nsteps; ++it) s.fd_potential(it, V[it - 1]);
  int saved = num_threads();
  int nt = nthreads > 0 ? nthreads : saved;
  num_threads() = 1;  // inner FFTs single-threaded; parallel over probes
  auto worker = [&](int tid) {
    std::vector<C> Pm, Ps, Pp;
    for (int p = tid; p < nprobe; p += nt) {
      s.probe(posx[p], posy[p], Pm);
      Ps = Pm;
      s.propagate(Ps, Real(s.dzFD));
      for (int it = 0; it < nsteps; ++it) s.fd_update(V[it], Pm, Ps, Pp);
      adf[p] = (double)s.adf(Ps, adf_in, adf_out);
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nt; ++i) th.emplace_back(worker, i);
  for (auto& x : th) x.join();
  num_threads() = saved;
  return 0;
}

// Differentiable multislice, FP64 adjoint.  psi_i = K(t_i psi_{i-1}) with
// t_i = exp(i sigma V_i), K = IFFT P FFT nf^2 (the conventional step of
// U.cu:2085-2124 with a given potential).  Returns the exit intensities and,
// for a loss L(I), dL/dV from the upstream gradient dL/dI.
template <class Real>
int diff_run(Sim<Real>& s, int nslices, const float* V, int nprobe, const double* psi0, double* intensity,
             const double* dLdI, double* grad, float dz) {
  using C = std::complex<Real>;
  size_t np = s.npix();
  std::vector<C> P;
  s.make_prop(P, Real(dz));
  if (grad) std::fill(grad, grad + (size_t)nslices * np, 0.0);
  std::vector<std::vector<C>> store(nslices + 1);
  for (int p = 0; p < nprobe; ++p) {
    std::vector<C> w;
    to_complex(psi0 + (size_t)p * 2 * np, np, w);
    store[0] = w;
    for (int i = 0; i < nslices; ++i) {
      const float* v = V + (size_t)i * np;
      for (size_t k = 0; k < np; ++k) w[k] *= C(std::cos(s.sigma2 * Real(v[k])), std::sin(s.sigma2 * Real(v[k])));
      s.fft.run(w.data(), false);
      for (size_t k = 0; k < np; ++k) w[k] *= P[k];
      s.fft.run(w.data(), true);
      for (size_t k = 0; k < np; ++k) w[k] *= s.nf * s.nf;
      store[i + 1] = w;
    }
    for (size_t k = 0; k < np; ++k) intensity[(size_t)p * np + k] = (double)std::norm(w[k]);
    if (!dLdI || !grad) continue;
    // G = dL/dRe + i dL/dIm of the exit wave
    std::vector<C> G(np);
    for (size_t k = 0; k < np; ++k) G[k] = Real(2) * Real(dLdI[(size_t)p * np + k]) * w[k];
    for (int i = nslices - 1; i >= 0; --i) {
      // adjoint of K: IFFT conj(P) FFT nf^2
      s.fft.run(G.data(), false);
      for (size_t k = 0; k < np; ++k) G[k] *= std::conj(P[k]);
      s.fft.run(G.data(), true);
      for (size_t k = 0; k < np; ++k) G[k] *= s.nf * s.nf;
      const float* v = V + (size_t)i * np;
      double* g = grad + (size_t)i * np;
      for (size_t k = 0; k < np; ++k) {
        C t(std::cos(s.sigma2 * Real(v[k])), std::sin(s.sigma2 * Real(v[k])));
        C u = t * store[i][k];
        // dL/dV = Re(conj(G_u) * i sigma u)
        g[k] += (double)(-s.sigma2 * std::imag(std::conj(G[k]) * u));
        G[k] *= std::conj(t);
      }
    }
  }
  return 0;
}

bool parse_xyz(const char* path, bool refmap, Atoms& a) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  if (!std::getline(in, line)) return false;
  int n = 0;
  {
    std::istringstream ls(line);
    if (!(ls >> n) || n <= 0) return false;
  }
  std::vector<std::string> rest;
  while (std::getline(in, line)) rest.push_back(line);
  size_t li = 0;
  // line 2: atom line if it parses as "symbol x y z", otherwise a comment
  if (!rest.empty()) {
    std::istringstream ls(rest[0]);
    std::string sym;
    float x, y, z;
    if (!(ls >> sym >> x >> y >> z)) li = 1;
  }
  for (int i = 0; i < n; ++i) {
    while (li < rest.size() && rest[li].find_first_not_of(" \t\r") == std::string::npos) ++li;
    if (li >= rest.size()) return false;
    std::istringstream ls(rest[li++]);
    std::string sym;
    float x, y, z;
    if (!(ls >> sym >> x >> y >> z)) return false;
    int Z = 0;
    if (refmap) Z = ref_znum(sym);
    else return false;
    a.Z.push_back(Z); a.x.push_back(x); a.y.push_back(y); a.z.push_back(z);
  }
  return true;
}

}  // namespace

extern "C" {

void* orc_create(int use_float) {
  auto* h = new Handle;
  h->is_float = use_float != 0;
  if (h->is_float) h->f.reset(new Sim<float>);
  else h->d.reset(new Sim<double>);
  return h;
}
void orc_destroy(void* p) { delete (Handle*)p; }
void orc_set_threads(int n) { num_threads() = n > 0 ? n : (int)std::max(1u, std::thread::hardware_concurrency()); }

int orc_set_atoms(void* p, int n, const int* Z, const float* x, const float* y, const float* z) {
  if (n <= 0) return -1;
  Atoms a;
  a.Z.assign(Z, Z + n); a.x.assign(x, x + n); a.y.assign(y, y + n); a.z.assign(z, z + n);
  dispatch((Handle*)p, [&](auto& s) { s.set_atoms(a); return 0; });
  return 0;
}

int orc_load_xyz_refmap(void* p, const char* path) {
  Atoms a;
  if (!parse_xyz(path, true, a)) return -1;
  dispatch((Handle*)p, [&](auto& s) { s.set_atoms(a); return 0; });
  return (int)a.x.size();
}

void orc_get_extents(void* p, float* out) {
  dispatch((Handle*)p, [&](auto& s) {
    out[0] = s.MaxX; out[1] = s.MinX; out[2] = s.MaxY; out[3] = s.MinY; out[4] = s.MaxZ; out[5] = s.MinZ;
    return 0;
  });
}

int orc_bin(void* p, const float* dx, const float* dy, const float* dz) {
  return dispatch((Handle*)p, [&](auto& s) { return s.bin(dx, dy, dz) ? s.nSlab : -1; });
}

int orc_get_bins(void* p, int* bids, int* zids) {
  return dispatch((Handle*)p, [&](auto& s) {
    std::copy(s.bids.begin(), s.bids.end(), bids);
    std::copy(s.zids.begin(), s.zids.end(), zids);
    return (int)s.bids.size();
  });
}

int orc_get_csr(void* p, int* start, float* x, float* y, float* z, int* Z) {
  return dispatch((Handle*)p, [&](auto& s) {
    std::copy(s.start.begin(), s.start.end(), start);
    std::copy(s.bx.begin(), s.bx.end(), x);
    std::copy(s.by.begin(), s.by.end(), y);
    std::copy(s.bz.begin(), s.bz.end(), z);
    std::copy(s.bZ.begin(), s.bZ.end(), Z);
    return (int)s.start.size();
  });
}

void orc_set_fparams(void* p, const float* fp) {
  dispatch((Handle*)p, [&](auto& s) { s.fparams.assign(fp, fp + 103 * 12); return 0; });
}

void orc_apply_microscope(void* p, float kV, float df, float a1, float a1p, float cs, float b, float d, float o) {
  dispatch((Handle*)p, [&](auto& s) { s.apply_microscope(kV, df, a1, a1p, cs, b, d, o); return 0; });
}

void orc_set_calc(void* p, float ps, float df, int rx, int ry, float sx, float sy, float sz, float ox, float oy,
                  int stem) {
  dispatch((Handle*)p, [&](auto& s) { s.set_calc(ps, df, rx, ry, sx, sy, sz, ox, oy, stem != 0); s.precalc(); return 0; });
}

// out: wavel, sigma2, dzFD, numberOfSlices, kmax, loadblocksx, loadblocksy,
// loadslicesz, wavel_f, nSlab
void orc_get_scalars(void* p, double* out) {
  dispatch((Handle*)p, [&](auto& s) {
    out[0] = (double)s.wavel; out[1] = (double)s.sigma2; out[2] = s.dzFD; out[3] = s.numberOfSlices;
    out[4] = s.kmax; out[5] = s.loadblocksx; out[6] = s.loadblocksy; out[7] = s.loadslicesz;
    out[8] = s.wavel_f; out[9] = s.nSlab;
    return 0;
  });
}

void orc_get_freqs(void* p, float* kx, float* ky) {
  dispatch((Handle*)p, [&](auto& s) {
    std::copy(s.kx.begin(), s.kx.end(), kx);
    std::copy(s.ky.begin(), s.ky.end(), ky);
    return 0;
  });
}

// mode 0: conventional 3D-integrated transmission, 1: projected transmission,
// 2: FD point potential (real), 3: conventional integrated potential (real),
// 4: projected potential (real)
void orc_potential(void* p, int mode, double z, double convdz, double* out) {
  dispatch((Handle*)p, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    std::vector<std::complex<R>> t;
    if (mode == 0) s.potential_conv(R(z), R(convdz), t);
    else if (mode == 1) s.potential_proj(R(z), R(convdz), t);
    else if (mode == 2) s.potential_fd(R(z), t);
    else if (mode == 3) s.potential_conv(R(z), R(convdz), t, 3);
    else if (mode == 5) s.potential_rs(R(z), t);
    else s.potential_proj(R(z), R(convdz), t, 3);
    from_complex(t, out);
    return 0;
  });
}

void orc_prop(void* p, double dz, double* out) {
  dispatch((Handle*)p, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    std::vector<std::complex<R>> P;
    s.make_prop(P, R(dz));
    from_complex(P, out);
    return 0;
  });
}

void orc_init_plane(void* p, float dz) { dispatch((Handle*)p, [&](auto& s) { s.init_plane(dz); return 0; }); }
void orc_init_plane_fd(void* p) { dispatch((Handle*)p, [&](auto& s) { s.init_plane_fd(); return 0; }); }
void orc_init_stem(void* p, int px, int py, float dz) {
  dispatch((Handle*)p, [&](auto& s) { s.init_stem(px, py); s.conventionaldz = dz; return 0; });
}
void orc_step_conv(void* p, int it, int potmode) { dispatch((Handle*)p, [&](auto& s) { s.step_conv(it, potmode); return 0; }); }
void orc_step_fd(void* p, int it) { dispatch((Handle*)p, [&](auto& s) { s.step_fd(it); return 0; }); }
// compat=physical (1) or the reference's behaviour (0); before set_calc
void orc_set_physical(void* p, int on) {
  dispatch((Handle*)p, [&](auto& s) { s.physical = on != 0; return 0; });
}
// a fixed real potential for step_fd (NULL restores the atoms' potential)
void orc_set_fd_potential(void* p, const double* V) {
  dispatch((Handle*)p, [&](auto& s) {
    if (!V) s.fd_potential_override.clear();
    else s.fd_potential_override.assign(V, V + s.npix());
    return 0;
  });
}
void orc_step_rs(void* p, int it) { dispatch((Handle*)p, [&](auto& s) { s.step_rs(it); return 0; }); }
// the real-space second-derivative stencil of MultisliceStepRS on a wave
void orc_rs_laplacian(void* p, const double* in, double* out) {
  dispatch((Handle*)p, [&](auto& s) {
    using Cx = typename std::decay_t<decltype(s)>::C;
    std::vector<Cx> w, o;
    to_complex(in, s.npix(), w);
    s.rs_laplacian(w, o);
    from_complex(o, out);
    return 0;
  });
}

void orc_probe(void* p, int px, int py, double* out) {
  dispatch((Handle*)p, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    std::vector<std::complex<R>> w;
    s.probe(px, py, w);
    from_complex(w, out);
    return 0;
  });
}

// which: 0 Psi, 1 PsiMinus, 2 PsiPlus
void orc_get_wave(void* p, int which, double* out) {
  dispatch((Handle*)p, [&](auto& s) { from_complex(which == 0 ? s.Psi : which == 1 ? s.PsiMinus : s.PsiPlus, out); return 0; });
}
void orc_set_wave(void* p, int which, const double* in) {
  dispatch((Handle*)p, [&](auto& s) { to_complex(in, s.npix(), which == 0 ? s.Psi : which == 1 ? s.PsiMinus : s.PsiPlus); return 0; });
}

void orc_add_tds(void* p) { dispatch((Handle*)p, [&](auto& s) { s.add_tds(); return 0; }); }
void orc_get_tds(void* p, double* out) {
  dispatch((Handle*)p, [&](auto& s) {
    for (size_t i = 0; i < s.TDS.size(); ++i) out[i] = (double)s.TDS[i];
    return 0;
  });
}

double orc_adf(void* p, float inner, float outer) {
  return dispatch((Handle*)p, [&](auto& s) { return (double)s.adf(s.Psi, inner, outer); });
}

void orc_ctf_image(void* p, double df, double* out) {
  dispatch((Handle*)p, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    std::vector<R> I;
    s.ctf_image(I, s.Psi, df);
    for (size_t i = 0; i < I.size(); ++i) out[i] = (double)I[i];
    return 0;
  });
}

int orc_fd_batch(void* p, int nprobe, const int* posx, const int* posy, int nsteps, float adf_in, float adf_out,
                 double* adf, int nthreads) {
  return dispatch((Handle*)p, [&](auto& s) {
    return fd_batch(s, nprobe, posx, posy, nsteps, adf_in, adf_out, adf, nthreads);
  });
}
int orc_stem_batch(void* p, int nprobe, const int* posx, const int* posy, float dz, int nslices, int potmode,
                   float adf_in, float adf_out, double* adf, double* exitw, double* cbed, int nthreads) {
  return dispatch((Handle*)p, [&](auto& s) {
    return stem_batch(s, nprobe, posx, posy, dz, nslices, potmode, adf_in, adf_out, adf, exitw, cbed, nthreads);
  });
}

int orc_diff(void* p, int nslices, const float* V, int nprobe, const double* psi0, double* intensity,
             const double* dLdI, double* grad, float dz) {
  return dispatch((Handle*)p, [&](auto& s) { return diff_run(s, nslices, V, nprobe, psi0, intensity, dLdI, grad, dz); });
}

// ------------------------------------------------------------ elementary
double orc_bessk0(double x) { return bessk0<double>(x); }
double orc_bessi0(double x) { return bessi0<double>(x); }
double orc_vatom(int Z, double r2, const float* fp) { return vatom<double>(Z, r2, fp); }
double orc_vzatom(int Z, double r2, const float* fp) { return vzatom<double>(Z, r2, fp); }

void orc_fft2(int ny, int nx, double* data, int inverse) {
  FFT2D<double> f;
  f.init(nx, ny);
  std::vector<std::complex<double>> w;
  to_complex(data, (size_t)nx * ny, w);
  f.run(w.data(), inverse != 0);
  from_complex(w, data);
}

void orc_philox4x64(const uint64_t* ctr, const uint64_t* key, uint64_t* out) { philox4x64(ctr, key, out); }

// detector model of SimulateImage (U.cu:2345-2428) on an intensity image
void orc_detector_image(int nx, int ny, const double* I, float dose, uint64_t seed, const float* dqe,
                        const float* ntf, int binning, double* out, double* stats) {
  std::vector<double> in(I, I + (size_t)nx * ny);
  std::vector<std::complex<double>> o;
  detector_image<double>(nx, ny, in, dose, seed, dqe, ntf, binning, o, stats);
  from_complex(o, out);
}
float orc_detector_table(const float* tab, int x, int y, int nx, int ny, int binning) {
  return detector_table(tab, x, y, nx, ny, binning);
}
double orc_detector_normal(uint64_t seed, uint64_t i) { return detector_normal(seed, i); }
float orc_readout_noise(uint64_t seed, uint64_t counter) { return readout_noise(seed, counter); }

// Frozen-phonon displacements for one configuration, U.cu:1549-1551 with the
// Philox/Box-Muller generator of msl_oracle.hpp.
void orc_tds_displacements(uint64_t seed, int config, int n, float sigma, float* dx, float* dy, float* dz) {
  for (int i = 0; i < n; ++i) {
    dx[i] = (float)((double)sigma * philox_normal(seed, (uint64_t)config, (uint64_t)i, 0));
    dy[i] = (float)((double)sigma * philox_normal(seed, (uint64_t)config, (uint64_t)i, 1));
    dz[i] = (float)((double)sigma * philox_normal(seed, (uint64_t)config, (uint64_t)i, 2));
  }
}

int orc_ref_znum(const char* sym) { return ref_znum(sym); }

}  // extern "C"
