// msl_internal.hpp -- launch interface between the host orchestration
// (msl_host.cpp) and the HIP kernels (k_pass.hip, k_physics.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace msl {

// ------------------------------------------------------------------ passes
// Row pass: one transform along x (contiguous) per (wave, row) line.
enum RowMode { ROW_FWD = 0, ROW_INV = 1, ROW_INV_MUL_FWD = 2, ROW_INV_FD_FWD = 3 };
// ROW_INV_FD_FWD: the finite-difference step of one row (finiteDifferenceKernel,
// U.cu:995-1023) between the transforms: IFFT_x of the Laplacian spectrum
// (src), psi+ from it, psi- (fd_pm), psi (fd_p0) and V (t, real part, slot
// indexing as RM_T), then FFT_x of psi+ into dst.
// pointwise factor applied between the (optional) inverse and forward
// transforms: none, a stored transmission t, or (RM_SEED, with ROW_FWD) twice
// a real array in the wave's own layout: the gradient 2 dL/dI psi of an
// intensity loss, transformed in the same pass
enum RowMul { RM_NONE = 0, RM_T = 1, RM_SEED = 2 };

// The launch structures are written once as field lists parameterised by a
// pointer qualifier G: empty for the host-visible structs below, the global
// address space for the copies device code reads from tables in memory
// (a pointer loaded from memory is otherwise generic to the compiler, and
// every access through it a flat instruction).
#define MSL_ROW_ARGS(G)                                                                                   \
  const float2 G* src;                                                                                   \
  float2 G* dst;                                                                                         \
  long long src_bstride, dst_bstride; /* elements between consecutive waves */                           \
  int nx, ny, nbatch;                                                                                    \
  /* RM_T: transmission of wave b at t + (b / t_group) * t_bstride, or at t + t_index[b] * t_bstride */ \
  const float2 G* t;                                                                                     \
  long long t_bstride;                                                                                   \
  int t_group;                                                                                           \
  float sigma;                                                                                           \
  float scale;               /* applied with the pointwise factor */                                     \
  int prune_in, prune_out;   /* band-limited input / output (only x < N/4, x >= 3N/4) */                 \
  const float2 G* tw;        /* twiddles for length nx */                                                \
  const int G* t_index;      /* optional per-wave transmission slot (device) */                          \
  /* ROW_INV_MUL_FWD: also store the real-space wave (before the multiply) here, same indexing as dst */ \
  float2 G* aux;                                                                                         \
  /* ROW_INV_FD_FWD: psi- and psi in real space (dst indexing), slice step and wavelength */             \
  const float2 G* fd_pm;                                                                                 \
  const float2 G* fd_p0;                                                                                 \
  float fd_dz, fd_wavel;                                                                                 \
  const float G* seed;       /* RM_SEED: real factor per element, src indexing */
struct RowArgs {
  MSL_ROW_ARGS()
};

// Column pass: one transform along y (stride nx) per (wave, column).
enum ColMode { COL_FWD = 0, COL_INV = 1, COL_FWD_MUL_INV = 2, COL_FWD_MUL = 3, COL_FWD_DETECT = 4, COL_MUL_INV = 5 };
// k-space factor: none (scale only), Fresnel propagator with strict
// |k| < kmax, band limit |k| <= kmax, spectral Laplacian, conjugate propagator
enum ColMul { CM_SCALE = 0, CM_PROP = 1, CM_BAND = 2, CM_GRAD = 3, CM_PROPCONJ = 4, CM_BANDGRAD = 5,
              // CM_PROP / CM_PROPCONJ where rows ny/4 .. 3ny/4 - 1 lie outside the band for every
              // column (ColArgs::zb_rows): the inverse transform skips their zeros.  Chosen by
              // launch_col from zb_rows; callers pass CM_PROP / CM_PROPCONJ.
              CM_PROP_ZB = 6, CM_PROPCONJ_ZB = 7 };
// CM_BANDGRAD: band limit and spectral Laplacian in one factor (the Laplacian
// of a band-limited wave, from the spectrum before the band limit)

#define MSL_COL_ARGS(G)                                                                                   \
  const float2 G* src;                                                                                   \
  float2 G* dst;                                                                                         \
  long long src_bstride, dst_bstride;                                                                    \
  int nx, ny, nbatch;                                                                                    \
  const float G* kx;         /* FP32 spatial frequencies (1/A) */                                        \
  const float G* ky;                                                                                     \
  const float2 G* px;        /* separable propagator factors exp(-i pi lambda dz k^2) */                 \
  const float2 G* py;                                                                                    \
  /* CM_PROP / CM_PROPCONJ: column x is inside |k| < kmax at rows r <= prop_rows[x].x and */            \
  /* r >= prop_rows[x].y (make_prop_rows) */                                                             \
  const int2 G* prop_rows;                                                                               \
  float kmax;                                                                                            \
  float scale;                                                                                           \
  const float2 G* tw;        /* twiddles for length ny */                                                \
  /* COL_FWD_DETECT: annular sum of |v|^2 over kin <= |k| < kout, times det_scale, into */              \
  /* partial[b * (nx / C) + tile]; optional per-wave |v|^2 map (fftshifted) into inten */                \
  float kin, kout, det_scale;                                                                            \
  float G* partial;                                                                                      \
  float G* inten;                                                                                        \
  /* > 0: launch only the tiles_lo column tiles at each end of kx (band-limited data); */                \
  /* COL_FWD_DETECT partials are then indexed by launched tile */                                        \
  int tiles_lo;                                                                                          \
  /* 1: rows ny/4 .. 3ny/4 - 1 are outside |k| < kmax for every column (prop_rows_quarter) */            \
  int zb_rows;
struct ColArgs {
  MSL_COL_ARGS()
};

// rows of each column inside the propagator band |k| < kmax (nx entries,
// device), built on s and waited for; hipErrorInvalidValue if some column's
// band is not a run from each end of ky
hipError_t make_prop_rows(const float* kx, const float* ky, int nx, int ny, float kmax, int2* rows, hipStream_t s);
int col_tile(int ny);  // columns per column-pass workgroup (detector layout)
// true when rows ny/4 .. 3ny/4 - 1 fail the propagator test |k| < kmax for
// every column: the row ny/4 fails it at kx = 0 (the FP32 test of
// k_prop_rows; |ky| only grows towards ny/2 and |k| >= |ky|)
bool prop_rows_quarter(const float* hky, int ny, float kmax);
int band_tiles_lo(int nx);  // column tiles per end covering x < N/4, x >= 3N/4 (0: no pruning)
// pass layout variants per length (kind 0 row, 1 column), see k_pass.hip
int pass_variant(int kind, int n);
void set_pass_variant(int kind, int n, int v);
int num_pass_variants(int kind);
hipError_t launch_row_v(int mode, int mul, const RowArgs& a, hipStream_t s, int var);
hipError_t launch_col_v(int mode, int mul, const ColArgs& a, hipStream_t s, int var);
// grids whose sides are not powers of two (k_generic.hip): unnormalised 2D
// FFT of nbatch [ny][nx] waves in place (scratch of the same size; twx / twy
// the tables exp(-2 pi i k / n)), k-space multipliers (ColMul, one wave
// layout, scale included), elementwise product, and the annular |F|^2 sum
// (partial: gen_annulus_blocks floats) into out[0]
hipError_t launch_gen_fft2(float2* data, float2* scratch, int nx, int ny, int nbatch, int inverse,
                           const float2* twx, const float2* twy, hipStream_t s);
hipError_t launch_gen_kmul(float2* w, int nx, int ny, int nbatch, int mul, const float* kx, const float* ky,
                           const float2* px, const float2* py, float kmax, float scale, hipStream_t s);
hipError_t launch_gen_cmul(const float2* src, const float2* t, float2* dst, long long n, hipStream_t s);
int gen_annulus_blocks(long long np);
hipError_t launch_gen_annulus(const float2* F, int nx, int ny, const float* kx, const float* ky, float kin,
                              float kout, float scale, float* partial, float* out, hipStream_t s);
// deterministic pseudo-random fill: unit != 0 gives exp(i theta) values
hipError_t launch_hash_fill(float2* w, long long n, unsigned seed, int unit, hipStream_t s);
// streaming copy of `bytes` (multiple of 16)
hipError_t launch_copy16(const void* src, void* dst, long long bytes, hipStream_t s);

hipError_t launch_row(int mode, int mul, const RowArgs& a, hipStream_t s);
// One row / column launch over the waves of several grid sizes (one segment
// per length 256, 512, 1024 or 2048; at most kGroupMaxSeg segments, square
// grids).  Row passes: ROW_INV_MUL_FWD with RM_T, band-limited output, the
// same prune flags in every segment.  Column passes: COL_FWD_MUL_INV with
// CM_PROP.  Segments with nbatch 0 are skipped.
// build_*_group validates the segments and lays out their blocks into a
// table; the launch reads the table from device memory (d, a copy of h that
// the caller uploaded) and sizes its persistent grid from h.
constexpr int kGroupMaxSeg = 4;
template <class A> struct PassGroup {
  A a[kGroupMaxSeg];
  long long blk0[kGroupMaxSeg + 1];  // first block of each segment; blk0[nseg] = total
  int nbx[kGroupMaxSeg];             // columns: launched tiles per wave
  int nseg;
};
hipError_t build_row_group(const RowArgs* segs, int nseg, PassGroup<RowArgs>& g);
hipError_t build_col_group(const ColArgs* segs, int nseg, PassGroup<ColArgs>& g);
hipError_t launch_row_group(int mode, int mul, const PassGroup<RowArgs>& h, const PassGroup<RowArgs>* d,
                            hipStream_t s);
hipError_t launch_col_group(int mode, int mul, const PassGroup<ColArgs>& h, const PassGroup<ColArgs>* d,
                            hipStream_t s);
hipError_t launch_col(int mode, int mul, const ColArgs& a, hipStream_t s);
bool fft_size_supported(int n);

// ------------------------------------------------------------------ physics
struct BinArgs {
  const float *x, *y, *z;       // input coordinates (already displaced)
  const int* Z;
  int n;
  float MaxX, MinX, MaxY, MinY, MaxZ, MinZ;
  int nslab;
  int* key;                     // zid*2500 + bid, or -1 if outside
  int* count;                   // nslab*2500 (+1) zero-initialised
  int* start;                   // nslab*2500 + 1
  int* fill;                    // nslab*2500 scratch
  int* perm;                    // n
  float *bx, *by, *bz;          // binned, relative to Minimum*
  int* bZ;
  int* bad;                     // set to 1 if any atom is outside the bin grid
  int* scan_tmp;                // block sums scratch
};
hipError_t launch_bin(const BinArgs& a, hipStream_t s);

// Potential synthesis.  mode 0: 20-point z integration of the 3D potential
// (slabs +-3), 1: projected Bessel/Gaussian potential (slice's own slabs),
// 2: point 3D potential on the plane (finite-difference mode, no offset).
// out_kind 0: transmission exp(i sigma V) as float2, 1: real V as float2
// (V, 0), 2: real V as float.
#define MSL_POT_ARGS(G)                                                                                   \
  const float G* bx;                                                                                     \
  const float G* by;                                                                                     \
  const float G* bz;                                                                                     \
  const int G* bZ;                                                                                       \
  const int G* start;                                                                                    \
  int nslab;                                                                                             \
  const float G* ztab;       /* per-Z transformed Kirkland constants (103 x 12) */                       \
  float MaxX, MinX, MaxY, MinY, MaxZ, MinZ;                                                              \
  int nx, ny;                                                                                            \
  float ps, offx, offy;                                                                                  \
  float dz;                  /* slice thickness (modes 0, 1) */                                          \
  int first_slice, nslices;  /* slice i uses z = (MaxZ-MinZ) - dz * i (modes 0, 1) */                    \
  float zplane;              /* modes 2, 3: plane position (single slice) */                             \
  int loadslicesz;                                                                                       \
  int physical;              /* modes 0, 1: compat=physical slice position and slab window */            \
  int tile;                  /* pixel tile shape (potential_tile_variants(); 0: 16x16) */               \
  float sigma;                                                                                           \
  /* mode 0 3D-integrated slice, 1 projected, 2 FD plane, 3 real-space 1 A slab; */                     \
  /* out_kind 0 exp(i sigma V), 1 (V, 0), 2 V */                                                         \
  int mode, out_kind;                                                                                    \
  void G* out;               /* [nslices][ny][nx] */                                                     \
  long long out_sstride;     /* elements between slices */
struct PotArgs {
  MSL_POT_ARGS()
};
hipError_t launch_potential(const PotArgs& a, hipStream_t s);
// one slice (first_slice = first, nslices 1) of several jobs in one launch:
// tab[j] (device) holds job j's arguments, blk0[0..njobs] (device) the first
// workgroup of each job, job j owning potential_tiles(nx, ny) workgroups
int potential_tiles(int nx, int ny);
int potential_tile_variants();
hipError_t launch_potential_group(const PotArgs* tab, const long long* blk0, int njobs, long long nblocks, int first,
                                  hipStream_t s);
void kirkland_transform(const float* fparams, float* ztab);  // host: 103x12 -> per-Z constants

// Probe in k-space (U.cu:354-383) for nprobe positions: batch b gets
// (posx[b], posy[b]); or posx == nullptr for the single position (px0, py0).
struct ProbeArgs {
  float2* out;
  long long bstride;
  int nx, ny, nprobe;
  const float* kx;
  const float* ky;
  const int* posx;
  const int* posy;
  int px0, py0;
  float apert, ps, df, Cs, wavel;
  // > 0: write only the columns x < band_cols and x >= nx - band_cols (the
  // caller has checked that the aperture lies inside them; the other
  // columns are left as they are)
  int band_cols;
};
hipError_t launch_probe_kspace(const ProbeArgs& a, hipStream_t s);

// sum |w| over n elements per wave (deterministic), then w /= sum.
hipError_t launch_l1_normalise(float2* w, long long n, int nbatch, long long bstride, float* scratch, hipStream_t s);
// as launch_l1_normalise, dividing by sqrt(sum |w|^2) when l2 != 0; the
// divisor of wave b lands in scratch[nbatch * ceil(n / 4096) + b]
hipError_t launch_probe_normalise(float2* w, long long n, int nbatch, long long bstride, float* scratch, int l2,
                                  hipStream_t s);
hipError_t launch_fill(float2* w, long long n, float2 v, hipStream_t s);
hipError_t launch_scale(float2* w, long long n, float s, hipStream_t st);
hipError_t launch_cmul(const float2* a, const float2* b, float2* out, long long n, float s, hipStream_t st);
hipError_t launch_reduce_rows(const float* partial, int nbatch, int ntiles, float* out, int accumulate, hipStream_t s);
hipError_t launch_abs(const float2* w, float* out, long long n, hipStream_t s);

// CTF (imagingKernel2, U.cu:1044-1080) on a k-space wave for ndf defocus
// values: out[d] = E_s E_d exp(-i chi) in[...] inside the aperture.
struct CtfArgs {
  const float2* in;
  float2* out;
  int nx, ny, ndf;
  const float* kx;
  const float* ky;
  const float* dfs;
  const float* css;  // optional per-image Cs (null: Cs for every image)
  float Cs, A1, A1phi, A2, A2phi, obj, wavel, beta, delta, scale;
};
hipError_t launch_ctf(const CtfArgs& a, hipStream_t s);

// one term of the real-space multislice series (MultisliceStepRS, U.cu:2126-2178)
hipError_t launch_rs_term(const float2* w, float2* out, const float2* V, int nx, int ny, float ps, float pre,
                          hipStream_t s);

// stochastic detector model (SimulateImage, U.cu:2345-2428)
hipError_t launch_detector_filter(float2* F, int nx, int ny, const float* tab, int binning, int root, float scale,
                                  hipStream_t s);
hipError_t launch_shot_noise(const float2* z, float2* out, long long n, float zscale, float dose,
                             unsigned long long seed, hipStream_t s);
hipError_t launch_real_to_complex(const float* in, float2* out, long long n, hipStream_t s);
hipError_t launch_intensity(const float2* w, float* out, long long n, float s, hipStream_t st);
// t[i] = exp(i sigma V[i])
hipError_t launch_transmission(const float* V, float2* t, long long n, float sigma, hipStream_t s);

// TDS / CBED: tds[y][x] += |F[(y+ny/2)%ny][(x+nx/2)%nx]| (U.cu:2734-2766)
hipError_t launch_tds_accum(const float2* F, float* tds, int nx, int ny, hipStream_t s);
// intensity != 0: tds += scale |F|^2 (compat=physical) instead of |F|
hipError_t launch_tds_accum_mode(const float2* F, float* tds, int nx, int ny, int intensity, float scale,
                                 hipStream_t s);
// sum over nbatch maps of n floats in fixed order: out (+)= sum_b in[b]
hipError_t launch_sum_maps(const float* in, int nbatch, long long n, float* out, int accumulate, hipStream_t s);

// Finite-difference update (finiteDifferenceKernel, U.cu:995-1023)
// psi+ of n points; V is periodic with period vper (one potential for a
// batch of waves; vper = n: one per wave)
hipError_t launch_fd_update(const float2* grad, const float2* psim, float2* psip, const float2* psi,
                            const float2* V, long long n, long long vper, float sigma, float wavel, float dz,
                            hipStream_t s);

}  // namespace msl
