/* msl.h -- C-ABI of the MI355X multislice engine (libmsl.so).
 *
 * Drop-in boundary for the reference's native engine class
 * UnmanagedMultisliceSimulation ("GUI Simulation/unmanagedGUISimulation.cuh",
 * U.cuh:62-199; methods defined in "GUI Simulation/UnmanagedGUISimulation.cu",
 * U.cu).  Each entry point below names the reference method it replaces.
 * The C++ facade in UnmanagedMultisliceSimulation.h keeps the reference's
 * method names and signatures on top of these calls.
 *
 * Conventions: plain pointers and sizes; every call returns an msl_status and
 * records a message readable with msl_last_error().  Results are copied into
 * caller-owned buffers.  A handle owns its device memory and one HIP stream;
 * a handle is not thread-safe, distinct handles are independent.  Units as
 * in the reference: angstrom, kV, mrad (apertures, beta), radians (angles).
 * Waves are complex64 stored as interleaved (re, im) float pairs, row-major
 * [resY][resX].  Grid sizes must be powers of two from 16 to 2048.
 */
#ifndef MSL_H
#define MSL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct msl_sim msl_sim;
typedef int msl_status;

enum {
  MSL_OK = 0,
  MSL_ERR_ARG = 1,         /* invalid argument */
  MSL_ERR_STATE = 2,       /* call out of sequence (e.g. step before init) */
  MSL_ERR_IO = 3,          /* file missing or malformed */
  MSL_ERR_HIP = 4,         /* HIP runtime error */
  MSL_ERR_UNSUPPORTED = 5, /* size or mode not supported */
  MSL_ERR_OOM = 6,         /* device allocation failed */
  MSL_ERR_RANGE = 7,       /* atom outside the bin grid / no stable FD step */
  MSL_ERR_CANCELLED = 8    /* msl_cancel() was called */
};

/* compatibility flags (msl_set_compat) */
enum {
  MSL_COMPAT_REFERENCE_SYMBOLS = 1, /* XYZ symbols outside GetZNum's table map to H (U.cu:1188-1243) */
  /* compat=physical: reference quirks fixed instead of reproduced (default off):
   *  - conventional slice i integrates [H - i dz, H - (i-1) dz), the top slab
   *    included (the reference starts one slice lower, U.cu:2097, 636);
   *  - the integrated potential reads every slab within sqrt(5) A of the
   *    slice's samples (the reference's +-3 one-angstrom slabs, U.cu:602-605,
   *    truncate slices thicker than ~0.76 A);
   *  - band limit (propagator, FD band limit) at 2/3 of the Nyquist frequency
   *    (U.cu:1929-1953 takes about half of it);
   *  - STEM probes normalised to sum |psi|^2 = 1 (U.cu:2698-2705 divides by
   *    sum |psi|);
   *  - AddTDSWaves accumulates |FFT psi|^2 / (nx ny) (U.cu:2741-2755
   *    accumulates |FFT psi|).
   * The band limit follows the flag at the next PreCalculateFrequencies, the
   * rest at the next potential, probe or TDS call.  Single-handle and batched
   * STEM paths; the serving and differentiable engines keep the reference's
   * behaviour. */
  MSL_COMPAT_PHYSICAL = 2
};

/* potential modes (msl_set_potential_mode) */
enum {
  MSL_POT_INTEGRATED = 0, /* 20-point z integral of the 3D potential (binnedAtomPKernel3, U.cu:582-653) */
  MSL_POT_PROJECTED = 1   /* Bessel/Gaussian projected potential (binnedAtomPKernelConventional3, U.cu:730-795) */
};

/* ---- lifetime: constructor / destructor, U.cu:2551-2561 ---- */
msl_status msl_create(msl_sim** out, int device);
void msl_destroy(msl_sim* sim);
const char* msl_last_error(const msl_sim* sim);
const char* msl_status_string(msl_status st);
msl_status msl_set_compat(msl_sim* sim, unsigned flags);
msl_status msl_set_potential_mode(msl_sim* sim, int mode);
msl_status msl_cancel(msl_sim* sim); /* polled between slices by batched runs */

/* ---- structure: LoadAtomFile, U.cu:1293-1356 ---- */
msl_status msl_load_atom_file(msl_sim* sim, const char* path);
/* atoms given directly (Z in 1..103); extents follow LoadAtomFile's rule */
msl_status msl_set_atoms(msl_sim* sim, int n, const int* Z, const float* x, const float* y, const float* z);
/* UploadBinnedAtoms, U.cu:1358-1536 */
msl_status msl_upload_binned_atoms(msl_sim* sim);
/* UploadBinnedAtomsTDS, U.cu:1538-1716: each call bins the next frozen-phonon
 * configuration, displaced by seeded Philox4x64-10 Gaussians (sigma per axis)
 * in place of the reference's unseeded rand() */
msl_status msl_set_tds(msl_sim* sim, uint64_t seed, float sigma);
msl_status msl_upload_binned_atoms_tds(msl_sim* sim);
/* bin with caller-given displacements (any may be NULL) */
msl_status msl_upload_binned_atoms_displaced(msl_sim* sim, const float* dx, const float* dy, const float* dz);
/* the displacements msl_upload_binned_atoms_tds uses for `config` */
msl_status msl_tds_displacements(uint64_t seed, int config, int n, float sigma, float* dx, float* dy, float* dz);
/* GetParameterisation, U.cu:1718-1747: path NULL or "" selects the embedded table */
msl_status msl_get_parameterisation(msl_sim* sim, const char* path);
/* copy the 103x12 table in use */
msl_status msl_get_fparams(const msl_sim* sim, float* out);

/* ---- microscope and sampling ---- */
/* ApplyMicroscopeParameters, U.cu:1749-1761 */
msl_status msl_apply_microscope_parameters(msl_sim* sim, float kV, float df, float Mod2f, float Arg2f, float Cs,
                                           float beta, float delta, float aperture);
/* SetCalculationVariables, U.cu:1763-1875 */
msl_status msl_set_calculation_variables(msl_sim* sim, float ps, float df, int resX, int resY, float sizeX,
                                         float sizeY, float sizeZ, float offx, float offy);
/* STEMSetCalculationVariables, U.cu:2567-2670 */
msl_status msl_stem_set_calculation_variables(msl_sim* sim, float ps, float df, int resX, int resY, float sizeX,
                                              float sizeY, float sizeZ);
/* PreCalculateFrequencies, U.cu:1877-1964 */
msl_status msl_precalculate_frequencies(msl_sim* sim);

/* ---- waves ---- */
/* InitialiseWavefunctions(), U.cu:1966-2001 (finite-difference start) */
msl_status msl_initialise_wavefunctions(msl_sim* sim);
/* InitialiseWavefunctions(float dz), U.cu:2003-2038 (conventional) */
msl_status msl_initialise_wavefunctions_dz(msl_sim* sim, float dz);
/* STEMInitialiseWavefunctions, U.cu:2672-2724 */
msl_status msl_stem_initialise_wavefunctions(msl_sim* sim, int posx, int posy);
/* conventional slice thickness used by STEM runs (the reference leaves it
 * at its last TEM value, U.cu:2005) */
msl_status msl_set_conventional_dz(msl_sim* sim, float dz);
/* MultisliceStep (finite difference), U.cu:2040-2083 */
msl_status msl_multislice_step(msl_sim* sim, int iteration);
/* MultisliceStepConv, U.cu:2085-2124 */
msl_status msl_multislice_step_conv(msl_sim* sim, int iteration);
/* MultisliceStepRS, U.cu:2126-2178 (in the C++ API, not used by the GUI):
 * real potential of the 1 A slab below z = H - iteration (10 samples), then
 * the 10-term real-space series 1 + i p1 V, w += i p_m (L w + V w),
 * p_m = lambda / (4 pi m), L the reference's 5-point second differences with
 * out-of-grid neighbours reading element 0.  As in the reference the series
 * restarts from 1 every step: Psi = PsiPlus = w10, and from the second step
 * PsiMinus = w10 * (previous Psi). */
msl_status msl_multislice_step_rs(msl_sim* sim, int iteration);
/* slice count as ManagedMultislice::GetSlices computes it (MM.cpp:105-111) */
msl_status msl_get_slices(const msl_sim* sim, int conventional, int* out);

/* ---- results ---- */
/* GetExitWave, U.cu:2231-2284: out (2*resX*resY floats) may be NULL; also
 * sets maxposEW / minposEW */
msl_status msl_get_exit_wave(msl_sim* sim, float* out);
/* GetSTEMExitWave, U.cu:2180-2229: fftshifted FFT of the wave */
msl_status msl_get_stem_exit_wave(msl_sim* sim, float* out);
/* diffraction log display of that centred spectrum C (the diffpattern branch
 * of conjugationKernel, U.cu:230-231, the mode U.cu:2198 marks "Should be
 * 1"): out[y][x] = log(1 + 3000 |C[y3][x3] + conj(C[ny-1-y3][nx-1-x3])|^2),
 * (x3, y3) the half-grid shift of (x, y); out is resX*resY floats; also sets
 * maxposEW / minposEW as msl_get_stem_exit_wave */
msl_status msl_get_diffraction_display(msl_sim* sim, float* out);
/* AllocTDS / AddTDSWaves / GetTDS*, U.cu:2726-2791 */
msl_status msl_alloc_tds(msl_sim* sim);
msl_status msl_add_tds_waves(msl_sim* sim);
msl_status msl_get_tds(msl_sim* sim, float* out);
/* SimulateImage (U.cu:2322-2428): the CTF image |IFFT(CTF . FFT(psi))|^2,
 * stored for msl_get_image; with the detector model enabled it continues
 * with the reference's K2 detector: DQE filter (sqrt of the curve), shot
 * noise round(dose a + sqrt(dose a) g) with g a seeded clamped normal
 * (Philox instead of rand()), NTF filter, 8 counts per electron; the complex
 * result is read by msl_get_detector_image and maxposIm / minposIm /
 * imagecontrast describe its modulus */
msl_status msl_simulate_image(msl_sim* sim, float dose);
msl_status msl_get_image(msl_sim* sim, float* out);
/* detector model on / off (off by default: msl_simulate_image gives the CTF
 * image only) and the seed of its noise */
msl_status msl_set_detector(msl_sim* sim, int enabled, uint64_t seed);
/* SimulatedImage of the reference: [resY][resX][2] */
msl_status msl_get_detector_image(msl_sim* sim, float* out);
/* readout noise of GetImValue (U.cu:2497-2512): 4 + 1.33 g, g the clamped
 * normal of draw `counter` of the detector's noise stream */
msl_status msl_readout_noise(const msl_sim* sim, long long counter, float* out);
/* CTF images for several defoci from the current wave: out is ndf*resX*resY */
msl_status msl_ctf_sweep(msl_sim* sim, int ndf, const float* dfs, float* out);
/* the GUI's batch Cs / defocus dialog (BatchProcessing.xaml.cs:29-135, whose
 * run handler the reference leaves empty): CTF images of the current wave for
 * every (Cs, df) of cs_steps x df_steps values start + i (end - start) /
 * (steps - 1) (Cs in A); out (optional) is [cs_steps][df_steps][resY][resX];
 * with folder != NULL each image is also written as
 * <folder>/Cs_<Cs>_df_<df>.tiff (32-bit float) */
msl_status msl_cs_df_sweep(msl_sim* sim, float cs_start, float cs_end, int cs_steps, float df_start, float df_end,
                           int df_steps, float* out, const char* folder);
/* annular dark-field signal of the current wave: sum of |FFT_unitary(psi)|^2
 * over inner <= lambda |k| < outer (mrad) */
msl_status msl_adf(msl_sim* sim, float inner_mrad, float outer_mrad, float* out);

typedef struct {
  float MaximumX, MinimumX, MaximumY, MinimumY, MaximumZ, MinimumZ; /* U.cuh:127-132 */
  int numberOfSlices;                                             /* FD slices, U.cuh:133 */
  float kmax, conventionaldz;                                     /* U.cuh:134-135 */
  float maxposIm, maxposEW, minposEW, minposIm;                   /* U.cuh:137-140 */
  float imagecontrast;                                            /* U.cuh:197 */
  float wavelength, sigma2, dzFD;
  int natoms, nslab, resX, resY;
  int band_tiles;  /* > 0: propagation skips the kx columns outside the propagator band (N/4 <= x < 3N/4) */
} msl_info;
msl_status msl_get_info(const msl_sim* sim, msl_info* out);

/* ---- batched STEM + frozen phonons (no counterpart: the reference runs one
 * probe position, MW.cs:360-399) ---- */
typedef struct {
  int nprobe;
  const int* posx;        /* probe positions in pixels, as STEMInitialiseWavefunctions */
  const int* posy;
  int nconfig;            /* frozen-phonon configurations; 0 or 1 with sigma 0 = static */
  uint64_t seed;          /* Philox seed of the displacements */
  float sigma;            /* displacement per axis (A) */
  float dz;               /* slice thickness (A) */
  int nslices;            /* 0: ceil((MaxZ - MinZ) / dz) */
  float adf_inner, adf_outer; /* mrad */
  int batch;              /* probes per launch, 0: automatic */
  int use_graph;          /* capture one chunk (probe synthesis, slice loop, detector) in a
                             hipGraph per chunk size and replay it for every chunk and
                             configuration; ignored with exit_out or while per-kernel
                             profiling is on */
  float* adf_out;         /* [nprobe], mean over configurations */
  float* pacbed_out;      /* optional [resY][resX]: mean over probes and configurations of
                             |Psi_hat|^2, Psi_hat the unnormalised FFT of the exit wave
                             (unshifted; sums to resX*resY times the mean intensity) */
  float* exit_out;        /* optional [nconfig][nprobe][resY][resX][2] exit waves (small runs) */
} msl_stem_batch_desc;
msl_status msl_stem_batched(msl_sim* sim, const msl_stem_batch_desc* desc);
/* batch width autotuner: runs one chunk of B of the descriptor's probes over
 * nslices slices for B = 64, 128, ... up to nprobe (and ~32 GiB of waves),
 * the faster of two timed runs each, and keeps the widest B within 2 % of the
 * fastest per probe-slice as the width of later batch = 0 runs on this grid
 * size.  On input *ncand is the capacity of widths /
 * us_per_probe_slice (may be 0); on output the number of candidates. */
msl_status msl_autotune_batch(msl_sim* sim, const msl_stem_batch_desc* desc, int nslices, int* best, int* widths,
                              float* us_per_probe_slice, int* ncand);

/* time spent in the last msl_stem_batched, split by stage (milliseconds):
 * [0] binning, [1] potentials, [2] propagation, [3] total.  The frozen-phonon
 * displacements of every configuration after the first are drawn on the host
 * while the previous configuration propagates, inside [2]. */
msl_status msl_last_timings(const msl_sim* sim, float* out4);
/* per-kernel timing of batched runs with HIP events on the handle's stream
 * (enabling resets the counters); out8 = {row ms, row launches, column ms,
 * column launches, detector ms, detector launches, potential ms, potential
 * launches} summed over the runs since enabling */
msl_status msl_set_profiling(msl_sim* sim, int on);
msl_status msl_kernel_stats(const msl_sim* sim, double* out8);

/* ---- differentiable multislice (no counterpart in the reference) ---- */
typedef struct msl_diff msl_diff;
/* psi_i = IFFT(P . FFT(exp(i sigma V_i) psi_{i-1})) / (nx ny), P the Fresnel
 * propagator of U.cu:399-428 for slice thickness dz; V given per slice.
 * checkpoint_every K: the forward keeps every K-th wave and the backward
 * recomputes the others segment by segment; 0 keeps every slice's wave (no
 * recomputation; (3 + nslices) x nprobe x nx x ny x 8 bytes); < 0 picks the
 * longest segment whose waves fit in half the device's free memory.  The
 * forward also keeps every slice's transmission exp(i sigma V_i) (nslices x
 * nx x ny x 8 bytes), which the backward reads instead of V. */
msl_status msl_diff_create(msl_diff** out, int device, int nx, int ny, float ps, float kV, float dz, int nslices,
                           int nprobe, int checkpoint_every);
void msl_diff_destroy(msl_diff* d);
const char* msl_diff_last_error(const msl_diff* d);
/* device pointers; work is queued on `stream` (a hipStream_t; NULL is the
 * null stream, ordered with all other default-stream work such as torch's
 * default stream) and not waited for */
msl_status msl_diff_forward(msl_diff* d, const float* V, const void* psi0, float* intensity, void* stream);
/* the same intensities without keeping anything for a backward (no
 * checkpoints, no stored real-space waves): inference, and the reference
 * point of the backward's cost; msl_diff_backward then returns MSL_ERR_STATE
 * until the next msl_diff_forward */
msl_status msl_diff_forward_only(msl_diff* d, const float* V, const void* psi0, float* intensity, void* stream);
msl_status msl_diff_backward(msl_diff* d, const float* dLdI, float* gradV, void* stream);
msl_status msl_diff_info(const msl_diff* d, float* wavelength, float* sigma, float* kmax, long long* workspace_bytes);
/* probes per adjoint chunk of the backward (0: automatic, all probes in one
 * chunk); every chunk runs through a whole segment before the next, and
 * gradient rows are summed in chunk order */
msl_status msl_diff_set_chunk(msl_diff* d, int probes);

/* ---- pass layouts and autotuning (no counterpart in the reference) ----
 * kind 0: row pass, 1: column pass.  Variants are alternative LDS layouts and
 * synchronisation scopes of the same transform; see the variant notes in
 * k_pass.hip.  msl_bench_pass times `reps` launches on an n x n batch of
 * nbatch waves and, with maxdiff, compares against variant 0;
 * msl_autotune_passes selects the fastest agreeing variant per pass for n
 * (out[16]: [0..7] row and [8..15] column timings in ms per variant, -1 where
 * rejected or absent).  Timings use band-pruned passes where the length allows. */
msl_status msl_set_pass_variant(int kind, int n, int variant);
msl_status msl_get_pass_variant(int kind, int n, int* variant);
msl_status msl_bench_pass(int kind, int variant, int n, int nbatch, int reps, float* ms, float* maxdiff);
msl_status msl_autotune_passes(int n, int nbatch, int reps, int* row_var, int* col_var, float* out);
/* plain streaming device copy (bytes a multiple of 16); the reference point
 * for hardware traffic counters */
msl_status msl_copy_device(void* dst, const void* src, long long bytes, void* stream);

/* ---- on-disk output (the GUI's TIFF exports, MW.cs:684-890) ----
 * 32-bit float TIFF: one sample per pixel, IEEE float, one strip, top-left
 * origin, min-is-black, uncompressed, little-endian.  msl_write_tiff writes a
 * caller's height x width image (host only); msl_save_tiff writes one of the
 * GUI's exports of the handle's state: the simulated image as GetImValue
 * reads it (detector image + readout noise, or the CTF image with the
 * detector model off), the exit wave's real or imaginary part, or the CBED
 * accumulator. */
enum { MSL_TIFF_IMAGE = 0, MSL_TIFF_EW_RE = 1, MSL_TIFF_EW_IM = 2, MSL_TIFF_TDS = 3 };
msl_status msl_write_tiff(const char* path, int width, int height, const float* data);
msl_status msl_save_tiff(msl_sim* sim, int what, const char* path);

/* ---- serving: a mix of independent jobs of uneven size (no counterpart in
 * the reference, whose GUI runs one simulation at a time, MW.cs:231-403) ----
 * Each job is the reference's call sequence on its own atoms: extents and
 * binning (LoadAtomFile / UploadBinnedAtoms), SetCalculationVariables with a
 * zero block offset (plane wave) or STEMSetCalculationVariables (probes),
 * PreCalculateFrequencies, then nslices MultisliceStepConv (slice thickness
 * dz) or nslices finite-difference MultisliceStep, from InitialiseWavefunctions
 * / STEMInitialiseWavefunctions.  Jobs sharing a grid size (and, for finite
 * differences, the plane-wave / probe step size) share one batch: every
 * slice is one row and one column launch over all their waves, the waves of
 * jobs with fewer slices leave the batch when they finish.  Batches are cut
 * to a device-memory budget. */
enum { MSL_METHOD_CONVENTIONAL = 0, MSL_METHOD_FD = 1 };
typedef struct {
  int n;                  /* grid n x n, power of two 16..2048 */
  int method;             /* MSL_METHOD_* */
  int nslices;            /* conventional slices, or finite-difference steps */
  int natoms;
  const int* Z;           /* atomic numbers 1..103 */
  const float* x;         /* positions (A) */
  const float* y;
  const float* z;
  int nprobe;             /* 0: plane wave; else probes at posx/posy (pixels) */
  const int* posx;
  const int* posy;
  float* exit_out;        /* plane wave: [n][n][2] exit wave (optional) */
  float* adf_out;         /* probes: [nprobe] ADF, sum |Psi_unitary|^2 in the annulus */
} msl_job;
typedef struct {
  float kV;               /* all jobs */
  float ps;               /* pixel size (A), all jobs */
  float dz;               /* conventional slice thickness (A) */
  float aperture;         /* probe convergence semi-angle (mrad); Cs = df = 0 */
  float adf_inner, adf_outer; /* mrad */
  int potential_mode;     /* MSL_POT_* */
  long long budget_bytes; /* device memory for the waves of one batch (0: 32 GiB) */
  int group_sizes;        /* 1: conventional jobs of grid 256..2048 share batches, with one
                             potential, one row and one column launch per slice for all
                             sizes; 0: every size class runs its own launches */
} msl_serve_params;
typedef struct msl_server msl_server;
msl_status msl_server_create(msl_server** out, int device);
void msl_server_destroy(msl_server* srv);
const char* msl_server_last_error(const msl_server* srv);
/* runs every job; stats (optional, 8 doubles): total ms, set-up ms (handles,
 * binning, tables), propagation ms, batches, waves, wave-slices, launches of
 * the row/column passes, largest batch (bytes) */
msl_status msl_server_run(msl_server* srv, const msl_serve_params* prm, const msl_job* jobs, int njobs,
                          double* stats);

/* ---- device-array primitives (used by the torch binding and tests) ---- */
/* batched 2D FFT of nbatch [ny][nx] complex64 device arrays, in place,
 * unnormalised; inverse != 0 for exp(+2 pi i) */
msl_status msl_fft2_device(int nx, int ny, int nbatch, void* data, int inverse, void* stream);
/* msl_fft2_device without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  No counterpart in the reference.  The twiddle tables that
 * msl_fft2_device allocates per call live in one static 32 KiB device array
 * per GPU instead: nx, ny <= 2048 (MSL_ERR_UNSUPPORTED beyond), nbatch <=
 * 65535.  Work is queued on `stream` and not waited for.  Every call on a GPU
 * rewrites that array: calls that overlap on one GPU race and must be
 * serialised by the caller. */
msl_status msl_fft2_device_noalloc(int nx, int ny, int nbatch, void* data, int inverse, void* stream);
/* pixel tile of the potential kernel: 0 16x16, 1 32x8, 2 64x4,
 * 3 8x32 threads (default); all give the same bits.  msl_autotune_potential times
 * each on the handle's binned atoms and grid (nslices slices of thickness
 * dz from the top) and keeps the fastest whose potentials equal the 16x16
 * tile's; ms_per_variant (optional, 4 floats) gets the times (-1 where
 * rejected). */
msl_status msl_set_potential_tile(msl_sim* sim, int variant);
msl_status msl_autotune_potential(msl_sim* sim, float dz, int nslices, int* best, float* ms_per_variant);
/* bin / potential on device arrays held by a handle (debug access) */
msl_status msl_get_binned(const msl_sim* sim, int* start, float* x, float* y, float* z, int* Z);
msl_status msl_get_bins(const msl_sim* sim, int* keys);
/* transmission (mode MSL_POT_*) or, with real != 0, the summed potential of
 * conventional slice `iteration` (z = MaxZ - MinZ - dz*iteration) */
msl_status msl_potential_slice(msl_sim* sim, int iteration, float dz, int real, float* out);
msl_status msl_set_wave(msl_sim* sim, int which, const float* in); /* 0 Psi, 1 PsiMinus, 2 PsiPlus */
msl_status msl_get_wave(msl_sim* sim, int which, float* out);

#ifdef __cplusplus
}
#endif
#endif /* MSL_H */
