"""Python mirror of the reference's simulation object over libmsl.

`MultisliceSimulation` keeps the method names of UnmanagedMultisliceSimulation
(U.cuh:154-195) and raises MslError on any failing status instead of the
reference's silent failures; results come back as numpy arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import MslError, lib, msl_info, msl_stem_batch_desc


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class MultisliceSimulation:
    COMPAT_REFERENCE_SYMBOLS = 1
    COMPAT_PHYSICAL = 2

    def __init__(self, device: int = 0, compat_symbols: bool = True, physical: bool = False):
        """physical: compat=physical, the reference quirks of slice position,
        slab window, band limit, probe norm and CBED fixed (msl.h)"""
        h = C.c_void_p()
        st = lib().msl_create(C.byref(h), device)
        if st != 0:
            raise MslError(st, "msl_create failed (no HIP device?)")
        self._h = h
        flags = (self.COMPAT_REFERENCE_SYMBOLS if compat_symbols else 0) | (self.COMPAT_PHYSICAL if physical else 0)
        if flags != self.COMPAT_REFERENCE_SYMBOLS:
            self._c(lib().msl_set_compat(self._h, flags))

    def close(self):
        if getattr(self, "_h", None):
            lib().msl_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, st):
        if st != 0:
            raise MslError(st, lib().msl_last_error(self._h).decode())

    # ------------------------------------------------------------ structure
    def LoadAtomFile(self, path: str):
        self._c(lib().msl_load_atom_file(self._h, path.encode()))

    def set_atoms(self, Z, x, y, z):
        Z = np.ascontiguousarray(Z, np.int32)
        x, y, z = (np.ascontiguousarray(v, np.float32) for v in (x, y, z))
        self._c(lib().msl_set_atoms(self._h, Z.size, _p(Z), _p(x), _p(y), _p(z)))

    def UploadBinnedAtoms(self):
        self._c(lib().msl_upload_binned_atoms(self._h))

    def set_tds(self, seed: int, sigma: float):
        self._c(lib().msl_set_tds(self._h, seed, sigma))

    def UploadBinnedAtomsTDS(self):
        self._c(lib().msl_upload_binned_atoms_tds(self._h))

    def upload_binned_atoms_displaced(self, dx, dy, dz):
        d = [np.ascontiguousarray(v, np.float32) for v in (dx, dy, dz)]
        self._c(lib().msl_upload_binned_atoms_displaced(self._h, *(_p(v) for v in d)))

    def GetParameterisation(self, path: str | None = None):
        self._c(lib().msl_get_parameterisation(self._h, path.encode() if path else None))

    def fparams(self):
        o = np.zeros(103 * 12, np.float32)
        self._c(lib().msl_get_fparams(self._h, _p(o)))
        return o.reshape(103, 12)

    def set_potential_mode(self, mode: int):
        self._c(lib().msl_set_potential_mode(self._h, mode))

    # ------------------------------------------------------------ setup
    def ApplyMicroscopeParameters(self, kV, df=0.0, Mod2f=0.0, Arg2f=0.0, Cs=0.0, B=0.0, D=0.0, obj=0.0):
        self._c(lib().msl_apply_microscope_parameters(self._h, kV, df, Mod2f, Arg2f, Cs, B, D, obj))

    def SetCalculationVariables(self, ps, df, resX, resY, sizeX, sizeY, sizeZ, offx=0.0, offy=0.0):
        self._c(lib().msl_set_calculation_variables(self._h, ps, df, resX, resY, sizeX, sizeY, sizeZ, offx, offy))
        self.resX, self.resY = resX, resY

    def STEMSetCalculationVariables(self, ps, df, resX, resY, sizeX, sizeY, sizeZ):
        self._c(lib().msl_stem_set_calculation_variables(self._h, ps, df, resX, resY, sizeX, sizeY, sizeZ))
        self.resX, self.resY = resX, resY

    def PreCalculateFrequencies(self):
        self._c(lib().msl_precalculate_frequencies(self._h))

    def InitialiseWavefunctions(self, dz: float | None = None):
        if dz is None:
            self._c(lib().msl_initialise_wavefunctions(self._h))
        else:
            self._c(lib().msl_initialise_wavefunctions_dz(self._h, dz))

    def STEMInitialiseWavefunctions(self, posx: int, posy: int):
        self._c(lib().msl_stem_initialise_wavefunctions(self._h, posx, posy))

    def set_conventional_dz(self, dz: float):
        self._c(lib().msl_set_conventional_dz(self._h, dz))

    def MultisliceStep(self, it: int):
        self._c(lib().msl_multislice_step(self._h, it))

    def MultisliceStepRS(self, it: int):
        self._c(lib().msl_multislice_step_rs(self._h, it))

    def MultisliceStepConv(self, it: int):
        self._c(lib().msl_multislice_step_conv(self._h, it))

    def GetSlices(self, conventional: bool = True) -> int:
        n = C.c_int()
        self._c(lib().msl_get_slices(self._h, 1 if conventional else 0, C.byref(n)))
        return n.value

    # ------------------------------------------------------------ results
    def _cplx(self):
        return np.zeros((self.resY, self.resX), np.complex64)

    def GetExitWave(self):
        o = self._cplx()
        self._c(lib().msl_get_exit_wave(self._h, _p(o)))
        return o

    def GetSTEMExitWave(self):
        o = self._cplx()
        self._c(lib().msl_get_stem_exit_wave(self._h, _p(o)))
        return o

    def diffraction_display(self):
        """log(1 + 3000 |.|^2) display of GetSTEMExitWave's centred spectrum
        (U.cu:230-231)"""
        o = np.zeros((self.resY, self.resX), np.float32)
        self._c(lib().msl_get_diffraction_display(self._h, _p(o)))
        return o

    def AllocTDS(self):
        self._c(lib().msl_alloc_tds(self._h))

    def AddTDSWaves(self):
        self._c(lib().msl_add_tds_waves(self._h))

    def GetTDS(self):
        o = np.zeros((self.resY, self.resX), np.float32)
        self._c(lib().msl_get_tds(self._h, _p(o)))
        return o

    def SimulateImage(self, dose: float = 1.0):
        """Runs SimulateImage; returns the CTF intensity (the detector model's
        output, when enabled, is read by detector_image())."""
        self._c(lib().msl_simulate_image(self._h, dose))
        o = np.zeros((self.resY, self.resX), np.float32)
        self._c(lib().msl_get_image(self._h, _p(o)))
        return o

    def set_detector(self, enabled: bool = True, seed: int = 0x4D534C):
        """Stochastic K2 detector model in SimulateImage (U.cu:2345-2428)."""
        self._c(lib().msl_set_detector(self._h, 1 if enabled else 0, seed))

    def detector_image(self):
        o = self._cplx()
        self._c(lib().msl_get_detector_image(self._h, _p(o)))
        return o

    def readout_noise(self, counter: int) -> float:
        v = C.c_float()
        self._c(lib().msl_readout_noise(self._h, counter, C.byref(v)))
        return v.value

    TIFF_IMAGE, TIFF_EW_RE, TIFF_EW_IM, TIFF_TDS = 0, 1, 2, 3

    def save_tiff(self, what: int, path: str):
        """One of the GUI's 32-bit float TIFF exports (MW.cs:684-890)."""
        self._c(lib().msl_save_tiff(self._h, what, path.encode()))

    def cs_df_sweep(self, cs=(0.0, 0.0, 1), df=(0.0, 0.0, 1), folder: str | None = None):
        """CTF images over a Cs x defocus grid ((start, end, steps) each); with
        folder, also one 32-bit TIFF per image."""
        o = np.zeros((cs[2], df[2], self.resY, self.resX), np.float32)
        self._c(lib().msl_cs_df_sweep(self._h, cs[0], cs[1], cs[2], df[0], df[1], df[2], _p(o),
                                      folder.encode() if folder else None))
        return o

    def ctf_sweep(self, dfs):
        dfs = np.ascontiguousarray(dfs, np.float32)
        o = np.zeros((dfs.size, self.resY, self.resX), np.float32)
        self._c(lib().msl_ctf_sweep(self._h, dfs.size, _p(dfs), _p(o)))
        return o

    def adf(self, inner, outer) -> float:
        v = C.c_float()
        self._c(lib().msl_adf(self._h, inner, outer, C.byref(v)))
        return v.value

    def info(self) -> dict:
        i = msl_info()
        self._c(lib().msl_get_info(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in msl_info._fields_}

    # ------------------------------------------------------------ debug access
    def binned(self):
        inf = self.info()
        n, nb = inf["natoms"], inf["nslab"] * 2500 + 1
        st = np.zeros(nb, np.int32)
        x, y, z = (np.zeros(n, np.float32) for _ in range(3))
        Z = np.zeros(n, np.int32)
        self._c(lib().msl_get_binned(self._h, _p(st), _p(x), _p(y), _p(z), _p(Z)))
        return st, x, y, z, Z

    def bin_keys(self):
        k = np.zeros(self.info()["natoms"], np.int32)
        self._c(lib().msl_get_bins(self._h, _p(k)))
        return k

    def potential_slice(self, it: int, dz: float, real: bool = False):
        if real:
            o = np.zeros((self.resY, self.resX), np.float32)
        else:
            o = np.zeros((self.resY, self.resX), np.complex64)
        self._c(lib().msl_potential_slice(self._h, it, dz, 1 if real else 0, _p(o)))
        return o

    def autotune_potential(self, dz: float, nslices: int = 4):
        """times the potential kernel's pixel-tile shapes on this handle's
        atoms and grid, keeps the fastest; returns (best, ms per shape)"""
        best = C.c_int(0)
        ms = np.zeros(4, np.float32)
        self._c(lib().msl_autotune_potential(self._h, dz, int(nslices), C.byref(best), _p(ms)))
        return best.value, ms.tolist()

    def set_potential_tile(self, variant: int):
        self._c(lib().msl_set_potential_tile(self._h, int(variant)))

    def set_wave(self, w, which=0):
        w = np.ascontiguousarray(w, np.complex64)
        self._c(lib().msl_set_wave(self._h, which, _p(w)))

    def get_wave(self, which=0):
        o = self._cplx()
        self._c(lib().msl_get_wave(self._h, which, _p(o)))
        return o

    # ------------------------------------------------------------ batched STEM
    def stem_batched(self, posx, posy, dz, nconfig=1, seed=0x4D534C, sigma=0.0, nslices=0,
                     adf=(60.0, 200.0), batch=0, use_graph=False, pacbed=False, exit_waves=False):
        posx = np.ascontiguousarray(posx, np.int32)
        posy = np.ascontiguousarray(posy, np.int32)
        n = posx.size
        adf_out = np.zeros(n, np.float32)
        pac = np.zeros((self.resY, self.resX), np.float32) if pacbed else None
        ew = np.zeros((max(1, nconfig), n, self.resY, self.resX), np.complex64) if exit_waves else None
        d = msl_stem_batch_desc(n, _p(posx), _p(posy), nconfig, seed, sigma, dz, nslices, adf[0], adf[1], batch,
                                1 if use_graph else 0, _p(adf_out), _p(pac), _p(ew))
        self._c(lib().msl_stem_batched(self._h, C.byref(d)))
        out = {"adf": adf_out}
        if pacbed:
            out["pacbed"] = pac
        if exit_waves:
            out["exit"] = ew
        return out

    def autotune_batch(self, posx, posy, dz, nslices=8, adf=(60.0, 200.0)):
        """Times one chunk of B of the given probes over nslices slices for
        B = 64, 128, ... and keeps, for later stem_batched(batch=0) runs on
        this grid, the widest B within 2 % of the fastest per probe-slice;
        returns (best, {B: us per probe-slice})."""
        posx = np.ascontiguousarray(posx, np.int32)
        posy = np.ascontiguousarray(posy, np.int32)
        adf_out = np.zeros(posx.size, np.float32)
        d = msl_stem_batch_desc(posx.size, _p(posx), _p(posy), 1, 0, 0.0, dz, nslices, adf[0], adf[1], 0, 0,
                                _p(adf_out), None, None)
        best, n = C.c_int(0), C.c_int(32)
        widths = np.zeros(32, np.int32)
        cost = np.zeros(32, np.float32)
        self._c(lib().msl_autotune_batch(self._h, C.byref(d), int(nslices), C.byref(best), _p(widths), _p(cost),
                                         C.byref(n)))
        k = min(n.value, 32)
        return best.value, {int(w): float(c) for w, c in zip(widths[:k], cost[:k])}

    def set_profiling(self, on: bool = True):
        self._c(lib().msl_set_profiling(self._h, 1 if on else 0))

    def kernel_stats(self):
        o = np.zeros(8, np.float64)
        self._c(lib().msl_kernel_stats(self._h, _p(o)))
        names = ("row", "column", "detector", "potential")
        return {n: {"ms_total": float(o[2 * i]), "launches": int(o[2 * i + 1]),
                    "ms_avg": float(o[2 * i] / o[2 * i + 1]) if o[2 * i + 1] else None}
                for i, n in enumerate(names)}

    def last_timings(self):
        o = np.zeros(4, np.float32)
        self._c(lib().msl_last_timings(self._h, _p(o)))
        return dict(zip(("binning_ms", "potential_ms", "propagation_ms", "total_ms"), o.tolist()))


def tds_displacements(seed: int, config: int, n: int, sigma: float):
    dx, dy, dz = (np.zeros(n, np.float32) for _ in range(3))
    st = lib().msl_tds_displacements(seed, config, n, sigma, _p(dx), _p(dy), _p(dz))
    if st != 0:
        raise MslError(st, "msl_tds_displacements")
    return dx, dy, dz


def fft2_device(ptr: int, nx: int, ny: int, nbatch: int, inverse: bool, stream: int = 0):
    st = lib().msl_fft2_device(nx, ny, nbatch, C.c_void_p(ptr), 1 if inverse else 0, C.c_void_p(stream))
    if st != 0:
        raise MslError(st, "msl_fft2_device")


def write_tiff(path: str, image) -> None:
    """32-bit float TIFF of a 2-D array, as the GUI's exports write them."""
    a = np.ascontiguousarray(image, np.float32)
    if a.ndim != 2:
        raise ValueError("a 2-D image is required")
    st = lib().msl_write_tiff(path.encode(), a.shape[1], a.shape[0], _p(a))
    if st != 0:
        raise MslError(st, f"cannot write {path}")
