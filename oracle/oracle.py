"""ctypes wrapper of liboracle.so, the CPU restatement of CUDAMultislice.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py for its cpu_baseline leg.  The engine in
cudamultislice_amd/ never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])


def _lib():
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    vp, i, f, d = C.c_void_p, C.c_int, C.c_float, C.c_double
    sig = {
        "orc_create": ([i], vp), "orc_destroy": ([vp], None), "orc_set_threads": ([i], None),
        "orc_set_atoms": ([vp, i, _ip, _fp, _fp, _fp], i),
        "orc_load_xyz_refmap": ([vp, C.c_char_p], i),
        "orc_get_extents": ([vp, _fp], None),
        "orc_bin": ([vp, vp, vp, vp], i),
        "orc_get_bins": ([vp, _ip, _ip], i),
        "orc_get_csr": ([vp, _ip, _fp, _fp, _fp, _ip], i),
        "orc_set_fparams": ([vp, _fp], None),
        "orc_apply_microscope": ([vp, f, f, f, f, f, f, f, f], None),
        "orc_set_calc": ([vp, f, f, i, i, f, f, f, f, f, i], None),
        "orc_get_scalars": ([vp, _dp], None),
        "orc_get_freqs": ([vp, _fp, _fp], None),
        "orc_potential": ([vp, i, d, d, _dp], None),
        "orc_prop": ([vp, d, _dp], None),
        "orc_init_plane": ([vp, f], None), "orc_init_plane_fd": ([vp], None),
        "orc_init_stem": ([vp, i, i, f], None),
        "orc_step_conv": ([vp, i, i], None), "orc_step_fd": ([vp, i], None), "orc_step_rs": ([vp, i], None),
        "orc_set_fd_potential": ([vp, vp], None),
        "orc_set_physical": ([vp, i], None),
        "orc_rs_laplacian": ([vp, _dp, _dp], None),
        "orc_probe": ([vp, i, i, _dp], None),
        "orc_get_wave": ([vp, i, _dp], None), "orc_set_wave": ([vp, i, _dp], None),
        "orc_add_tds": ([vp], None), "orc_get_tds": ([vp, _dp], None),
        "orc_adf": ([vp, f, f], d),
        "orc_ctf_image": ([vp, d, _dp], None),
        "orc_stem_batch": ([vp, i, _ip, _ip, f, i, i, f, f, vp, vp, vp, i], i),
        "orc_fd_batch": ([vp, i, _ip, _ip, i, f, f, _dp, i], i),
        "orc_diff": ([vp, i, _fp, i, _dp, _dp, vp, vp, f], i),
        "orc_bessk0": ([d], d), "orc_bessi0": ([d], d),
        "orc_vatom": ([i, d, _fp], d), "orc_vzatom": ([i, d, _fp], d),
        "orc_fft2": ([i, i, _dp, i], None),
        "orc_philox4x64": ([_up, _up, _up], None),
        "orc_detector_image": ([i, i, _dp, f, C.c_uint64, _fp, _fp, i, _dp, _dp], None),
        "orc_detector_table": ([_fp, i, i, i, i, i], f),
        "orc_detector_normal": ([C.c_uint64, C.c_uint64], d),
        "orc_readout_noise": ([C.c_uint64, C.c_uint64], f),
        "orc_tds_displacements": ([C.c_uint64, i, i, f, _fp, _fp, _fp], None),
        "orc_ref_znum": ([C.c_char_p], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_L = None


def L():
    global _L
    if _L is None:
        _L = _lib()
    return _L


def _cplx(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.complex128)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n: int) -> None:
    L().orc_set_threads(int(n))


# ---------------------------------------------------------------- elementary
def bessk0(x): return L().orc_bessk0(float(x))
def bessi0(x): return L().orc_bessi0(float(x))
def vatom(Z, r2, fp): return L().orc_vatom(int(Z), float(r2), np.ascontiguousarray(fp, np.float32).ravel())
def vzatom(Z, r2, fp): return L().orc_vzatom(int(Z), float(r2), np.ascontiguousarray(fp, np.float32).ravel())
def ref_znum(sym: str) -> int: return L().orc_ref_znum(sym.encode())


def fft2(a, inverse=False):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    ny, nx = a.shape
    buf = a.view(np.float64).reshape(-1)
    L().orc_fft2(ny, nx, buf, 1 if inverse else 0)
    return buf.view(np.complex128).reshape(ny, nx)


def philox4x64(ctr, key):
    out = np.zeros(4, np.uint64)
    L().orc_philox4x64(np.asarray(ctr, np.uint64), np.asarray(key, np.uint64), out)
    return out


def detector_image(I, dose, seed, dqe, ntf, binning=4):
    """Detector model of SimulateImage (U.cu:2345-2428) on an intensity image:
    (complex image, [|max|, |min|, contrast])."""
    I = np.ascontiguousarray(I, np.float64)
    ny, nx = I.shape
    out = np.zeros((ny, nx), np.complex128)
    st = np.zeros(3)
    L().orc_detector_image(nx, ny, I.ravel(), dose, seed, np.ascontiguousarray(dqe, np.float32),
                           np.ascontiguousarray(ntf, np.float32), binning, out.view(np.float64).ravel(), st)
    return out, st


def detector_table(tab, x, y, nx, ny, binning=4):
    return L().orc_detector_table(np.ascontiguousarray(tab, np.float32), x, y, nx, ny, binning)


def detector_normal(seed, i):
    return L().orc_detector_normal(seed, i)


def readout_noise(seed, counter):
    return L().orc_readout_noise(seed, counter)


def tds_displacements(seed: int, config: int, n: int, sigma: float):
    dx, dy, dz = (np.zeros(n, np.float32) for _ in range(3))
    L().orc_tds_displacements(seed, config, n, sigma, dx, dy, dz)
    return dx, dy, dz


# ---------------------------------------------------------------- simulation
class OracleSim:
    """Mirror of UnmanagedMultisliceSimulation (U.cuh:62-199) on the CPU."""

    def __init__(self, use_float: bool = False, physical: bool = False):
        """physical: compat=physical, the reference quirks Q2, Q3, Q5, Q6 and
        Q7 fixed (top slab integrated, full slab window, 2/3-Nyquist band
        limit, unit-intensity probe, CBED of unitary intensity)"""
        self._h = L().orc_create(1 if use_float else 0)
        self.resX = self.resY = 0
        if physical:
            L().orc_set_physical(self._h, 1)

    def __del__(self):
        if getattr(self, "_h", None):
            L().orc_destroy(self._h)
            self._h = None

    # structure
    def set_atoms(self, Z, x, y, z):
        Z = np.ascontiguousarray(Z, np.int32); n = Z.size
        r = L().orc_set_atoms(self._h, n, Z, *(np.ascontiguousarray(v, np.float32) for v in (x, y, z)))
        assert r == 0
        self.natoms = n

    def load_xyz_refmap(self, path):
        n = L().orc_load_xyz_refmap(self._h, path.encode())
        if n < 0:
            raise ValueError("bad xyz")
        self.natoms = n
        return n

    def extents(self):
        e = np.zeros(6, np.float32); L().orc_get_extents(self._h, e); return e

    def bin(self, disp=None):
        if disp is None:
            r = L().orc_bin(self._h, None, None, None)
        else:
            d = [np.ascontiguousarray(v, np.float32) for v in disp]
            r = L().orc_bin(self._h, *(_ptr(v) for v in d))
        if r < 0:
            raise ValueError("atom outside bin grid")
        self.nslab = r
        return r

    def bins(self):
        b = np.zeros(self.natoms, np.int32); z = np.zeros(self.natoms, np.int32)
        L().orc_get_bins(self._h, b, z); return b, z

    def csr(self):
        st = np.zeros(self.nslab * 2500 + 1, np.int32)
        x, y, z = (np.zeros(self.natoms, np.float32) for _ in range(3))
        Z = np.zeros(self.natoms, np.int32)
        L().orc_get_csr(self._h, st, x, y, z, Z)
        return st, x, y, z, Z

    def set_fparams(self, fp):
        L().orc_set_fparams(self._h, np.ascontiguousarray(fp, np.float32).ravel())

    def apply_microscope(self, kV, df=0.0, A1=0.0, A1phi=0.0, Cs=0.0, beta=0.0, delta=0.0, obj=0.0):
        L().orc_apply_microscope(self._h, kV, df, A1, A1phi, Cs, beta, delta, obj)

    def set_calc(self, ps, df, resX, resY, sx, sy, sz, ox=0.0, oy=0.0, stem=False):
        L().orc_set_calc(self._h, ps, df, resX, resY, sx, sy, sz, ox, oy, 1 if stem else 0)
        self.resX, self.resY = resX, resY

    def scalars(self):
        o = np.zeros(10); L().orc_get_scalars(self._h, o)
        keys = ["wavel", "sigma2", "dzFD", "numberOfSlices", "kmax", "loadblocksx", "loadblocksy",
                "loadslicesz", "wavel_f", "nSlab"]
        return dict(zip(keys, o.tolist()))

    def freqs(self):
        kx = np.zeros(self.resX, np.float32); ky = np.zeros(self.resY, np.float32)
        L().orc_get_freqs(self._h, kx, ky); return kx, ky

    def _wave(self):
        return np.zeros(self.resX * self.resY * 2)

    def potential(self, mode, z, convdz=1.0):
        o = self._wave(); L().orc_potential(self._h, mode, z, convdz, o)
        return o.view(np.complex128).reshape(self.resY, self.resX)

    def prop(self, dz):
        o = self._wave(); L().orc_prop(self._h, dz, o)
        return o.view(np.complex128).reshape(self.resY, self.resX)

    def init_plane(self, dz): L().orc_init_plane(self._h, dz)
    def init_plane_fd(self): L().orc_init_plane_fd(self._h)
    def init_stem(self, px, py, dz=1.0): L().orc_init_stem(self._h, px, py, dz)
    def step_conv(self, it, potmode=0): L().orc_step_conv(self._h, it, potmode)
    def step_fd(self, it): L().orc_step_fd(self._h, it)
    def step_rs(self, it): L().orc_step_rs(self._h, it)

    def set_fd_potential(self, V=None):
        """fixed real potential [resY][resX] for step_fd; None: the atoms'"""
        if V is None:
            L().orc_set_fd_potential(self._h, None)
        else:
            self._fdV = np.ascontiguousarray(V, np.float64)
            L().orc_set_fd_potential(self._h, self._fdV.ctypes.data_as(C.c_void_p))

    def rs_laplacian(self, w):
        w = np.ascontiguousarray(w, np.complex128)
        o = np.zeros_like(w)
        L().orc_rs_laplacian(self._h, w.view(np.float64).ravel(), o.view(np.float64).ravel())
        return o

    def probe(self, px, py):
        o = self._wave(); L().orc_probe(self._h, px, py, o)
        return o.view(np.complex128).reshape(self.resY, self.resX)

    def wave(self, which=0):
        o = self._wave(); L().orc_get_wave(self._h, which, o)
        return o.view(np.complex128).reshape(self.resY, self.resX)

    def set_wave(self, w, which=0):
        L().orc_set_wave(self._h, which, np.ascontiguousarray(w, np.complex128).view(np.float64).ravel())

    def add_tds(self): L().orc_add_tds(self._h)

    def tds(self):
        o = np.zeros(self.resX * self.resY); L().orc_get_tds(self._h, o)
        return o.reshape(self.resY, self.resX)

    def adf(self, inner, outer): return L().orc_adf(self._h, inner, outer)

    def ctf_image(self, df):
        o = np.zeros(self.resX * self.resY); L().orc_ctf_image(self._h, df, o)
        return o.reshape(self.resY, self.resX)

    def stem_batch(self, posx, posy, dz, nslices, adf_in=60.0, adf_out=200.0, potmode=0,
                   want_exit=False, want_cbed=False, nthreads=0):
        posx = np.ascontiguousarray(posx, np.int32); posy = np.ascontiguousarray(posy, np.int32)
        n = posx.size
        adf = np.zeros(n)
        ew = np.zeros(n * self.resX * self.resY * 2) if want_exit else None
        cb = np.zeros(self.resX * self.resY) if want_cbed else None
        L().orc_stem_batch(self._h, n, posx, posy, dz, nslices, potmode, adf_in, adf_out,
                           _ptr(adf), _ptr(ew), _ptr(cb), nthreads)
        out = {"adf": adf}
        if want_exit:
            out["exit"] = ew.view(np.complex128).reshape(n, self.resY, self.resX)
        if want_cbed:
            out["cbed"] = cb.reshape(self.resY, self.resX)
        return out

    def fd_batch(self, posx, posy, nsteps, adf_in=60.0, adf_out=200.0, nthreads=0):
        """ADF of finite-difference STEM runs (STEMInitialiseWavefunctions then
        nsteps MultisliceStep) from every probe position, threads over probes"""
        posx = np.ascontiguousarray(posx, np.int32); posy = np.ascontiguousarray(posy, np.int32)
        adf = np.zeros(posx.size)
        L().orc_fd_batch(self._h, posx.size, posx, posy, int(nsteps), adf_in, adf_out, adf, nthreads)
        return adf

    def diff(self, V, psi0, dLdI=None, dz=2.0):
        V = np.ascontiguousarray(V, np.float32)
        ns = V.shape[0]
        psi0 = np.ascontiguousarray(psi0, np.complex128)
        nprobe = psi0.shape[0]
        inten = np.zeros(psi0.shape, np.float64)
        grad = np.zeros(V.shape, np.float64) if dLdI is not None else None
        d = None if dLdI is None else np.ascontiguousarray(dLdI, np.float64)
        L().orc_diff(self._h, ns, V.ravel(), nprobe, psi0.view(np.float64).ravel(), inten.ravel(),
                     _ptr(d), _ptr(grad), dz)
        return inten, grad


# ---------------------------------------------------------------- serving jobs
def job_reference(Z, x, y, z, n, method, nslices, posx, posy, fp, kV, ps, dz, aperture, adf):
    """One job of a serving mix as the reference runs a simulation: extents
    and binning of its atoms, SetCalculationVariables with a zero block offset
    (plane wave, posx None) or STEMSetCalculationVariables (probes at
    posx/posy), then nslices conventional steps (method 0, U.cu:2085-2124) or
    finite-difference steps (method 1, U.cu:2040-2083).  Returns {"exit":
    wave} or {"adf": per probe}."""
    o = OracleSim()
    o.set_atoms(Z, x, y, z)
    o.set_fparams(fp)
    o.bin()
    e = o.extents()
    stem = posx is not None and np.asarray(posx).size > 0
    o.apply_microscope(kV, obj=aperture if stem else 0.0)
    o.set_calc(ps, 0.0, n, n, 1, 1, float(e[4] - e[5]), 0.0, 0.0, stem)
    if not stem:
        if method == 0:
            o.init_plane(dz)
            for it in range(1, nslices + 1):
                o.step_conv(it)
        else:
            o.init_plane_fd()
            for it in range(1, nslices + 1):
                o.step_fd(it)
        return {"exit": o.wave()}
    if method == 0:
        return {"adf": o.stem_batch(posx, posy, dz, nslices, adf[0], adf[1])["adf"]}
    return {"adf": o.fd_batch(posx, posy, nslices, adf[0], adf[1])}


def serving_sample(jobs):
    """Indices of the jobs of a serving mix checked against the oracle: at
    grids 256, 512 and 1024 the cheapest job (pixel-slices times probes,
    finite differences counted twice) of each (method, plane wave / probes)
    class, and at 2048 the cheapest plane-wave job of each method (14 of the
    256-job mix, 5.5 %)."""
    pick, cost = {}, {}
    for k, j in enumerate(jobs):
        if j.n == 2048 and j.nprobe > 0:
            continue
        key = (j.n, j.method, j.nprobe > 0)
        c = j.n * j.n * j.nslices * max(1, j.nprobe) * (2 if j.method == 1 else 1)
        if key not in pick or c < cost[key]:
            pick[key], cost[key] = k, c
    return sorted(pick.values())
