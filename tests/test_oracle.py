"""Pins the CPU oracle (oracle/) against checks independent of its own code:
scipy special functions, numpy FFT and Philox, closed forms, brute force and
invariants.  CPU only."""
import os

import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as sp

from oracle import oracle as O
from cudamultislice_amd import structures as S

REF_FPARAMS = "/root/reference/fparams.dat"


# ---------------------------------------------------------------- data
def test_embedded_fparams_match_reference_file(fp):
    if not os.path.exists(REF_FPARAMS):
        pytest.skip("reference tree not mounted")
    rows = [ln.split() for ln in open(REF_FPARAMS) if ln.strip()]
    ref = np.array([[np.float32(float(v)) for v in r] for r in rows], np.float32)
    assert ref.shape == (103, 12)
    assert np.array_equal(ref.view(np.uint32), fp.view(np.uint32))


# ---------------------------------------------------------------- Bessel
@pytest.mark.parametrize("x", np.concatenate([np.linspace(0.01, 2.0, 37), np.linspace(2.0, 30.0, 41)]))
def test_bessk0_abramowitz_stegun_bounds(x):
    k = O.bessk0(x)
    if x <= 2.0:  # A&S 9.8.5: |eps| < 1e-8 absolute on K0 + ln(x/2) I0
        assert abs(k - sp.k0(x)) < 1e-7 + 2e-7 * abs(sp.k0(x))
    else:  # A&S 9.8.6: |eps| < 1.9e-7 on sqrt(x) e^x K0
        assert abs(np.sqrt(x) * np.exp(x) * (k - sp.k0(x))) < 2.5e-7


@pytest.mark.parametrize("x", np.concatenate([np.linspace(-3.75, 3.75, 31), np.linspace(3.75, 40, 31)]))
def test_bessi0_abramowitz_stegun_bounds(x):
    v = O.bessi0(x)
    if abs(x) <= 3.75:  # 9.8.1: |eps| < 1.6e-7
        assert abs(v - sp.i0(x)) < 1.6e-7 * max(1.0, abs(sp.i0(x))) + 1e-9
    else:  # 9.8.2: |eps| < 1.9e-7 on sqrt(x) e^-x I0
        assert abs(np.sqrt(x) * np.exp(-x) * (v - sp.i0(x))) < 2.5e-7


def test_bessk0_zero_sentinel():
    assert O.bessk0(0.0) == 1e20


# ---------------------------------------------------------------- Kirkland
def _split_table(fp, Z, part):
    t = fp.copy()
    if part == "lorentz":
        t[Z - 1, 6::2] = 0.0
    else:
        t[Z - 1, 0:6:2] = 0.0
    return t


@pytest.mark.parametrize("Z", [6, 8, 14, 22, 38, 79])
@pytest.mark.parametrize("r", [0.3, 0.6, 1.0, 1.7])
def test_vatom_integrates_to_vzatom(fp, Z, r):
    """Projected potential = z-integral of the 3D potential (Kirkland App. C).
    The Lorentzian terms integrate to 2 K0 exactly; the Gaussian terms carry
    the ratio 266.5157269 / (150.4121417 sqrt(pi)) of the two prefactors the
    reference uses (U.cu:151 vs U.cu:192)."""
    for part, scale in (("lorentz", 1.0), ("gauss", 266.5157269 / (150.4121417 * np.sqrt(np.pi)))):
        t = _split_table(fp, Z, part)
        f = lambda z: O.vatom(Z, r * r + z * z, t)
        integ = 2 * si.quad(f, 0, 40, limit=400, epsabs=1e-12, epsrel=1e-11)[0]
        proj = O.vzatom(Z, r * r, t)
        assert integ == pytest.approx(scale * proj, rel=3e-6)


def test_vatom_clamps_radius(fp):
    assert O.vatom(14, 0.0, fp) == O.vatom(14, 0.0625, fp) == O.vatom(14, 0.01, fp)
    assert O.vzatom(14, 0.0, fp) == O.vzatom(14, 0.0625, fp)
    assert O.vatom(14, 0.07, fp) < O.vatom(14, 0.0625, fp)


def test_vatom_large_r_decays(fp):
    v = [O.vatom(14, r * r, fp) for r in (0.5, 1.0, 2.0, 4.0)]
    assert all(a > b > 0 for a, b in zip(v, v[1:]))


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 8), (16, 32), (64, 64), (6, 10), (12, 9)])
@pytest.mark.parametrize("inv", [False, True])
def test_fft2_matches_numpy(shape, inv):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = O.fft2(a, inv)
    want = np.fft.ifft2(a) * a.size if inv else np.fft.fft2(a)
    assert np.max(np.abs(got - want)) < 1e-10 * np.max(np.abs(want))


# ---------------------------------------------------------------- Philox
def test_philox4x64_matches_numpy():
    for ctr, key in (([0, 0, 0, 0], [0, 0]), ([5, 6, 7, 8], [11, 12]),
                     ([2**64 - 1, 2**63 + 5, 3, 9], [0x4D534C, 77])):
        bg = np.random.Philox(counter=np.array(ctr, np.uint64), key=np.array(key, np.uint64))
        # numpy increments the 256-bit counter before each block
        want = bg.random_raw(4)
        c = np.array(ctr, dtype=np.uint64)
        c1 = c.copy()
        c1[0] = np.uint64((int(c[0]) + 1) % 2**64)
        if int(c[0]) == 2**64 - 1:
            c1[1] = np.uint64((int(c[1]) + 1) % 2**64)
        got = O.philox4x64(c1, np.array(key, np.uint64))
        assert np.array_equal(got, want)


def test_tds_displacements_statistics():
    dx, dy, dz = O.tds_displacements(S.SEED, 3, 200000, 0.08)
    for d in (dx, dy, dz):
        assert abs(d.mean()) < 0.08 * 0.01
        assert d.std() == pytest.approx(0.08, rel=0.01)
    # configurations differ, same seed repeats
    a, _, _ = O.tds_displacements(S.SEED, 4, 1000, 0.08)
    b, _, _ = O.tds_displacements(S.SEED, 4, 1000, 0.08)
    assert np.array_equal(a, b) and not np.array_equal(a, dx[:1000])
    assert np.corrcoef(dx, dy)[0, 1] < 0.01


# ---------------------------------------------------------------- physics
def _sim_with(Z, x, y, z, fp, use_float=False):
    s = O.OracleSim(use_float)
    s.set_atoms(Z, x, y, z)
    s.bin()
    s.set_fparams(fp)
    return s


def test_wavelength_and_sigma(fp):
    s = _sim_with(*S.silicon(1, 1, 1), fp)
    for kV, lam, sig in ((200.0, 0.025116, 7.2753e-4), (300.0, 0.019718, 6.5134e-4)):
        s.apply_microscope(kV)
        s.set_calc(0.1, 0.0, 64, 64, 10, 10, 10)
        sc = s.scalars()
        assert sc["wavel"] == pytest.approx(lam, rel=2e-5)
        assert sc["sigma2"] == pytest.approx(sig, rel=2e-5)
        # physical relativistic wavelength with CODATA constants; the
        # reference's rounded constants stay within 0.2 %
        h, m, e, c = 6.62607015e-34, 9.1093837e-31, 1.602176634e-19, 2.99792458e8
        V = kV * 1e3
        lam_true = h / np.sqrt(2 * m * e * V * (1 + e * V / (2 * m * c * c))) * 1e10
        assert sc["wavel"] == pytest.approx(lam_true, rel=2e-3)


@pytest.mark.parametrize("N,ps", [(64, 0.1), (256, 0.084844), (512, 0.078125), (100, 0.2)])
def test_frequency_grid_and_kmax(fp, N, ps):
    s = _sim_with(*S.silicon(1, 1, 1), fp)
    s.apply_microscope(200.0)
    s.set_calc(ps, 0.0, N, N, 10, 10, 10)
    kx, ky = s.freqs()
    i = np.arange(N)
    want = np.where(i <= N // 2, i, i - N) / np.float32(np.float32(ps) * N)
    assert np.allclose(kx, want, rtol=1e-6) and np.allclose(ky, want, rtol=1e-6)
    assert kx[N // 2] > 0  # Nyquist bin positive
    assert s.scalars()["kmax"] == pytest.approx((N // 2 - 1) / (2 * N * ps), rel=1e-6)


def test_fd_step_size_matches_quadratic(fp):
    s = _sim_with(*S.silicon(1, 1, 1), fp)
    s.apply_microscope(300.0)
    s.set_calc(0.1, 0.0, 128, 128, 10, 10, 20.0)
    sc = s.scalars()
    assert 0 < sc["dzFD"] < 0.1
    assert sc["numberOfSlices"] == int(np.ceil(np.float32(20.0) / np.float32(sc["dzFD"])))
    s.set_calc(0.02, 0.0, 128, 128, 10, 10, 20.0, stem=True)
    assert s.scalars()["dzFD"] <= np.float32(0.06)


# ---------------------------------------------------------------- binning
def test_binning_edges_exact(fp):
    # extents: x in [0, 46] -> Min -2, Max 48 -> width 50 A -> 1 A bins
    x = np.array([0.0, 1.0, 2.5, 46.0, 20.0, 20.0], np.float32)
    y = np.array([0.0, 46.0, 3.0, 10.0, 20.0, 20.0], np.float32)
    z = np.array([0.0, 0.0, 5.0, 3.0, 1.0, 1.0], np.float32)
    Z = np.array([6, 8, 14, 38, 22, 22], np.int32)
    s = O.OracleSim()
    s.set_atoms(Z, x, y, z)
    assert list(s.extents()) == [48, -2, 48, -2, 7, -2]
    nslab = s.bin()
    assert nslab == 9
    b, zz = s.bins()
    assert list(b) == [2 + 50 * 2, 3 + 50 * 48, 4 + 50 * 5, 48 + 50 * 12, 22 + 50 * 22, 22 + 50 * 22]
    assert list(zz) == [7, 7, 2, 4, 6, 6]
    st, bx, by, bz, bZ = s.csr()
    # CSR in (slab, y, x) order, stable in input order, shifted by Minimum
    key = zz * 2500 + b
    order = np.argsort(key, kind="stable")
    assert np.array_equal(bZ, Z[order])
    assert np.array_equal(bx, x[order] - np.float32(-2))
    assert np.array_equal(np.diff(st) >= 0, np.ones(st.size - 1, bool))
    assert st[-1] == 6 and np.array_equal(np.bincount(key, minlength=nslab * 2500), np.diff(st))


def test_binning_random_matches_float32_brute_force(fp):
    Z, x, y, z = S.srtio3(6, 6, 8)
    rng = np.random.default_rng(5)
    dx, dy, dz = (rng.normal(0, 0.08, x.size).astype(np.float32) for _ in range(3))
    s = O.OracleSim()
    s.set_atoms(Z, x, y, z)
    e = s.extents()
    s.bin((dx, dy, dz))
    b, zz = s.bins()
    f32 = np.float32
    X, Y, Zp = x + dx, y + dy, z + dz
    bxw = np.floor(((X - e[1]) / f32(e[0] - e[1])) * f32(50)).astype(int)
    byw = np.floor(((Y - e[3]) / f32(e[2] - e[3])) * f32(50)).astype(int)
    zw = np.floor((f32(e[4]) - Zp) / f32(1)).astype(int)
    assert np.array_equal(b, bxw + 50 * byw) and np.array_equal(zz, zw)


# ---------------------------------------------------------------- potential
def _tiny_specimen():
    # three atoms near the centre and two far anchors fixing the extents
    Z = np.array([14, 38, 8, 6, 6], np.int32)
    x = np.array([5.0, 6.1, 5.4, -30.0, 40.0], np.float32)
    y = np.array([5.2, 4.9, 6.3, -30.0, 40.0], np.float32)
    z = np.array([3.0, 4.2, 2.1, 0.0, 9.0], np.float32)
    return Z, x, y, z


def _f32_fma(a, b, c):
    # float32 fused multiply-add: the double product of two floats is exact
    return np.float32(float(a) * float(b) + float(c))


def _brute_conv(s, Z, x, y, z, zc, convdz, N, ps, off, window=True):
    """No bins: every atom in the +-3 slab window (window=False: every atom),
    20 z samples, r^2 < 5 (decided in FP32 as the reference kernel decides
    it)."""
    e = s.extents()
    H = float(e[4] - e[5])
    kc = int(np.floor(H - zc))
    zid = np.floor((np.float32(e[4]) - z) / np.float32(1)).astype(int)
    X = (x - np.float32(e[1])).astype(np.float64)
    Y = (y - np.float32(e[3])).astype(np.float64)
    Zr = (z - np.float32(e[5])).astype(np.float64)
    fpar = s._fp
    ps, off = float(np.float32(ps)), float(np.float32(off))  # PixelScale is a float (U.cuh:107)
    out = np.zeros((N, N))
    for iy in range(N):
        for ix in range(N):
            px, py = off + ix * ps, off + iy * ps
            pxf, pyf = _f32_fma(ix, ps, off), _f32_fma(iy, ps, off)
            acc = 0.0
            for a in range(len(Z)):
                if window and abs(zid[a] - kc) > 3:
                    continue
                dxf, dyf = np.float32(pxf - np.float32(X[a])), np.float32(pyf - np.float32(Y[a]))
                rxyf = _f32_fma(dyf, dyf, np.float32(dxf * dxf))
                for m in range(1, 21):
                    z2f = np.float32(np.float32(zc) - np.float32(np.float32(m) * np.float32(convdz)) / np.float32(20))
                    ddf = np.float32(z2f - np.float32(Zr[a]))
                    if not _f32_fma(ddf, ddf, rxyf) < np.float32(5):
                        continue
                    z2 = zc - m * convdz / 20.0
                    r2 = (px - X[a]) ** 2 + (py - Y[a]) ** 2 + (z2 - Zr[a]) ** 2
                    acc += convdz / 20.0 * O.vatom(int(Z[a]), r2, fpar)
            out[iy, ix] = acc
    return out


def test_conventional_potential_brute_force(fp):
    Z, x, y, z = _tiny_specimen()
    s = _sim_with(Z, x, y, z, fp)
    s._fp = fp
    N, ps, off = 12, 0.55, 32.0
    s.apply_microscope(200.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 1, off, off)
    e = s.extents()
    H = float(e[4] - e[5])
    for zc, dz in ((H - 6.0, 1.0), (H - 8.0, 2.0)):
        got = s.potential(3, zc, dz).real
        want = _brute_conv(s, Z, x, y, z, zc, dz, N, ps, off)
        assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
        assert np.max(want) > 1.0  # the slice actually hits atoms
        t = s.potential(0, zc, dz)
        assert np.allclose(np.abs(t), 1.0, atol=1e-14)
        assert np.allclose(np.angle(t), np.angle(np.exp(1j * s.scalars()["sigma2"] * got)), atol=1e-12)


def test_potential_linear_in_atoms(fp):
    Z, x, y, z = _tiny_specimen()
    N, ps, off = 16, 0.4, 32.0

    def pot(mask, mode):
        keep = mask | np.array([0, 0, 0, 1, 1], bool)
        s = _sim_with(Z[keep], x[keep], y[keep], z[keep], fp)
        s.apply_microscope(300.0)
        s.set_calc(ps, 0.0, N, N, 1, 1, 1, off, off)
        e = s.extents()
        return s.potential(mode, float(e[4] - e[5]) - 7.0, 2.0).real

    for mode in (3, 4):
        a = pot(np.array([1, 0, 0, 0, 0], bool), mode)
        b = pot(np.array([0, 1, 1, 0, 0], bool), mode)
        ab = pot(np.array([1, 1, 1, 0, 0], bool), mode)
        assert np.allclose(a + b, ab, rtol=1e-12, atol=1e-12)
        assert ab.max() > 0


def test_projected_mode_counts_only_the_slices_own_slabs(fp):
    """The projected mode puts an atom's whole projection into the slice
    whose slabs hold it, and nothing into the neighbouring slice."""
    Z = np.array([38, 6, 6], np.int32)
    x = np.array([5.0, -30.0, 40.0], np.float32)
    y = np.array([5.0, -30.0, 40.0], np.float32)
    z = np.array([10.5, 0.0, 30.0], np.float32)
    s = _sim_with(Z, x, y, z, fp)
    N, ps, off = 16, 0.25, 33.0
    s.apply_microscope(300.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 1, off, off)
    e = s.extents()
    H = float(e[4] - e[5])
    # atom at relative z 12.5 (MinZ=-2); slab holding it for dz=1: H - zc = floor(H-12.5)
    zc = H - np.floor(H - 12.5)
    proj = s.potential(4, zc, 1.0).real
    assert proj.max() > 10
    # the projected mode only counts atoms in the slice's own slabs
    assert np.all(s.potential(4, zc - 1.0, 1.0).real == 0)


# ---------------------------------------------------------------- propagation
def _vacuum_sim(fp, N=64, ps=0.2, kV=200.0):
    Z = np.array([6, 6], np.int32)
    x = np.array([-60.0, 200.0], np.float32)
    y = np.array([-60.0, 200.0], np.float32)
    z = np.array([0.0, 20.0], np.float32)
    s = _sim_with(Z, x, y, z, fp)
    s.apply_microscope(kV)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 62.0, 62.0)
    return s


def test_vacuum_plane_wave_stays_unity(fp):
    s = _vacuum_sim(fp)
    s.init_plane(1.0)
    for it in range(1, 6):
        s.step_conv(it)
    assert np.max(np.abs(s.wave() - 1.0)) < 1e-13


def test_fresnel_gaussian_closed_form(fp):
    N, ps = 128, 0.2
    s = _vacuum_sim(fp, N, ps)
    s.init_plane(3.0)
    lam = s.scalars()["wavel"]
    c = np.arange(N) * ps
    c = c - c[N // 2]
    X, Y = np.meshgrid(c, c)
    w = 1.6
    psi0 = np.exp(-(X ** 2 + Y ** 2) / w ** 2)
    s.set_wave(np.fft.ifftshift(psi0))
    s.step_conv(2)  # iteration > 1 propagates Psi by conventionaldz in vacuum
    got = np.fft.fftshift(s.wave())
    q = w ** 2 + 1j * lam * 3.0 / np.pi
    want = (w ** 2 / q) * np.exp(-(X ** 2 + Y ** 2) / q)
    assert np.max(np.abs(got - want)) < 1e-9


# ---------------------------------------------------------------- probe
def test_probe_closed_form_and_shift(fp):
    N, ps = 64, 0.25
    s = _vacuum_sim(fp, N, ps, kV=300.0)
    s.apply_microscope(300.0, obj=25.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 0, 0, stem=True)
    p0 = s.probe(0, 0)
    assert np.sum(np.abs(p0)) == pytest.approx(1.0, rel=1e-12)  # L1 normalised (U.cu:2698-2705)
    F = np.fft.fft2(p0)
    kx, ky = s.freqs()
    kk = np.sqrt(kx[None, :].astype(np.float64) ** 2 + ky[:, None] ** 2)
    inside = kk < 25e-3 / s.scalars()["wavel"]
    assert inside.sum() > 10
    amp = np.abs(F[inside])
    assert np.allclose(amp, amp[0], rtol=1e-10)  # aberration-free: flat disc
    assert np.max(np.abs(F[~inside])) < 1e-12 * amp[0]
    # shift theorem: position r0 lands the probe at -r0 (mod N)
    p = s.probe(5, 9)
    # (the FP32 k-grid of U.cu:1889-1903 limits the phase ramp to ~1e-7)
    assert np.max(np.abs(p - np.roll(p0, (-9, -5), axis=(0, 1)))) < 1e-7 * np.max(np.abs(p0))


# ---------------------------------------------------------------- CTF
def test_ctf_identity_and_weak_phase_chi(fp):
    N, ps = 64, 0.25
    s = _vacuum_sim(fp, N, ps, kV=200.0)
    s.apply_microscope(200.0, obj=1000.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 62.0, 62.0)
    s.init_plane(1.0)
    rng = np.random.default_rng(2)
    kx, ky = s.freqs()
    kk = np.sqrt(kx[None, :].astype(np.float64) ** 2 + ky[:, None] ** 2)
    band = np.fft.fft2(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    band[kk > 0.5 * kk.max()] = 0
    psi = np.fft.ifft2(band)
    s.set_wave(psi)
    assert np.max(np.abs(s.ctf_image(0.0) - np.abs(psi) ** 2)) < 1e-12 * np.max(np.abs(psi) ** 2)
    # weak phase object exp(i eps cos(2 pi k0 x)) imaged through psi_hat e^{-i chi}
    # (U.cu:1073-1074): psi' = 1 + i eps e^{-i chi} cos(.), so
    # I = 1 + 2 eps sin(chi(k0)) cos(2 pi k0 x) to first order in eps
    Cs, df = 1.2e7, -500.0
    s.apply_microscope(200.0, df=df, Cs=Cs, obj=1000.0)
    s.set_calc(ps, df, N, N, 1, 1, 22, 62.0, 62.0)
    lam = s.scalars()["wavel"]
    for j in (3, 7, 12):
        x = np.arange(N) * ps
        eps = 1e-4
        ph = np.tile(np.exp(1j * eps * np.cos(2 * np.pi * j * np.arange(N) / N)), (N, 1))
        s.set_wave(ph)
        I = s.ctf_image(df)
        k0 = j / (N * ps)
        chi = np.pi * lam * k0 ** 2 * (0.5 * Cs * lam ** 2 * k0 ** 2 + df)
        want = 1 + 2 * eps * np.sin(chi) * np.cos(2 * np.pi * j * np.arange(N) / N)
        assert np.max(np.abs(I[0] - want)) < 5e-8


# ---------------------------------------------------------------- detectors
def test_adf_parseval_and_vacuum_zero(fp):
    N, ps = 64, 0.2
    s = _vacuum_sim(fp, N, ps, kV=300.0)
    s.apply_microscope(300.0, obj=20.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 0, 0, stem=True)
    s.init_stem(32, 32, 1.0)
    w = s.wave(1)
    s.set_wave(w)
    assert s.adf(0.0, 1e9) == pytest.approx(np.sum(np.abs(w) ** 2), rel=1e-12)
    assert s.adf(25.0, 200.0) < 1e-25
    assert s.adf(0.0, 20.0) == pytest.approx(np.sum(np.abs(w) ** 2), rel=1e-10)


def test_tds_accumulates_centred_amplitude(fp):
    N, ps = 32, 0.2
    s = _vacuum_sim(fp, N, ps, kV=300.0)
    s.apply_microscope(300.0, obj=20.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 0, 0, stem=True)
    s.init_stem(16, 16, 1.0)
    w = s.wave(1)
    s.set_wave(w)
    s.add_tds(); s.add_tds()
    assert np.allclose(s.tds(), 2 * np.abs(np.fft.fftshift(np.fft.fft2(w))), rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- FD
def test_fd_vacuum_dispersion(fp):
    """FD scheme on a tilted plane wave obeys its characteristic equation
    c+ r^2 - (2 + 4 pi^2 k^2 dz^2) r + c- = 0 (pi = 3.14159 as U.cu:1006)."""
    N, ps = 32, 0.3
    s = _vacuum_sim(fp, N, ps, kV=200.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 0, 0)
    sc = s.scalars()
    dz, lam = np.float32(sc["dzFD"]), sc["wavel"]
    kx, _ = s.freqs()
    j = 3
    k = float(kx[j])
    assert k < sc["kmax"]
    pw = np.tile(np.exp(2j * np.pi * j * np.arange(N) / N), (N, 1))
    p9 = 3.14159
    a = 2 * p9 * float(dz) / lam
    cp, cm = 1 + 1j * a, 1 - 1j * a
    roots = np.roots([cp, -(2 + 4 * p9 ** 2 * k ** 2 * float(dz) ** 2), cm])
    psi_m, psi0 = 1.0 + 0j, np.exp(-1j * np.pi * float(dz) * lam * k * k)
    A = np.linalg.solve(np.array([[1, 1], roots]), np.array([psi_m, psi0]))
    s.init_plane_fd()
    s.set_wave(pw * psi_m, 1)
    s.set_wave(pw * psi0, 0)
    for n in range(1, 6):
        s.step_fd(n)
        want = A[0] * roots[0] ** (n + 1) + A[1] * roots[1] ** (n + 1)
        assert np.max(np.abs(s.wave() - want * pw)) < 1e-10


# ---------------------------------------------------------------- adjoint
def test_oracle_adjoint_matches_central_differences():
    N, ns, npb = 16, 4, 2
    s = _vacuum_sim(np.ones((103, 12), np.float32), N, 0.3, kV=300.0)
    rng = np.random.default_rng(7)
    V = (rng.random((ns, N, N)) * 40).astype(np.float32)
    psi0 = rng.standard_normal((npb, N, N)) + 1j * rng.standard_normal((npb, N, N))
    target = rng.random((npb, N, N))
    I, _ = s.diff(V, psi0, None)
    _, g = s.diff(V, psi0, 2 * (I - target))
    for (i, y, x) in [(0, 1, 2), (1, 5, 5), (3, 15, 0), (2, 8, 11)]:
        h = 1e-2
        Vp = V.astype(np.float64).copy()
        vals = []
        for sgn in (1, -1):
            Vq = V.copy()
            Vq[i, y, x] = np.float32(V[i, y, x] + sgn * h)
            Iq, _ = s.diff(Vq, psi0, None)
            vals.append(np.sum((Iq - target) ** 2))
        hh = float(np.float32(V[i, y, x] + h)) - float(np.float32(V[i, y, x] - h))
        fd = (vals[0] - vals[1]) / hh
        assert g[i, y, x] == pytest.approx(fd, rel=1e-5, abs=1e-10)


# ---------------------------------------------------------------- loader
def test_xyz_reference_symbol_map(tmp_path):
    assert O.ref_znum("Si") == 14 and O.ref_znum("Sr") == 38 and O.ref_znum("O") == 8
    assert O.ref_znum("Ti") == 1 and O.ref_znum("Au") == 1  # unknown -> H (U.cu:1241)
    p = tmp_path / "a.xyz"
    p.write_text("3\n\nSi 0 0 0\nTi 1.5 2 3\nO 4 5 6.25\n")
    s = O.OracleSim()
    assert s.load_xyz_refmap(str(p)) == 3
    assert list(s.extents()) == [6.0, -2.0, 7.0, -2.0, 8.25, -2.0]
    p.write_text("2\ncomment line here\nSi 0 0 0\nO 1 1 1\n")
    assert s.load_xyz_refmap(str(p)) == 2


def test_jittered_atoms_keep_min_distance_and_fill_box():
    from scipy.spatial import cKDTree
    from cudamultislice_amd import structures as S
    Z, x, y, z = S.jittered_atoms(3000, (30.0, 25.0, 40.0), seed=5)
    p = np.stack([x, y, z], 1).astype(float)
    d, _ = cKDTree(p).query(p, k=2)
    assert d[:, 1].min() >= 1.4 - 1e-5
    assert p.min(0).min() >= 0 and np.all(p.max(0) <= [30.0, 25.0, 40.0])
    assert Z.min() >= 6 and Z.max() <= 79 and len(set(Z.tolist())) > 60
    # spread: every eighth of the box holds atoms
    occ = np.histogramdd(p, bins=(2, 2, 2), range=((0, 30), (0, 25), (0, 40)))[0]
    assert occ.min() > 0.08 * len(p)


# ---------------------------------------------------------------- detector model
def test_detector_table_lookup_uses_signed_binned_radius():
    from cudamultislice_amd.detector import tables
    dqe, ntf = tables()
    assert dqe.size == ntf.size == 725 and dqe[0] == np.float32(0.72963) and ntf[0] == np.float32(1.14689)
    N = 256
    for k in (0, 1, 5, 31):
        assert O.detector_table(dqe, 4 * k, 0, N, N) == dqe[k]       # r = 4k / 4
        assert O.detector_table(dqe, 0, N - 4 * k if k else 0, N, N) == dqe[k]  # negative frequency
        assert O.detector_table(ntf, 4 * k + 3, 0, N, N) == ntf[k]   # floor
    # the index N/2 is the negative Nyquist: |-N/2| / 4
    assert O.detector_table(dqe, N // 2, 0, N, N) == dqe[N // 8]
    # 3-4-5: sqrt(12^2 + 16^2) / 4 = 5
    assert O.detector_table(dqe, 12, N - 16, N, N) == dqe[5]


def test_detector_normal_is_a_clamped_standard_normal():
    from scipy.stats import norm
    g = np.array([O.detector_normal(7, i) for i in range(100000)])
    assert g.min() >= -3 and g.max() <= 3 and np.sum(np.abs(g) == 3) > 0
    q = norm.sf(3)
    var = (1 - 2 * q) - 2 * 3 * norm.pdf(3) + 2 * 9 * q  # E[clamp(g, -3, 3)^2]
    assert abs(g.mean()) < 0.01 and abs(g.var() - var) < 0.01
    assert O.detector_normal(7, 5) == O.detector_normal(7, 5) != O.detector_normal(8, 5)


def test_detector_constant_image_mean_is_the_dc_gain():
    from cudamultislice_amd.detector import tables
    dqe, ntf = tables()
    N, I0, dose = 128, 0.8, 50.0
    out, st = O.detector_image(np.full((N, N), I0), dose, 3, dqe, ntf)
    want = 64 * float(ntf[0]) * dose * I0 * np.sqrt(float(dqe[0]))
    assert abs(out.real.mean() / want - 1) < 2e-3
    assert np.abs(out.imag).max() < 1e-9 * want
    assert st[0] == pytest.approx(np.abs(out).max()) and st[1] == pytest.approx(np.abs(out).min())
    assert st[2] == pytest.approx(np.abs(out).std(), rel=1e-9)


def test_detector_flat_curves_give_rounded_counts_with_shot_noise():
    N, I0, dose = 128, 0.5, 400.0
    ones = np.ones(725, np.float32)
    out, _ = O.detector_image(np.full((N, N), I0), dose, 11, ones, ones)
    c = out.real / 64
    assert np.abs(c - np.round(c)).max() < 1e-9  # integer counts
    from scipy.stats import norm
    q = norm.sf(3)
    gvar = (1 - 2 * q) - 2 * 3 * norm.pdf(3) + 2 * 9 * q
    m = dose * I0
    assert abs(c.mean() / m - 1) < 2e-3
    assert abs(c.var() / (m * gvar + 1 / 12) - 1) < 0.03


def test_detector_transfer_of_a_cosine_follows_the_curves():
    from cudamultislice_amd.detector import tables
    dqe, ntf = tables()
    N, eps, k0 = 256, 0.2, 37
    x = np.arange(N)
    I = np.tile(1 + eps * np.cos(2 * np.pi * k0 * x / N), (N, 1))
    out, _ = O.detector_image(I, 1e7, 5, dqe, ntf)
    F = np.fft.fft2(out.real)
    ratio = abs(F[0, k0]) / abs(F[0, 0])
    i = k0 // 4
    want = eps / 2 * np.sqrt(float(dqe[i]) / float(dqe[0])) * float(ntf[i]) / float(ntf[0])
    assert ratio == pytest.approx(want, rel=2e-3)


# ---------------------------------------------------------------- real-space multislice
def _brute_rs(s, Z, x, y, z, zc, N, ps, off):
    """No bins: binnedAtomPKernelRS restated per pixel over all atoms in the
    +-3 slab window, 10 samples z - m/10 of weight 1/10, r^2 < 5 in FP32."""
    e = s.extents()
    H = float(e[4] - e[5])
    kc = int(np.floor(np.float32(H) - np.float32(zc)))
    zid = np.floor((np.float32(e[4]) - z) / np.float32(1)).astype(int)
    X, Y = (x - np.float32(e[1])).astype(np.float64), (y - np.float32(e[3])).astype(np.float64)
    Zr = (z - np.float32(e[5])).astype(np.float64)
    ps, off = float(np.float32(ps)), float(np.float32(off))
    out = np.zeros((N, N))
    for iy in range(N):
        for ix in range(N):
            px, py = off + ix * ps, off + iy * ps
            pxf, pyf = _f32_fma(ix, ps, off), _f32_fma(iy, ps, off)
            for a in range(len(Z)):
                if abs(zid[a] - kc) > 3:
                    continue
                dxf, dyf = np.float32(pxf - np.float32(X[a])), np.float32(pyf - np.float32(Y[a]))
                rxyf = _f32_fma(dyf, dyf, np.float32(dxf * dxf))
                for m in range(1, 11):
                    z2f = np.float32(np.float32(zc) - np.float32(m) / np.float32(10))
                    ddf = np.float32(z2f - np.float32(Zr[a]))
                    if not _f32_fma(ddf, ddf, rxyf) < np.float32(5):
                        continue
                    r2 = (px - X[a]) ** 2 + (py - Y[a]) ** 2 + (zc - m / 10.0 - Zr[a]) ** 2
                    out[iy, ix] += O.vatom(int(Z[a]), r2, s._fp) / 10.0
    return out


def test_rs_potential_brute_force(fp):
    Z, x, y, z = _tiny_specimen()
    s = _sim_with(Z, x, y, z, fp)
    s._fp = fp
    N, ps, off = 12, 0.55, 32.0
    s.apply_microscope(200.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 1, off, off)
    H = float(s.extents()[4] - s.extents()[5])
    for zc in (H - 6.0, H - 7.5):
        zc = float(np.float32(zc))  # the step passes the FP32 plane H - it
        got = s.potential(5, zc, 1.0).real
        want = _brute_rs(s, Z, x, y, z, zc, N, ps, off)
        assert np.max(want) > 1.0
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_rs_stencil_delta_responses_and_boundary_rule(fp):
    Z, x, y, z = _tiny_specimen()
    s = _sim_with(Z, x, y, z, fp)
    N, ps = 16, 0.5
    s.apply_microscope(200.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 1, 0.0, 0.0)
    h2 = 1.0 / (ps * ps)
    d = np.zeros((N, N), complex)
    d[7, 6] = 1.0
    L = s.rs_laplacian(d)
    # interior: centre 2 (-4/3 + 1/12), first neighbours 4/3, second -1/12 (f subtracted once)
    assert L[7, 6] == pytest.approx(2 * (-4 / 3 + 1 / 12) * h2)
    for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        assert L[7 + dy, 6 + dx] == pytest.approx(4 / 3 * h2)
        assert L[7 + 2 * dy, 6 + 2 * dx] == pytest.approx(-1 / 12 * h2)
    assert np.count_nonzero(np.abs(L) > 1e-12) == 9
    # element 0 of a row stands in for every missing neighbour
    d = np.zeros((N, N), complex)
    d[5, 0] = 1.0
    L = s.rs_laplacian(d)
    row = L[5]
    assert row[0] == pytest.approx(2 * (-1 / 12) * h2 * 0 + (-4 / 3 + 1 / 12) * h2 * 0 + (-4 / 3 + 1 / 12) * h2)
    assert row[1] == pytest.approx((4 / 3 - 1 / 12) * h2)      # x-1 and x-2 both read element 0
    assert row[2] == pytest.approx(-1 / 12 * h2)
    assert row[N - 1] == pytest.approx((4 / 3 - 1 / 12) * h2)  # x+1 and x+2 read element 0
    assert row[N - 2] == pytest.approx(-1 / 12 * h2)


def test_rs_step_in_vacuum_follows_the_series_closed_form(fp):
    # atoms far outside the grid: V = 0, a constant wave sees L w = (5/2) w / ps^2
    Z = np.array([6, 6], np.int32)
    s = _sim_with(Z, np.array([-60, 200], np.float32), np.array([-60, 200], np.float32),
                  np.array([0, 30], np.float32), fp)
    N, ps = 16, 0.3
    s.apply_microscope(200.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 1, 0.0, 0.0)
    s.init_plane(1.0)
    lam = float(np.float32(s.scalars()["wavel"]))
    base = lam / (4 * float(np.float32(3.1415927)))
    p32 = float(np.float32(ps))  # PixelScale is a float
    c = 2.5 / (p32 * p32)
    w = 1.0 + 0j
    for m in range(2, 11):
        w *= 1 + 1j * base / m * c
    for it in (1, 2):
        s.step_rs(it)
        assert np.allclose(s.wave(0), w, rtol=1e-12, atol=0)
        assert np.allclose(s.wave(2), w, rtol=1e-12, atol=0)
    assert np.allclose(s.wave(1), w * w, rtol=1e-12)  # step 2: PsiMinus = PsiPlus * Psi


# ---------------------------------------------------------------- further pins
def _brute_fd(s, Z, x, y, z, zc, N, ps):
    """binnedAtomPKernelFD (U.cu:883-946) without bins: pixel (i ps, j ps)
    (no block offset), the atoms of 1 A slabs floor(H - z) +- 3, the 3D
    r^2 < 5 test taken in FP32 as the kernel takes it, vatom summed at the
    plane z."""
    e = s.extents()
    H = np.float32(e[4] - e[5])
    kc = int(np.floor(np.float32(H - np.float32(zc)) / np.float32(1)))
    zid = np.floor((np.float32(e[4]) - z) / np.float32(1)).astype(int)
    X, Y, Zr = (x - np.float32(e[1])), (y - np.float32(e[3])), (z - np.float32(e[5]))
    ps32 = np.float32(ps)
    out = np.zeros((N, N))
    for iy in range(N):
        for ix in range(N):
            pxf, pyf = np.float32(np.float32(ix) * ps32), np.float32(np.float32(iy) * ps32)
            acc = 0.0
            for a in range(len(Z)):
                if abs(zid[a] - kc) > 3:
                    continue
                dxf, dyf = np.float32(pxf - X[a]), np.float32(pyf - Y[a])
                ddf = np.float32(np.float32(zc) - Zr[a])
                if not _f32_fma(ddf, ddf, _f32_fma(dyf, dyf, np.float32(dxf * dxf))) < np.float32(5):
                    continue
                r2 = (float(pxf) - float(X[a])) ** 2 + (float(pyf) - float(Y[a])) ** 2 + (zc - float(Zr[a])) ** 2
                acc += O.vatom(int(Z[a]), r2, s._fp)
            out[iy, ix] = acc
    return out


def test_fd_potential_brute_force(fp):
    """The finite-difference potential (mode 2) against a bin-free sum over
    every atom: the same values wherever the slab window and cutoff admit
    an atom, zero elsewhere."""
    Z = np.array([14, 38, 8, 6, 6], np.int32)
    x = np.array([1.0, 2.1, 1.4, -2.0, 40.0], np.float32)
    y = np.array([1.2, 0.9, 2.3, -2.0, 40.0], np.float32)
    z = np.array([3.0, 4.2, 2.1, 0.0, 9.0], np.float32)
    s = _sim_with(Z, x, y, z, fp)
    s._fp = fp
    N, ps = 14, 0.5
    s.apply_microscope(200.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 1)
    e = s.extents()
    H = float(e[4] - e[5])
    for depth in (6.3, 7.05, 9.5):
        zc = float(np.float32(H - depth))
        got = s.potential(2, zc, 1.0)
        assert np.all(got.imag == 0)
        want = _brute_fd(s, Z, x, y, z, zc, N, ps)
        assert np.max(want) > 1.0
        assert np.max(np.abs(got.real - want)) <= 1e-12 * np.max(want)
        assert np.all((got.real == 0) == (want == 0))


def _fd_dz_reference_formula(kV, ps, N):
    """SetCalculationVariables' step (U.cu:1791-1834), restated in float64
    from the published constants: the root X of A X^2 + B X + 3 = 0 and
    0.99 sqrt(X)."""
    Pi, e, m, c2 = 3.1415926, 1.6e-19, 9.11e-31, 9e16
    eV = e * kV * 1e3
    lam = 6.63e-34 * 3e8 / np.sqrt(eV * (2 * m * c2 + eV)) * 1e10
    sig = 2 * Pi / (lam * kV * 1e3) * (m * c2 + eV) / (2 * m * c2 + eV)
    V2 = kV * 1e3
    p = N / (2 * ps * N)
    ke2 = 0.666666 * 2 * p * p
    u = ke2 - sig * V2 / (Pi * lam)
    A = 16 * Pi ** 4 * u * u
    B = 16 * Pi ** 2 * (u - 1 / (4 * lam * lam))
    X = (-B + np.sqrt(B * B - 12 * A)) / (2 * A)
    return 0.99 * np.sqrt(X), lam, sig


@pytest.mark.parametrize("kV,ps,N", [(300.0, 0.1, 128), (200.0, 0.2, 64), (300.0, 0.02, 128), (100.0, 0.05, 256)])
def test_fd_step_size_is_stable_and_follows_the_quadratic(fp, kV, ps, N):
    """dzFD (U.cu:1815-1834): equal to an independent float64 evaluation of
    the reference's quadratic (FP32 rounding only), and inside the von
    Neumann stability region of the scheme it drives.  For a mode of
    transverse k^2 in a potential V the update c+ r^2 - b r + c- = 0, with
    c+- = 1 +- 2 pi i dz / lambda and b = 2 + 4 pi^2 dz^2 (k^2 - sigma V /
    (pi lambda)), keeps both roots on the unit circle iff
    b^2 <= 4 |c+|^2.  At the step size this holds for every k^2 up to the
    band edge k_e^2 = 2/3 (kx_N^2 + ky_N^2) and V from 0 to the accelerating
    voltage (the formula's design range), and the step is a fixed fraction
    (0.70-0.75) of the stability limit at k_e^2, V = kV."""
    s = _sim_with(*S.silicon(1, 1, 1), fp)
    s.apply_microscope(kV)
    s.set_calc(ps, 0.0, N, N, 10, 10, 20.0)
    sc = s.scalars()
    want, lam, sig = _fd_dz_reference_formula(kV, ps, N)
    assert sc["dzFD"] == pytest.approx(want, rel=2e-5)
    dz = sc["dzFD"]
    ke2 = 0.666666 * 2 * (1 / (2 * ps)) ** 2
    for V in np.linspace(0.0, kV * 1e3, 7):
        for k2 in np.linspace(0.0, ke2, 9):
            u = k2 - sig * V / (np.pi * lam)
            b = 2 + 4 * np.pi ** 2 * dz ** 2 * u
            assert b * b <= 4 * (1 + (2 * np.pi * dz / lam) ** 2), (V, k2)
            r = np.roots([1 + 2j * np.pi * dz / lam, -b, 1 - 2j * np.pi * dz / lam])
            assert np.allclose(np.abs(r), 1.0, atol=1e-6)
    u = ke2 - sig * kV * 1e3 / (np.pi * lam)
    limit = np.sqrt((1 / lam ** 2 - u) / (np.pi ** 2 * u * u))
    assert 0.70 < dz / limit < 0.75, dz / limit


def test_fd_dispersion_in_a_constant_potential(fp):
    """FD scheme (U.cu:995-1023) on a tilted plane wave in a constant
    potential V0: the characteristic equation gains the coupling term,
    c+ r^2 - (2 + 4 pi^2 k^2 dz^2 - 4 pi sigma V0 dz^2 / lambda) r + c- = 0
    (pi = 3.14159 as the kernel writes it); a constant potential survives
    the band limit unchanged."""
    N, ps = 32, 0.3
    s = _vacuum_sim(fp, N, ps, kV=200.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 0, 0)
    sc = s.scalars()
    dz, lam, sig = float(np.float32(sc["dzFD"])), sc["wavel"], sc["sigma2"]
    kx, _ = s.freqs()
    j = 4
    k = float(kx[j])
    p9 = 3.14159
    for V0 in (15.0, 120.0):
        b = 2 + 4 * p9 ** 2 * k * k * dz ** 2 - 4 * p9 * sig * V0 * dz ** 2 / lam
        a = 2 * p9 * dz / lam
        roots = np.roots([1 + 1j * a, -b, 1 - 1j * a])
        pw = np.tile(np.exp(2j * np.pi * j * np.arange(N) / N), (N, 1))
        psi_m, psi0 = 1.0 + 0j, np.exp(-0.3j)
        A = np.linalg.solve(np.array([[1, 1], roots]), np.array([psi_m, psi0]))
        s.init_plane_fd()
        s.set_fd_potential(np.full((N, N), V0))
        s.set_wave(pw * psi_m, 1)
        s.set_wave(pw * psi0, 0)
        for n in range(1, 7):
            s.step_fd(n)
            want = A[0] * roots[0] ** (n + 1) + A[1] * roots[1] ** (n + 1)
            assert np.max(np.abs(s.wave() - want * pw)) < 1e-10, (V0, n)
        # the potential shifts the phase velocity: not the vacuum solution
        rv = np.roots([1 + 1j * a, -(2 + 4 * p9 ** 2 * k * k * dz ** 2), 1 - 1j * a])
        Av = np.linalg.solve(np.array([[1, 1], rv]), np.array([psi_m, psi0]))
        assert np.max(np.abs(s.wave() - (Av[0] * rv[0] ** 7 + Av[1] * rv[1] ** 7) * pw)) > 1e-6
    s.set_fd_potential(None)


def test_integrated_mode_over_all_slices_is_the_truncated_projection(fp):
    """Summed over the 1 A slices that cover an isolated atom, the
    integrated mode (20 samples per slice, U.cu:582-653) is the z integral
    of vatom over the sphere r^2 < 5 (Riemann sum of step dz/20), and the
    projected mode (U.cu:730-795) the whole projection vzatom: the two
    modes agree up to the tail beyond the cutoff, which a quadrature of
    vatom gives independently."""
    from scipy.integrate import quad
    Z = np.array([38, 6, 6], np.int32)
    x = np.array([5.0, -30.0, 40.0], np.float32)
    y = np.array([5.0, -30.0, 40.0], np.float32)
    z = np.array([10.5, 0.0, 30.0], np.float32)
    s = _sim_with(Z, x, y, z, fp)
    N, ps, off = 16, 0.25, 33.0
    s.apply_microscope(300.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 1, off, off)
    e = s.extents()
    H = float(e[4] - e[5])
    n = int(np.ceil(H))
    integ = sum(s.potential(3, float(np.float32(H - it)), 1.0).real for it in range(1, n + 1))
    proj = sum(s.potential(4, float(np.float32(H - it)), 1.0).real for it in range(1, n + 1))
    X = float(np.float32(5.0) - e[1])
    c = off + np.arange(N) * ps
    checked = 0
    for iy in range(N):
        for ix in range(N):
            r2 = (c[ix] - X) ** 2 + (c[iy] - X) ** 2
            if not 0.25 < r2 < 2.0:
                continue
            zmax = np.sqrt(5.0 - r2)
            full = O.vzatom(38, r2, fp)
            cut = quad(lambda t: O.vatom(38, r2 + t * t, fp), -zmax, zmax, limit=200)[0]
            assert proj[iy, ix] == pytest.approx(full, rel=1e-12)
            # a right-end Riemann sum of step 0.05 A of a function cut off at
            # its ends: within 1 % (one sample in 20 dropped would be 5 %)
            assert integ[iy, ix] == pytest.approx(cut, rel=1e-2)
            assert 0.8 * full < integ[iy, ix] < full
            checked += 1
    assert checked >= 12


def _grating(N, j, l, eps):
    yy, xx = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    return np.exp(1j * eps * np.cos(2 * np.pi * (j * xx + l * yy) / N)), np.cos(2 * np.pi * (j * xx + l * yy) / N)


def test_ctf_spatial_envelope_closed_form(fp):
    """Partial spatial coherence of imagingKernel2 (U.cu:1064-1066) against
    the textbook envelope E_s(k) = exp(-(pi alpha / lambda)^2 (Cs lambda^3 k^3
    + df lambda k)^2), alpha the illumination semi-angle (Kirkland, Advanced
    Computing in Electron Microscopy, ch. 3): a weak phase grating at k0 images
    as 1 + 2 eps E_s(k0) sin chi(k0) cos(2 pi k0 x)."""
    N, ps, Cs, df, beta = 64, 0.25, 1.2e7, -500.0, 0.6
    s = _vacuum_sim(fp, N, ps, kV=200.0)
    s.apply_microscope(200.0, df=df, Cs=Cs, beta=beta, obj=1000.0)
    s.set_calc(ps, df, N, N, 1, 1, 22, 62.0, 62.0)
    s.init_plane(1.0)
    lam = s.scalars()["wavel"]
    eps = 1e-4
    for j in (3, 7, 12, 17):
        ph, cosg = _grating(N, j, 0, eps)
        s.set_wave(ph)
        k0 = j / (N * ps)
        chi = np.pi * lam * k0 ** 2 * (0.5 * Cs * lam ** 2 * k0 ** 2 + df)
        Es = np.exp(-(np.pi * beta * 1e-3 / lam) ** 2 * (Cs * lam ** 3 * k0 ** 3 + df * lam * k0) ** 2)
        assert Es < 0.999 or j == 3
        want = 1 + 2 * eps * Es * np.sin(chi) * cosg
        assert np.max(np.abs(s.ctf_image(df) - want)) < 5e-8, j


def test_ctf_temporal_envelope_is_a_defocus_average(fp):
    """Partial temporal coherence of imagingKernel2 (U.cu:1067-1068): to first
    order in a weak phase grating, the image with spread delta equals the
    average of coherent images over Gaussian defocus offsets of standard
    deviation delta / sqrt(2) (Gauss-Hermite quadrature of the oracle's own
    delta = 0 images), independent of the envelope's formula."""
    N, ps, Cs, df, delta = 64, 0.25, 1.2e7, -500.0, 60.0
    s = _vacuum_sim(fp, N, ps, kV=200.0)
    s.set_calc(ps, df, N, N, 1, 1, 22, 62.0, 62.0)
    s.init_plane(1.0)
    xs, ws = np.polynomial.hermite_e.hermegauss(40)
    ws = ws / ws.sum()
    eps = 1e-4
    for j in (5, 11, 16):
        ph, _ = _grating(N, j, 0, eps)
        s.set_wave(ph)
        s.apply_microscope(200.0, df=df, Cs=Cs, delta=delta, obj=1000.0)
        I = s.ctf_image(df)
        s.apply_microscope(200.0, df=df, Cs=Cs, obj=1000.0)
        avg = sum(w * s.ctf_image(df + delta / np.sqrt(2) * xi) for xi, w in zip(xs, ws))
        assert np.max(np.abs(I - avg)) < 5e-8, j
        assert np.max(np.abs(I - s.ctf_image(df))) > 1e-6  # the spread acts at this k


def test_ctf_twofold_astigmatism_phase(fp):
    """Twofold astigmatism A1 in chi (U.cu:1070): chi = pi lambda k^2 A1
    sin(2 (phi - phi_A1)), phi = atan2(ky, kx); weak phase gratings along
    several directions image as 1 + 2 eps sin chi(k0) cos(2 pi k0 . r)."""
    N, ps, A1, A1phi = 64, 0.25, 150.0, 0.3
    s = _vacuum_sim(fp, N, ps, kV=200.0)
    s.apply_microscope(200.0, A1=A1, A1phi=A1phi, obj=1000.0)
    s.set_calc(ps, 0.0, N, N, 1, 1, 22, 62.0, 62.0)
    s.init_plane(1.0)
    lam = s.scalars()["wavel"]
    eps = 1e-4
    seen = []
    for j, l in ((6, 0), (0, 6), (5, 5), (-4, 7), (9, -3)):
        ph, cosg = _grating(N, j, l, eps)
        s.set_wave(ph)
        kx, ky = j / (N * ps), l / (N * ps)
        chi = np.pi * lam * (kx * kx + ky * ky) * A1 * np.sin(2 * (np.arctan2(ky, kx) - A1phi))
        seen.append(np.sin(chi))
        want = 1 + 2 * eps * np.sin(chi) * cosg
        assert np.max(np.abs(s.ctf_image(0.0) - want)) < 5e-8, (j, l)
    assert max(seen) > 0.3 and min(seen) < -0.3  # both signs of the twofold term


def test_aberrated_probe_phase(fp):
    """STEM probe with Cs and defocus (U.cu:368-375): its spectrum inside
    the aperture is a constant times exp(-i chi(k)), chi = pi lambda k^2
    (Cs lambda^2 k^2 / 2 + df), and zero outside."""
    N, ps = 64, 0.2
    s = _vacuum_sim(fp, N, ps, kV=300.0)
    for Cs, df in ((1e4, -50.0), (1.2e7, 300.0), (0.0, -700.0)):
        s.apply_microscope(300.0, df=df, Cs=Cs, obj=25.0)
        s.set_calc(ps, df, N, N, 1, 1, 22, 0, 0, stem=True)
        lam = s.scalars()["wavel"]
        p = s.probe(0, 0)
        F = np.fft.fft2(p)
        kx, ky = s.freqs()
        k2 = kx[None, :].astype(np.float64) ** 2 + ky[:, None].astype(np.float64) ** 2
        inside = np.sqrt(k2) < 25e-3 / s.scalars()["wavel_f"]
        chi = np.pi * lam * k2 * (0.5 * Cs * lam ** 2 * k2 + df)
        assert np.ptp(chi[inside]) > 2.0  # the aberration matters
        rel = F[inside] / F[0, 0] * np.exp(1j * chi[inside])
        # pi is 3.14159265 in the probe kernel: a phase error of 3.6e-9 chi
        assert np.all(np.abs(rel - 1.0) <= 4e-9 * np.abs(chi[inside]) + 1e-10), (Cs, df)
        assert np.max(np.abs(F[~inside])) < 1e-12 * np.abs(F[0, 0])
        assert np.sum(np.abs(p)) == pytest.approx(1.0, rel=1e-12)


def test_fd_batch_equals_probe_by_probe_steps(fp):
    """The probe-parallel finite-difference STEM batch gives each probe's
    ADF of STEMInitialiseWavefunctions followed by MultisliceStep calls."""
    Z, x, y, z = S.silicon(1, 1, 2)
    s = _sim_with(Z, x, y, z, fp)
    s.apply_microscope(300.0, obj=25.0)
    s.set_calc(0.25, 0.0, 32, 32, 1, 1, float(s.extents()[4] - s.extents()[5]), 0, 0, stem=True)
    px, py = np.array([3, 17, 30], np.int32), np.array([9, 0, 22], np.int32)
    first = 40
    batch = s.fd_batch(px, py, first, 5.0, 200.0)
    for p in range(3):
        s.init_stem(int(px[p]), int(py[p]))
        for it in range(1, first + 1):
            s.step_fd(it)
        assert batch[p] == pytest.approx(s.adf(5.0, 200.0), rel=1e-12)
    assert batch.min() > 0



# ---------------------------------------------------------------- compat=physical
def _phys(Z, x, y, z, fp):
    s = O.OracleSim(physical=True)
    s.set_atoms(Z, x, y, z)
    s.bin()
    s.set_fparams(fp)
    s._fp = fp
    return s


def test_physical_window_reaches_every_atom_of_a_thick_slice(fp):
    """Q3 (U.cu:602-605): at 2 A slices the reference's +-3 slab window
    misses atoms within reach of the slice's lower samples; under
    compat=physical the integrated potential equals the bin-free sum over
    every atom (3D cutoff only), where the reference mode does not."""
    Z = np.array([14, 38, 8, 6, 6, 14], np.int32)
    x = np.array([5.0, 6.1, 5.4, -30.0, 40.0, 5.6], np.float32)
    y = np.array([5.2, 4.9, 6.3, -30.0, 40.0, 5.5], np.float32)
    # the last Si atom sits in the slab four below the slice's own, within
    # reach of its lowest samples
    z = np.array([3.0, 4.2, 2.1, 0.0, 12.0, 0.9], np.float32)
    N, ps, off = 12, 0.55, 32.0
    sp, sr = _phys(Z, x, y, z, fp), _sim_with(Z, x, y, z, fp)
    sr._fp = fp
    for s in (sp, sr):
        s.apply_microscope(200.0)
        s.set_calc(ps, 0.0, N, N, 1, 1, 1, off, off)
    e = sp.extents()
    H = float(e[4] - e[5])
    zc = float(np.float32(H - 9.0))  # samples 9-11 A below the top; the atoms sit 10-13 A down
    want = _brute_conv(sp, Z, x, y, z, zc, 2.0, N, ps, off, window=False)
    got = sp.potential(3, zc, 2.0).real
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(want)
    ref = sr.potential(3, zc, 2.0).real
    assert np.max(np.abs(ref - _brute_conv(sr, Z, x, y, z, zc, 2.0, N, ps, off))) <= 1e-12 * np.max(want)
    assert np.max(want - ref) > 1e-4 * np.max(want)  # the reference window loses a contribution


def test_physical_slices_integrate_the_top_slab(fp):
    """Q2 (U.cu:2097, 636): the reference's slice i integrates
    [H - (i+1) dz, H - i dz), so the top slab [H - dz, H] is never reached
    (with the 2 A margin above the top atom and dz = 2 A, the upper half of
    the top layer's potential); under compat=physical slice i covers
    [H - i dz, H - (i-1) dz) and the slices together give the top atom's whole
    (cut-off) projection."""
    from scipy.integrate import quad
    Z = np.array([38, 6, 6], np.int32)
    x = np.array([5.0, -30.0, 40.0], np.float32)
    y = np.array([5.0, -30.0, 40.0], np.float32)
    z = np.array([29.6, 0.0, 28.0], np.float32)  # the Sr atom is the top atom
    N, ps, off, dz = 16, 0.25, 33.0, 2.0
    sums = {}
    for phys in (False, True):
        s = _phys(Z, x, y, z, fp) if phys else _sim_with(Z, x, y, z, fp)
        s.apply_microscope(300.0)
        s.set_calc(ps, 0.0, N, N, 1, 1, 1, off, off)
        e = s.extents()
        H = np.float32(e[4] - e[5])
        n = int(np.ceil(H / np.float32(dz)))
        acc = 0
        for it in range(1, n + 1):
            zt = float(H - np.float32(np.float32(dz) * np.float32(it - (1 if phys else 0))))
            acc = acc + s.potential(3, zt, dz).real
        sums[phys] = acc
    X = float(np.float32(5.0) - e[1])
    c = off + np.arange(N) * ps
    ix = int(np.argmin(np.abs(c - (X + 0.7))))
    r2 = (c[ix] - X) ** 2 + (c[ix] - X) ** 2
    cut = quad(lambda t: O.vatom(38, r2 + t * t, fp), -np.sqrt(5 - r2), np.sqrt(5 - r2), limit=200)[0]
    # right-end Riemann sums of step 0.1 A
    assert sums[True][ix, ix] == pytest.approx(cut, rel=2e-2)
    assert sums[False][ix, ix] < 0.7 * cut  # the reference loses the upper half


def test_physical_band_limit_probe_norm_and_cbed(fp):
    """Q5: kmax = 2/3 of the Nyquist frequency (1 / (3 ps)); Q6: the probe
    has unit intensity; Q7: AddTDSWaves accumulates the unitary diffracted
    intensity |FFT psi|^2 / (nx ny), centred."""
    Z, x, y, z = S.silicon(1, 1, 1)
    N, ps = 64, 0.25
    sp = _phys(Z, x, y, z, fp)
    sp.apply_microscope(300.0, obj=25.0)
    sp.set_calc(ps, 0.0, N, N, 1, 1, 10, 0, 0, stem=True)
    assert sp.scalars()["kmax"] == pytest.approx(1.0 / (3 * ps), rel=1e-6)
    p = sp.probe(10, 20)
    assert np.sum(np.abs(p) ** 2) == pytest.approx(1.0, rel=1e-12)
    sp.set_wave(p)
    sp.add_tds(); sp.add_tds()
    want = 2 * np.abs(np.fft.fftshift(np.fft.fft2(p))) ** 2 / (N * N)
    assert np.allclose(sp.tds(), want, rtol=1e-12, atol=1e-18)
    assert sp.tds().sum() == pytest.approx(2.0, rel=1e-12)  # Parseval
    sr = _sim_with(Z, x, y, z, fp)
    sr.apply_microscope(300.0, obj=25.0)
    sr.set_calc(ps, 0.0, N, N, 1, 1, 10, 0, 0, stem=True)
    assert sr.scalars()["kmax"] == pytest.approx((N // 2 - 1) / (2 * N * ps), rel=1e-6)
    assert np.sum(np.abs(sr.probe(10, 20))) == pytest.approx(1.0, rel=1e-12)


def test_physical_step_uses_the_slice_below_the_top(fp):
    """Under compat=physical MultisliceStepConv's slice 1 is the
    transmission of [H - dz, H) followed by the propagator; in reference mode
    the transmission of [H - 2 dz, H - dz)."""
    Z, x, y, z = S.silicon(2, 2, 2)
    N, ps, dz = 32, 0.4, 2.0
    for phys in (False, True):
        s = _phys(Z, x, y, z, fp) if phys else _sim_with(Z, x, y, z, fp)
        s.apply_microscope(200.0)
        s.set_calc(ps, 0.0, N, N, 1, 1, 1, 2.0, 2.0)
        e = s.extents()
        H = np.float32(e[4] - e[5])
        s.init_plane(dz)
        s.step_conv(1)
        zt = float(H - np.float32(0.0 if phys else dz))
        t = s.potential(0, zt, dz)
        want = np.fft.ifft2(np.fft.fft2(t) * s.prop(dz))
        assert np.max(np.abs(s.wave() - want)) < 1e-12
        assert np.std(np.angle(want)) > 1e-3  # the slice holds atoms
        waves = waves + [want] if phys else [want]
    assert np.max(np.abs(waves[0] - waves[1])) > 1e-3
