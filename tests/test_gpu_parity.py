"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle on
the same seeded inputs.  Integer work is compared bit for bit; waves within
relative L2 1e-4 of the FP64 oracle (the tolerance the project targets after
100 slices), potentials and images within 1e-5 / 1e-4 relative."""
import numpy as np
import pytest

from cudamultislice_amd import structures as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _both(Z, x, y, z, fp):
    from cudamultislice_amd import MultisliceSimulation
    g = MultisliceSimulation()
    g.set_atoms(Z, x, y, z)
    g.GetParameterisation()
    o = O.OracleSim()
    o.set_atoms(Z, x, y, z)
    o.set_fparams(fp)
    return g, o


def _setup(g, o, kV, ps, N, sz, off=(0.0, 0.0), stem=False, **micro):
    g.ApplyMicroscopeParameters(kV, micro.get("df", 0.0), micro.get("A1", 0.0), micro.get("A1phi", 0.0),
                                micro.get("Cs", 0.0), micro.get("beta", 0.0), micro.get("delta", 0.0),
                                micro.get("obj", 0.0))
    o.apply_microscope(kV, micro.get("df", 0.0), micro.get("A1", 0.0), micro.get("A1phi", 0.0),
                       micro.get("Cs", 0.0), micro.get("beta", 0.0), micro.get("delta", 0.0), micro.get("obj", 0.0))
    if stem:
        g.STEMSetCalculationVariables(ps, micro.get("df", 0.0), N, N, 1, 1, sz)
    else:
        g.SetCalculationVariables(ps, micro.get("df", 0.0), N, N, 1, 1, sz, off[0], off[1])
    o.set_calc(ps, micro.get("df", 0.0), N, N, 1, 1, sz, off[0], off[1], stem)
    g.PreCalculateFrequencies()


def test_binning_bit_exact(fp):
    Z, x, y, z = S.srtio3(5, 5, 7)
    g, o = _both(Z, x, y, z, fp)
    dx, dy, dz = O.tds_displacements(S.SEED, 2, Z.size, 0.08)
    g.upload_binned_atoms_displaced(dx, dy, dz)
    o.bin((dx, dy, dz))
    for a, b in zip(g.binned(), o.csr()):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    # the same without displacements through UploadBinnedAtoms
    g.UploadBinnedAtoms()
    o.bin()
    for a, b in zip(g.binned(), o.csr()):
        assert np.array_equal(a, b)


def test_tds_displacements_identical_to_oracle():
    from cudamultislice_amd import tds_displacements
    for cfg in (0, 5):
        a = tds_displacements(S.SEED, cfg, 1000, 0.08)
        b = O.tds_displacements(S.SEED, cfg, 1000, 0.08)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("mode", [0, 1])
def test_potential_slice_matches_oracle(fp, mode):
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    o.bin()
    N, ps = 64, 0.2
    _setup(g, o, 300.0, ps, N, 20.0, off=(2.0, 2.0))
    g.set_potential_mode(mode)
    e = o.extents()
    H = np.float32(e[4] - e[5])
    for it, dz in ((3, 1.0), (4, 2.0), (7, 1.0)):
        zf = float(H - np.float32(np.float32(dz) * np.float32(it)))
        got = g.potential_slice(it, dz, real=True)
        want = o.potential(3 if mode == 0 else 4, zf, dz).real
        assert want.max() > 1.0
        assert np.max(np.abs(got - want)) < 2e-5 * np.max(np.abs(want))
        t = g.potential_slice(it, dz, real=False)
        tw = o.potential(0 if mode == 0 else 1, zf, dz)
        assert rel_l2(t, tw) < 1e-5


def test_fft2_device_matches_numpy():
    import torch
    from cudamultislice_amd import fft2_device
    rng = np.random.default_rng(3)
    for ny, nx, nb in ((64, 64, 3), (256, 128, 2), (512, 512, 2), (1024, 1024, 1), (2048, 2048, 1), (16, 32, 4)):
        a = (rng.standard_normal((nb, ny, nx)) + 1j * rng.standard_normal((nb, ny, nx))).astype(np.complex64)
        for inv in (False, True):
            t = torch.from_numpy(a.copy()).cuda()
            fft2_device(t.data_ptr(), nx, ny, nb, inv)
            torch.cuda.synchronize()
            want = np.fft.ifft2(a) * (nx * ny) if inv else np.fft.fft2(a)
            assert rel_l2(t.cpu().numpy(), want) < 3e-6, (ny, nx, inv)


def _si_small():
    return S.silicon(2, 2, 3)


def test_tem_exit_wave_matches_oracle(fp):
    Z, x, y, z = _si_small()
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    o.bin()
    _setup(g, o, 200.0, 0.2, 64, 20.0, off=(2.0, 2.0))
    g.InitialiseWavefunctions(1.0)
    o.init_plane(1.0)
    n = g.GetSlices(True)
    e = o.extents()
    assert n == int(np.ceil(np.float32(e[4] - e[5]) / np.float32(1.0)))
    for it in range(1, n + 1):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    ew = g.GetExitWave()
    want = o.wave()
    assert rel_l2(ew, want) < 1e-4
    inf = g.info()
    assert inf["maxposEW"] == pytest.approx(np.abs(want).max(), rel=1e-4)
    assert inf["minposEW"] == pytest.approx(np.abs(want).min(), rel=1e-4)


def test_example_tem_program_full_si_case_matches_oracle(fp, tmp_path):
    """The C++ facade driven as the reference's GUI drives its engine
    (example_tem): Si 4x4x10 cells, 256^2 over 21.72 A, 1 A slices, 200 kV,
    CTF df -500 A, Cs 1.2e7 A, 20 mrad, against the oracle."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cudamultislice_amd", "example_tem")
    assert os.path.exists(exe), "build first (make -C cudamultislice_amd/csrc)"
    Z, x, y, z = S.silicon(4, 4, 10)
    path = tmp_path / "si.xyz"
    with open(path, "w") as f:
        f.write(f"{Z.size}\n\n")
        for a, b, c in zip(x, y, z):
            f.write(f"Si {a:.6f} {b:.6f} {c:.6f}\n")
    r = subprocess.run([exe, str(path), str(tmp_path / "out"), "256", "0.084844", "200", "1", "1000"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ew = np.fromfile(tmp_path / "out_ew.bin", np.complex64).reshape(256, 256)
    img = np.fromfile(tmp_path / "out_image.bin", np.float32).reshape(256, 256)
    o = O.OracleSim()
    assert o.load_xyz_refmap(str(path)) == Z.size
    o.set_fparams(fp)
    o.bin()
    e = o.extents()
    o.apply_microscope(200.0, -500.0, 0.0, 0.0, 1.2e7, 0.0, 0.0, 20.0)
    o.set_calc(0.084844, -500.0, 256, 256, e[0] - e[1], e[2] - e[3], e[4] - e[5], 2.0, 2.0, False)
    o.init_plane(1.0)
    n = int(np.ceil(np.float32(e[4] - e[5]) / np.float32(1.0)))
    assert f"slices {n} " in r.stdout
    for it in range(1, n + 1):
        o.step_conv(it)
    assert rel_l2(ew, o.wave()) < 1e-4
    # the facade's SimulateImage runs the detector model (seed 0x4D534C, 1000
    # electrons per pixel): integer counts, so FP32 vs FP64 inputs flip a few
    # pixels by one count
    from cudamultislice_amd.detector import tables
    det, _ = O.detector_image(o.ctf_image(-500.0), 1000.0, 0x4D534C, *tables())
    assert rel_l2(img, np.abs(det)) < 1e-3


def test_example_tem_program_under_address_sanitizer(tmp_path):
    """The facade program with AddressSanitizer / LeakSanitizer on every host
    object (kernels uninstrumented): no invalid access and no leaked host
    allocation of ours across the whole GUI sequence (load, bin, physics,
    frequencies, slices, exit wave, CTF + detector image, FreeExitWave,
    destruction).  The ROCm runtime's own allocations held until exit are
    suppressed (csrc/lsan.supp)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "cudamultislice_amd", "example_tem_asan")
    if not os.path.exists(exe):
        pytest.skip("example_tem_asan not built (make -C cudamultislice_amd/csrc)")
    Z, x, y, z = S.silicon(2, 2, 4)
    path = tmp_path / "si.xyz"
    with open(path, "w") as f:
        f.write(f"{Z.size}\n\n")
        for a, b, c in zip(x, y, z):
            f.write(f"Si {a:.6f} {b:.6f} {c:.6f}\n")
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:exitcode=23"
    env["LSAN_OPTIONS"] = "suppressions=" + os.path.join(root, "cudamultislice_amd", "csrc", "lsan.supp")
    r = subprocess.run([exe, str(path), str(tmp_path / "out"), "128", "0.1", "200", "1", "1000"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.startswith("slices ")


def test_tiff_exports_hold_the_handle_state(fp, tmp_path):
    """The GUI's four TIFF exports (MW.cs:684-890) read back with PIL equal
    the exit wave, the CBED accumulator and GetImValue's image."""
    from PIL import Image
    Z, x, y, z = _si_small()
    g, _ = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    g.ApplyMicroscopeParameters(200.0, -500.0, 0, 0, 1.2e7, 0, 0, 20.0)
    g.SetCalculationVariables(0.2, -500.0, 64, 64, 1, 1, 20, 2.0, 2.0)
    g.PreCalculateFrequencies()
    g.InitialiseWavefunctions(1.0)
    g.AllocTDS()
    for it in range(1, 6):
        g.MultisliceStepConv(it)
    ew = g.GetExitWave()
    g.AddTDSWaves()
    g.set_detector(True, 9)
    g.SimulateImage(500.0)
    read = lambda p: np.asarray(Image.open(p), np.float32)  # noqa: E731
    for what, want in ((g.TIFF_EW_RE, ew.real), (g.TIFF_EW_IM, ew.imag), (g.TIFF_TDS, g.GetTDS())):
        p = str(tmp_path / f"{what}.tiff")
        g.save_tiff(what, p)
        assert np.array_equal(read(p), np.asarray(want, np.float32)), what
    p = str(tmp_path / "image.tiff")
    g.save_tiff(g.TIFF_IMAGE, p)
    img = read(p)
    noise = np.array([g.readout_noise(i) for i in range(img.size)], np.float32).reshape(img.shape)
    assert np.allclose(img - noise, np.abs(g.detector_image()), rtol=1e-6, atol=1e-3)


def test_stem_probe_steps_and_tds_match_oracle(fp):
    Z, x, y, z = _si_small()
    g, o = _both(Z, x, y, z, fp)
    N, ps, dz = 64, 0.25, 2.0
    _setup(g, o, 300.0, ps, N, 20.0, stem=True, obj=25.0, Cs=1e4, df=-50.0)
    g.AllocTDS()
    g.set_tds(S.SEED, 0.08)
    for cfg in range(2):
        g.UploadBinnedAtomsTDS()
        o.bin(O.tds_displacements(S.SEED, cfg, Z.size, 0.08))
        g.STEMInitialiseWavefunctions(20, 33)
        g.set_conventional_dz(dz)
        o.init_stem(20, 33, dz)
        assert rel_l2(g.get_wave(1), o.wave(1)) < 1e-5  # probe
        n = g.GetSlices(True)
        for it in range(1, n + 1):
            g.MultisliceStepConv(it)
            o.step_conv(it)
        assert rel_l2(g.GetExitWave(), o.wave()) < 1e-4
        g.AddTDSWaves()
        o.add_tds()
        # annulus inside the band limit (kmax lambda = 19 mrad here)
        assert o.adf(5.0, 18.0) > 1e-3 * o.adf(0.0, 1e3)
        assert g.adf(5.0, 18.0) == pytest.approx(o.adf(5.0, 18.0), rel=1e-4)
    assert rel_l2(g.GetTDS(), o.tds()) < 1e-4


def test_stem_exit_wave_readback_and_diffraction_display(fp):
    """GetSTEMExitWave (U.cu:2180-2229): the unnormalised spectrum of the exit
    wave, centred by a half-grid shift, with maxposEW / minposEW the largest
    and smallest moduli; the diffraction log display of that spectrum
    (U.cu:230-231) equals log(1 + 3000 |F(k) + conj(F(-k-1))|^2) on the
    uncentred spectrum F (the display's two index maps composed)."""
    Z, x, y, z = _si_small()
    g, o = _both(Z, x, y, z, fp)
    _setup(g, o, 300.0, 0.25, 64, 20.0, stem=True, obj=25.0)
    g.UploadBinnedAtoms()
    o.bin()
    g.STEMInitialiseWavefunctions(20, 33)
    g.set_conventional_dz(2.0)
    o.init_stem(20, 33, 2.0)
    for it in range(1, 6):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    F = O.fft2(o.wave())
    centred = np.roll(F, (32, 32), axis=(0, 1))
    got = g.GetSTEMExitWave()
    assert rel_l2(got, centred) < 1e-4
    inf = g.info()
    assert inf["maxposEW"] == pytest.approx(np.abs(centred).max(), rel=1e-4)
    assert inf["minposEW"] <= np.abs(centred).max() * 1e-3
    disp = g.diffraction_display()
    mirror = F[(-1 - np.arange(64)) % 64][:, (-1 - np.arange(64)) % 64]
    want = np.log1p(3000.0 * np.abs(F + np.conj(mirror)) ** 2)
    assert np.max(np.abs(disp - want)) < 1e-4 * np.max(want)
    assert disp.min() >= 0 and disp.max() > 1.0


def test_ctf_image_matches_oracle(fp, tmp_path):
    Z, x, y, z = _si_small()
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    o.bin()
    kw = dict(df=-500.0, Cs=1.2e7, obj=20.0, beta=0.5, delta=30.0, A1=20.0, A1phi=0.3)
    _setup(g, o, 200.0, 0.2, 64, 20.0, off=(2.0, 2.0), **kw)
    g.InitialiseWavefunctions(1.0)
    o.init_plane(1.0)
    for it in range(1, 9):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    img = g.SimulateImage(1.0)
    want = o.ctf_image(-500.0)
    assert rel_l2(img, want) < 1e-4
    dfs = np.array([-800, -400, 0, 200], np.float32)
    sw = g.ctf_sweep(dfs)
    for k, d in enumerate(dfs):
        assert rel_l2(sw[k], o.ctf_image(float(d))) < 1e-4
    # batch Cs x defocus grid, images also written as TIFF files
    import os
    from PIL import Image
    grid = g.cs_df_sweep(cs=(0.0, 2.4e7, 2), df=(-800.0, 200.0, 3), folder=str(tmp_path))
    for i, csv in enumerate((0.0, 2.4e7)):
        o.apply_microscope(200.0, -500.0, 20.0, 0.3, csv, 0.5, 30.0, 20.0)
        for j, d in enumerate((-800.0, -300.0, 200.0)):
            assert rel_l2(grid[i, j], o.ctf_image(d)) < 1e-4, (csv, d)
            f = tmp_path / f"Cs_{csv:g}_df_{d:g}.tiff"
            assert np.array_equal(np.asarray(Image.open(f), np.float32), grid[i, j])
    assert len(os.listdir(tmp_path)) == 6


def test_detector_model_matches_oracle(fp):
    """The K2 detector model of SimulateImage on the engine's own CTF image,
    against the oracle's FP64 model with the same seed and curves; readout
    noise draw by draw."""
    from cudamultislice_amd.detector import tables
    Z, x, y, z = _si_small()
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    kw = dict(df=-500.0, Cs=1.2e7, obj=20.0)
    _setup(g, o, 200.0, 0.2, 128, 20.0, off=(2.0, 2.0), **kw)
    g.InitialiseWavefunctions(1.0)
    for it in range(1, 9):
        g.MultisliceStepConv(it)
    dqe, ntf = tables()
    for seed, dose in ((5, 300.0), (0x4D534C, 20.0)):
        g.set_detector(True, seed)
        I = g.SimulateImage(dose).astype(np.float64)  # the CTF intensity the model starts from
        det = g.detector_image()
        want, st = O.detector_image(I, dose, seed, dqe, ntf)
        # integer counts: FP32 vs FP64 rounding of dose * a flips rare pixels
        assert rel_l2(det, want) < 1e-4, (seed, dose)
        inf = g.info()
        assert inf["maxposIm"] == pytest.approx(st[0], rel=1e-3)
        assert inf["imagecontrast"] == pytest.approx(st[2], rel=1e-3)
    for k in (0, 1, 77):
        assert g.readout_noise(k) == O.readout_noise(0x4D534C, k)
    g.set_detector(False)
    g.SimulateImage(1.0)
    import ctypes as C
    from cudamultislice_amd._lib import lib
    assert lib().msl_get_detector_image(g._h, None) != 0  # model off: no detector image


def test_fd_steps_match_oracle(fp):
    Z, x, y, z = _si_small()
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    o.bin()
    _setup(g, o, 200.0, 0.2, 64, 20.0)
    g.InitialiseWavefunctions()
    o.init_plane_fd()
    e = o.extents()
    # start inside the specimen so the potential is non-zero
    first = int(0.5 * (e[4] - e[5]) / o.scalars()["dzFD"])
    for it in range(first, first + 25):
        g.MultisliceStep(it)
        o.step_fd(it)
    assert rel_l2(g.get_wave(0), o.wave(0)) < 1e-4
    assert np.max(np.abs(o.wave(0) - 1.0)) > 1e-3  # the potential acted


def _raster(N, n, step):
    xs = (np.arange(n) * step) % N
    px, py = np.meshgrid(xs, xs)
    return px.ravel().astype(np.int32), py.ravel().astype(np.int32)


def test_rs_steps_match_oracle(fp):
    """MultisliceStepRS: the real-space series and its wave bookkeeping."""
    Z, x, y, z = _si_small()
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    o.bin()
    _setup(g, o, 200.0, 0.2, 64, 20.0, off=(2.0, 2.0))
    g.InitialiseWavefunctions(1.0)
    o.init_plane(1.0)
    for it in range(1, 6):
        g.MultisliceStepRS(it)
        o.step_rs(it)
        assert rel_l2(g.get_wave(0), o.wave(0)) < 1e-5, it
        if it > 1:
            assert rel_l2(g.get_wave(1), o.wave(1)) < 1e-5, it
    assert np.std(np.abs(o.wave(0))) > 1e-3  # the potential acted


def test_batched_stem_matches_oracle(fp):
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, o = _both(Z, x, y, z, fp)
    N, ps, dz = 64, 0.2, 2.0
    _setup(g, o, 300.0, ps, N, 20.0, stem=True, obj=25.0)
    px, py = _raster(N, 3, 17)
    # kmax lambda = 24 mrad at this sampling: annulus 10-200 mrad keeps 10-24
    r = g.stem_batched(px, py, dz, nconfig=2, seed=S.SEED, sigma=0.08, adf=(10.0, 200.0), exit_waves=True)
    e = o.extents()
    ns = int(np.ceil(np.float32(e[4] - e[5]) / np.float32(dz)))
    adf = np.zeros(px.size)
    for cfg in range(2):
        o.bin(O.tds_displacements(S.SEED, cfg, Z.size, 0.08))
        ro = o.stem_batch(px, py, dz, ns, 10.0, 200.0, want_exit=True)
        adf += ro["adf"] / 2
        for p in range(px.size):
            assert rel_l2(r["exit"][cfg, p], ro["exit"][p]) < 1e-4
    assert adf.min() > 1e-4
    assert np.max(np.abs(r["adf"] - adf) / adf) < 1e-4


def test_batched_fused_detector_equals_exit_path_and_per_probe(fp):
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, o = _both(Z, x, y, z, fp)
    N, ps, dz = 128, 0.15, 2.0
    _setup(g, o, 300.0, ps, N, 20.0, stem=True, obj=25.0)
    px, py = _raster(N, 4, 29)
    # kmax lambda = 32 mrad at this sampling
    fused = g.stem_batched(px, py, dz, nconfig=1, adf=(15.0, 200.0), pacbed=True)
    slow = g.stem_batched(px, py, dz, nconfig=1, adf=(15.0, 200.0), exit_waves=True)
    assert fused["adf"].min() > 1e-4
    assert np.max(np.abs(fused["adf"] - slow["adf"]) / slow["adf"]) < 1e-5
    # per-probe path through the reference-shaped API
    g.UploadBinnedAtoms()
    n = int(np.ceil((g.info()["MaximumZ"] - g.info()["MinimumZ"]) / dz))
    for p in (0, 5, 15):
        g.STEMInitialiseWavefunctions(int(px[p]), int(py[p]))
        g.set_conventional_dz(dz)
        for it in range(1, n + 1):
            g.MultisliceStepConv(it)
        assert g.adf(15.0, 200.0) == pytest.approx(float(fused["adf"][p]), rel=1e-4)
    # PACBED: mean |Psi|^2 sums to the mean transmitted intensity within the band
    assert fused["pacbed"].sum() > 0


def test_every_pass_variant_matches_oracle_in_batched_stem(fp):
    """Each row / column layout and FFT plan the autotuner may select, in a
    band-pruned batched run (probe, slices, fused detector) against the
    oracle, and each agrees with variant 0 on the tuner's own benchmark."""
    import ctypes as C
    from cudamultislice_amd._lib import lib
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, o = _both(Z, x, y, z, fp)
    N, ps, dz = 64, 0.2, 2.0
    _setup(g, o, 300.0, ps, N, 20.0, stem=True, obj=25.0)
    assert g.info()["band_tiles"] > 0
    px, py = _raster(N, 3, 17)
    e = o.extents()
    ns = int(np.ceil(np.float32(e[4] - e[5]) / np.float32(dz)))
    o.bin()
    want = o.stem_batch(px, py, dz, ns, 10.0, 200.0)["adf"]
    L = lib()
    nrow, ncol = 5, 8
    try:
        for rv in range(nrow):
            for cv in range(ncol):
                assert L.msl_set_pass_variant(0, N, rv) == 0 and L.msl_set_pass_variant(1, N, cv) == 0
                r = g.stem_batched(px, py, dz, nconfig=1, adf=(10.0, 200.0))
                assert np.max(np.abs(r["adf"] - want) / want) < 1e-4, (rv, cv)
        for n in (16, 32, 64, 128, 256, 512, 1024, 2048):
            for kind, nv in ((0, nrow), (1, ncol)):
                for v in range(1, nv):
                    ms, md = C.c_float(0), C.c_float(1)
                    nb = 16 if n <= 1024 else 2  # row variant 4 needs two waves per line
                    assert L.msl_bench_pass(kind, v, n, nb, 1, C.byref(ms), C.byref(md)) == 0
                    assert md.value < 1e-5, (n, kind, v, md.value)
    finally:
        L.msl_set_pass_variant(0, N, 1)
        L.msl_set_pass_variant(1, N, 0)


def test_row_variant_4_agrees_with_variant_2_and_matches_the_oracle(fp):
    """Row variant 4 (a workgroup per row looping over the waves of a batch,
    the shared transmission row in registers, the next wave's row requested
    before the transforms) runs variant 2's transforms and products in the
    same order; the compiler may contract a different set of multiply-adds
    in the two kernels, so they agree to FP32 rounding (1e-6), on a 512^2
    batch wide enough for it (32 probes, 8 lines of 32 threads per
    workgroup), and its exit waves are within 1e-4 of the oracle."""
    Z, x, y, z = S.srtio3(3, 3, 3)
    g, o = _both(Z, x, y, z, fp)
    N, ps, dz = 512, 0.04, 2.0
    _setup(g, o, 300.0, ps, N, 20.0, stem=True, obj=25.0)
    assert g.info()["band_tiles"] > 0
    rng = np.random.default_rng(11)
    px = rng.integers(0, N, 32).astype(np.int32)
    py = rng.integers(0, N, 32).astype(np.int32)
    from cudamultislice_amd._lib import lib
    L = lib()
    runs = {}
    try:
        for rv in (2, 4):
            assert L.msl_set_pass_variant(0, N, rv) == 0
            runs[rv] = g.stem_batched(px, py, dz, nconfig=1, adf=(10.0, 200.0), batch=32, exit_waves=True)
    finally:
        L.msl_set_pass_variant(0, N, 4)
    assert np.max(np.abs(runs[4]["adf"] - runs[2]["adf"]) / runs[2]["adf"]) < 1e-6
    assert rel_l2(runs[4]["exit"], runs[2]["exit"]) < 1e-6
    e = o.extents()
    ns = int(np.ceil(np.float32(e[4] - e[5]) / np.float32(dz)))
    o.bin()
    sel = np.array([0, 17])
    ro = o.stem_batch(px[sel], py[sel], dz, ns, 10.0, 200.0, want_exit=True)
    for k, p in enumerate(sel):
        assert rel_l2(runs[4]["exit"][0, p], ro["exit"][k]) < 1e-4, p


def test_batch_width_autotuner_keeps_results(fp):
    """msl_autotune_batch times the candidate widths and keeps the widest
    within 2 % of the fastest; a batch = 0 run then uses it and gives the
    bits of a single-launch run (each probe's arithmetic does not depend on
    the chunk it is in)."""
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, o = _both(Z, x, y, z, fp)
    _setup(g, o, 300.0, 0.2, 64, 20.0, stem=True, obj=25.0)
    rng = np.random.default_rng(5)
    px = rng.integers(0, 64, 300).astype(np.int32)
    py = rng.integers(0, 64, 300).astype(np.int32)
    best, cost = g.autotune_batch(px, py, 2.0, nslices=4)
    assert sorted(cost) == [64, 128, 256, 300] and best in cost
    assert all(c > 0 for c in cost.values())
    fastest = np.float32(min(cost.values()))
    assert best == max(w for w, c in cost.items() if np.float32(c) <= np.float32(1.02) * fastest)
    a = g.stem_batched(px, py, 2.0, nconfig=1, adf=(10.0, 200.0))["adf"]
    b = g.stem_batched(px, py, 2.0, nconfig=1, adf=(10.0, 200.0), batch=300)["adf"]
    assert np.array_equal(a, b)


def test_batched_run_is_deterministic(fp):
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, _ = _both(Z, x, y, z, fp)
    g.ApplyMicroscopeParameters(300.0, 0, 0, 0, 0, 0, 0, 25.0)
    g.STEMSetCalculationVariables(0.2, 0.0, 64, 64, 1, 1, 20)
    g.PreCalculateFrequencies()
    px, py = _raster(64, 4, 13)
    a = g.stem_batched(px, py, 2.0, nconfig=2, sigma=0.08, pacbed=True)
    b = g.stem_batched(px, py, 2.0, nconfig=2, sigma=0.08, pacbed=True)
    assert np.array_equal(a["adf"], b["adf"]) and np.array_equal(a["pacbed"], b["pacbed"])


def test_batched_graph_replays_equal_eager_launches(fp):
    """The captured chunk (probe synthesis, slice loop, fused detector) is
    replayed for every chunk and configuration, with eager binning and
    potential synthesis between replays; each replay must give exactly the
    eager result (ragged last chunk included: 40 probes in chunks of 16)."""
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, _ = _both(Z, x, y, z, fp)
    g.ApplyMicroscopeParameters(300.0, 0, 0, 0, 0, 0, 0, 25.0)
    g.STEMSetCalculationVariables(0.2, 0.0, 64, 64, 1, 1, 20)
    g.PreCalculateFrequencies()
    px, py = _raster(64, 7, 9)
    px, py = px[:40], py[:40]
    for nconf in (1, 3):
        eager = g.stem_batched(px, py, 2.0, nconfig=nconf, sigma=0.08, adf=(10.0, 200.0), batch=16, pacbed=True)
        graph = g.stem_batched(px, py, 2.0, nconfig=nconf, sigma=0.08, adf=(10.0, 200.0), batch=16, pacbed=True,
                               use_graph=True)
        assert eager["adf"].min() > 0
        assert np.array_equal(eager["adf"], graph["adf"])
        assert np.array_equal(eager["pacbed"], graph["pacbed"])
    # one configuration per call: each config's result equals its own eager run
    for cfg_seed in (S.SEED, S.SEED + 7):
        a = g.stem_batched(px, py, 2.0, nconfig=1, seed=cfg_seed, sigma=0.08, adf=(10.0, 200.0), batch=16)
        b = g.stem_batched(px, py, 2.0, nconfig=1, seed=cfg_seed, sigma=0.08, adf=(10.0, 200.0), batch=16,
                           use_graph=True)
        assert np.array_equal(a["adf"], b["adf"])


def _serve_mix():
    from cudamultislice_amd.serve import CONVENTIONAL, FD, Job
    si = S.srtio3(2, 2, 2)
    big = S.srtio3(3, 3, 3)
    ra = S.random_atoms(60, (8.0, 8.0, 6.0), seed=S.SEED + 3)
    rb = S.random_atoms(150, (20.0, 20.0, 8.0), seed=S.SEED + 4)
    pr = lambda *v: np.array(v, np.int32)  # noqa: E731
    return [
        Job(64, *si, nslices=6),                                           # plane wave, conventional
        Job(64, *ra, nslices=9, posx=pr(3, 40, 63), posy=pr(7, 22, 50)),    # probes, conventional
        Job(128, *rb, nslices=5, posx=pr(0, 100), posy=pr(64, 9)),         # other grid size
        Job(128, *big, nslices=3),                                         # short plane wave
        Job(64, *si, nslices=10, method=FD),                               # plane wave, finite differences
        Job(64, *ra, nslices=12, method=FD, posx=pr(30, 5), posy=pr(31, 60)),  # probes, finite differences
    ]


def _serve_oracle(job, fp, kV, ps, dz, ap, adf):
    from serve_oracle import job_reference
    return job_reference(job, fp, kV, ps, dz, ap, adf)


def test_serving_mix_of_uneven_jobs_matches_oracle(fp):
    """Uneven jobs (grid 64 / 128, conventional / finite differences, plane
    wave / probes, 3-12 slices) in shared batches, against the oracle run job
    by job; a tiny memory budget (one wave per batch) gives the same bits."""
    from cudamultislice_amd.serve import serve
    kV, ps, dz, ap, adf = 300.0, 0.2, 2.0, 25.0, (10.0, 200.0)
    jobs = _serve_mix()
    st = serve(jobs, kV=kV, ps=ps, dz=dz, aperture=ap, adf=adf)
    assert st["batches"] == 4 and st["waves"] == 10
    for j in jobs:
        want = _serve_oracle(j, fp, kV, ps, dz, ap, adf)
        if "exit" in want:
            assert rel_l2(j.result["exit"], want["exit"]) < 1e-4, j.n
        else:
            assert want["adf"].min() > 1e-5
            assert np.max(np.abs(j.result["adf"] - want["adf"]) / want["adf"]) < 1e-4, (j.n, j.method)
    again = _serve_mix()
    st2 = serve(again, kV=kV, ps=ps, dz=dz, aperture=ap, adf=adf, budget_bytes=64 * 64 * 8 * 6)
    assert st2["batches"] > st["batches"]
    for a, b in zip(jobs, again):
        for k in a.result:
            assert np.array_equal(a.result[k], b.result[k]), k


def test_grouped_sizes_share_launches_and_match_oracle(fp):
    """Conventional jobs of grid 256, 512, 1024 and 2048 in grouped batches
    (one potential, row and column launch per slice for every size) agree
    with the per-size launches to rounding (same plans and tiles; the
    compiler may fuse multiply-adds differently in the two kernels) and match
    the oracle job by job (2048 against the per-size run only); a 64-point
    job beside them runs per size."""
    from cudamultislice_amd.serve import Job, serve
    kV, ps, dz, ap, adf = 300.0, 0.2, 2.0, 25.0, (10.0, 200.0)
    si = S.srtio3(2, 2, 2)
    ra = S.random_atoms(60, (8.0, 8.0, 6.0), seed=S.SEED + 5)
    rb = S.random_atoms(120, (30.0, 30.0, 6.0), seed=S.SEED + 6)
    pr = lambda *v: np.array(v, np.int32)  # noqa: E731

    def mix():
        return [
            Job(256, *ra, nslices=4, posx=pr(3, 200, 255), posy=pr(7, 100, 0)),
            Job(256, *si, nslices=6),
            Job(512, *rb, nslices=3, posx=pr(0, 300), posy=pr(64, 9)),
            Job(1024, *si, nslices=2),
            Job(2048, *ra, nslices=2),
            Job(64, *si, nslices=5),
        ]
    grouped, per_size = mix(), mix()
    sg = serve(grouped, kV=kV, ps=ps, dz=dz, aperture=ap, adf=adf, group_sizes=True)
    sp = serve(per_size, kV=kV, ps=ps, dz=dz, aperture=ap, adf=adf, group_sizes=False)
    assert sg["launches"] < sp["launches"] and sg["wave_slices"] == sp["wave_slices"]
    for a, b in zip(grouped, per_size):
        for k in a.result:
            if k == "adf":
                assert np.max(np.abs(a.result[k] - b.result[k]) / b.result[k]) < 1e-5, (a.n, k)
            else:
                assert rel_l2(a.result[k], b.result[k]) < 1e-5, (a.n, k)
    for j in grouped:
        if j.n == 2048:
            assert np.isfinite(j.result["exit"]).all()
            continue
        want = _serve_oracle(j, fp, kV, ps, dz, ap, adf)
        if "exit" in want:
            assert rel_l2(j.result["exit"], want["exit"]) < 1e-4, j.n
        else:
            assert np.max(np.abs(j.result["adf"] - want["adf"]) / want["adf"]) < 1e-4, j.n


def test_diff_forward_and_gradient_match_fp64_adjoint():
    import torch
    from cudamultislice_amd.diff import DiffEngine, MultisliceIntensity
    N, ns, B, K, ps, kV, dz = 64, 12, 3, 5, 0.25, 300.0, 2.0
    rng = np.random.default_rng(11)
    V = (rng.random((ns, N, N)) * 30.0).astype(np.float32)
    psi0 = (rng.standard_normal((B, N, N)) + 1j * rng.standard_normal((B, N, N))).astype(np.complex64)
    target = rng.random((B, N, N)).astype(np.float32) * 2
    eng = DiffEngine(N, N, ps, kV, dz, ns, B, checkpoint_every=K)
    Vt = torch.from_numpy(V).cuda().requires_grad_(True)
    pt = torch.from_numpy(psi0).cuda()
    I = MultisliceIntensity.apply(Vt, pt, eng)
    loss = ((I - torch.from_numpy(target).cuda()) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    # oracle: same physics from a handle with matching scalars
    o = O.OracleSim()
    o.set_atoms(np.array([6, 6], np.int32), np.array([-60, 200], np.float32), np.array([-60, 200], np.float32),
                np.array([0, 20], np.float32))
    o.set_fparams(np.ones((103, 12), np.float32))
    o.apply_microscope(kV)
    o.set_calc(ps, 0.0, N, N, 1, 1, 22, 0, 0)
    Io, _ = o.diff(V, psi0.astype(np.complex128), None, dz)
    assert rel_l2(I.detach().cpu().numpy(), Io) < 1e-4
    _, go = o.diff(V, psi0.astype(np.complex128), 2 * (Io - target), dz)
    g = Vt.grad.cpu().numpy()
    assert rel_l2(g, go) < 1e-3
    # central finite differences of the FP64 loss at 32 random voxels
    pc = psi0.astype(np.complex128)
    idx = [tuple(int(rng.integers(0, n)) for n in V.shape) for _ in range(32)]
    fd, gv = [], []
    for ix in idx:
        Vp, Vm = V.copy(), V.copy()
        Vp[ix] += np.float32(1e-2)
        Vm[ix] -= np.float32(1e-2)
        Lp = float(((o.diff(Vp, pc, None, dz)[0] - target) ** 2).sum())
        Lm = float(((o.diff(Vm, pc, None, dz)[0] - target) ** 2).sum())
        fd.append((Lp - Lm) / (float(Vp[ix]) - float(Vm[ix])))
        gv.append(g[ix])
    fd, gv = np.array(fd), np.array(gv)
    assert np.abs(fd).max() > 0
    assert np.linalg.norm(gv - fd) / np.linalg.norm(fd) < 1e-3


def test_diff_checkpoint_interval_does_not_change_gradient():
    import torch
    from cudamultislice_amd.diff import DiffEngine
    N, ns, B = 64, 11, 2
    rng = np.random.default_rng(12)
    V = torch.from_numpy((rng.random((ns, N, N)) * 20).astype(np.float32)).cuda()
    p = torch.from_numpy((rng.standard_normal((B, N, N)) + 0j).astype(np.complex64)).cuda()
    gI = torch.from_numpy(rng.standard_normal((B, N, N)).astype(np.float32)).cuda()
    grads = []
    for K in (1, 3, 11):
        e = DiffEngine(N, N, 0.2, 200.0, 1.5, ns, B, checkpoint_every=K)
        e.forward(V, p)
        grads.append(e.backward(gI).cpu().numpy())
        if K == 3:  # the same interval twice: bit-identical
            e.forward(V, p)
            assert np.array_equal(e.backward(gI).cpu().numpy(), grads[-1])
    # other intervals move the real-space waves through other passes (rounding only)
    assert rel_l2(grads[1], grads[0]) < 1e-6 and rel_l2(grads[2], grads[0]) < 1e-6
    # the automatic interval keeps every wave of a problem this small: the
    # workspace and the bits of checkpoint_every = nslices
    e = DiffEngine(N, N, 0.2, 200.0, 1.5, ns, B, checkpoint_every=-1)
    assert e.info()["workspace_bytes"] == DiffEngine(N, N, 0.2, 200.0, 1.5, ns, B, 0).info()["workspace_bytes"]
    e.forward(V, p)
    assert np.array_equal(e.backward(gI).cpu().numpy(), grads[2])


def test_diff_forward_only_matches_forward_and_blocks_backward():
    """The inference forward runs the same passes minus the stores kept for a
    backward: the same intensities; a backward after it is refused, and the
    next keeping forward restores the same gradient."""
    import torch
    from cudamultislice_amd._lib import MslError
    from cudamultislice_amd.diff import DiffEngine
    N, ns, B = 64, 7, 2
    rng = np.random.default_rng(13)
    V = torch.from_numpy((rng.random((ns, N, N)) * 20).astype(np.float32)).cuda()
    p = torch.from_numpy((rng.standard_normal((B, N, N)) + 0j).astype(np.complex64)).cuda()
    gI = torch.from_numpy(rng.standard_normal((B, N, N)).astype(np.float32)).cuda()
    for K in (0, 3):
        e = DiffEngine(N, N, 0.2, 200.0, 1.5, ns, B, checkpoint_every=K)
        I_keep = e.forward(V, p).cpu().numpy()
        g = e.backward(gI).cpu().numpy()
        I_only = e.forward(V, p, keep=False).cpu().numpy()
        # the keeping passes are separate kernels (they also store w_i): rounding only
        assert rel_l2(I_only, I_keep) < 1e-6
        with pytest.raises(MslError) as err:
            e.backward(gI)
        assert err.value.code == 2
        e.forward(V, p)
        assert np.array_equal(e.backward(gI).cpu().numpy(), g)
        e.close()


def test_invalid_calls_return_status_codes(fp):
    """Error behaviour of the C-ABI (the reference has none: void methods,
    ignored CUDA statuses, exit(1)): every misuse below returns a status and a
    message, leaves the handle usable, and launches nothing."""
    from cudamultislice_amd import MultisliceSimulation
    from cudamultislice_amd._lib import MslError
    from cudamultislice_amd.serve import Job, serve
    ARG, STATE, UNSUPPORTED = 1, 2, 5

    def code(f, *a, **k):
        with pytest.raises(MslError) as e:
            f(*a, **k)
        return e.value.code

    Z, x, y, z = S.srtio3(2, 2, 2)
    with MultisliceSimulation() as g:
        assert code(g.set_atoms, Z[:0], x[:0], y[:0], z[:0]) == ARG          # empty structure
        assert code(g.MultisliceStepConv, 1) == STATE                          # step before set-up
        g.set_atoms(Z, x, y, z)
        g.GetParameterisation()
        g.UploadBinnedAtoms()
        g.ApplyMicroscopeParameters(300.0)
        assert code(g.SetCalculationVariables, 0.2, 0.0, 10000, 64, 1, 1, 8.0) == UNSUPPORTED  # > 8192
        assert code(g.SetCalculationVariables, 0.2, 0.0, 64, 1, 1, 1, 8.0) == UNSUPPORTED      # < 2
        g.SetCalculationVariables(0.2, 0.0, 64, 64, 1, 1, 8.0)
        g.PreCalculateFrequencies()
        assert code(g.MultisliceStepConv, 1) == STATE                          # before InitialiseWavefunctions
        g.InitialiseWavefunctions(2.0)
        g.MultisliceStepConv(1)                                                # the handle still works
        assert code(g.stem_batched, np.zeros(0, np.int32), np.zeros(0, np.int32), 2.0) == ARG
    bad = Job(64, Z, x, y, z, nslices=2, posx=np.array([64], np.int32), posy=np.array([0], np.int32))
    assert code(serve, [bad]) == ARG                                            # probe outside the grid
    assert code(serve, [Job(100, Z, x, y, z, nslices=2)]) == UNSUPPORTED        # grid not 2^k


def test_handles_release_their_device_memory(fp):
    """Every handle frees what it allocated (the reference leaks its plans
    and frequency arrays on each PreCalculateFrequencies): device memory in
    use returns to its level after repeated create / run / destroy cycles of
    the simulation, batched STEM, differentiable and serving paths."""
    import torch
    from cudamultislice_amd import MultisliceSimulation
    from cudamultislice_amd.diff import DiffEngine, MultisliceIntensity
    from cudamultislice_amd.serve import Job, serve
    Z, x, y, z = S.srtio3(2, 2, 3)
    px, py = np.array([3, 40], np.int32), np.array([7, 50], np.int32)

    def cycle():
        with MultisliceSimulation() as g:
            g.set_atoms(Z, x, y, z)
            g.GetParameterisation()
            g.UploadBinnedAtoms()
            g.ApplyMicroscopeParameters(300.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 20.0)
            g.SetCalculationVariables(0.2, 0.0, 128, 128, 1, 1, 12.0, 0.0, 0.0)
            g.PreCalculateFrequencies()
            g.PreCalculateFrequencies()
            g.InitialiseWavefunctions(2.0)
            for i in range(1, 4):
                g.MultisliceStepConv(i)
            g.GetExitWave()
            g.STEMSetCalculationVariables(0.2, 0.0, 64, 64, 1, 1, 12.0)
            g.PreCalculateFrequencies()
            g.stem_batched(px, py, 2.0, nconfig=1, adf=(10.0, 200.0))
        eng = DiffEngine(64, 64, 0.25, 300.0, 2.0, 4, 2, checkpoint_every=2)
        V = torch.rand(4, 64, 64, device="cuda", requires_grad=True)
        psi = torch.ones(2, 64, 64, dtype=torch.complex64, device="cuda")
        MultisliceIntensity.apply(V, psi, eng).sum().backward()
        eng.close()
        serve([Job(64, Z, x, y, z, nslices=3, posx=px, posy=py), Job(64, Z, x, y, z, nslices=2)],
              kV=300.0, ps=0.2, dz=2.0, aperture=25.0, adf=(10.0, 200.0))
        torch.cuda.synchronize()

    def used():
        torch.cuda.synchronize()
        free, total = torch.cuda.mem_get_info()
        return total - free - torch.cuda.memory_reserved()

    cycle()  # first use: runtime code objects and pools
    base = used()
    for _ in range(3):
        cycle()
    assert used() - base < 8 << 20, (used() - base)


def test_long_run_100_slices_within_1e4(fp):
    """Exit wave after >= 100 slices against the FP64 oracle."""
    Z, x, y, z = S.silicon(3, 3, 18)
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    o.bin()
    _setup(g, o, 200.0, 0.16, 128, 110.0, off=(2.0, 2.0))
    g.InitialiseWavefunctions(1.0)
    o.init_plane(1.0)
    n = g.GetSlices(True)
    assert n >= 100
    for it in range(1, n + 1):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    assert rel_l2(g.GetExitWave(), o.wave()) < 1e-4


# ---------------------------------------------------------------- compat=physical
@pytest.mark.parametrize("physical", [False, True])
def test_compat_modes_match_the_oracle_quirk_by_quirk(fp, physical):
    """Both compat modes against the oracle in the same mode, quirk by quirk:
    Q2/Q3 the integrated potential of 2 A slices (slice position, slab
    window), Q5 the band limit, Q6 the probe norm, Q7 the CBED accumulator,
    and a TEM exit wave through all slices at dz = 2 A."""
    from cudamultislice_amd import MultisliceSimulation
    Z, x, y, z = S.srtio3(3, 3, 4)
    g = MultisliceSimulation(physical=physical)
    g.set_atoms(Z, x, y, z)
    g.GetParameterisation()
    o = O.OracleSim(physical=physical)
    o.set_atoms(Z, x, y, z)
    o.set_fparams(fp)
    g.UploadBinnedAtoms()
    o.bin()
    N, ps, dz = 64, 0.2, 2.0
    _setup(g, o, 300.0, ps, N, 20.0, off=(2.0, 2.0))
    # Q5
    inf = g.info()
    want_kmax = 1.0 / (3 * ps) if physical else (N // 2 - 1) / (2 * N * ps)
    assert inf["kmax"] == pytest.approx(want_kmax, rel=1e-6)
    assert o.scalars()["kmax"] == pytest.approx(inf["kmax"], rel=1e-6)
    assert (inf["band_tiles"] > 0) == (not physical)
    # Q2 / Q3: slice it of the engine is the oracle's potential at conv_z
    e = o.extents()
    H = np.float32(e[4] - e[5])
    for it in (1, 3, 5):
        zf = float(H - np.float32(np.float32(dz) * np.float32(it - (1 if physical else 0))))
        got = g.potential_slice(it, dz, real=True)
        want = o.potential(3, zf, dz).real
        assert want.max() > 1.0
        assert np.max(np.abs(got - want)) < 2e-5 * np.max(np.abs(want)), it
    # exit wave through every slice
    g.InitialiseWavefunctions(dz)
    o.init_plane(dz)
    n = g.GetSlices(True)
    for it in range(1, n + 1):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    assert rel_l2(g.GetExitWave(), o.wave()) < 1e-4
    # Q6 / Q7 on a STEM wave
    g.ApplyMicroscopeParameters(300.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 25.0)
    g.STEMSetCalculationVariables(ps, 0.0, N, N, 1, 1, 20.0)
    g.PreCalculateFrequencies()
    o.apply_microscope(300.0, obj=25.0)
    o.set_calc(ps, 0.0, N, N, 1, 1, 20.0, 0, 0, True)
    g.AllocTDS()
    g.STEMInitialiseWavefunctions(20, 33)
    g.set_conventional_dz(dz)
    o.init_stem(20, 33, dz)
    p = g.get_wave(1)
    norm = np.sum(np.abs(p.astype(np.complex128)) ** 2) if physical else np.sum(np.abs(p.astype(np.complex128)))
    assert norm == pytest.approx(1.0, rel=1e-5)
    assert rel_l2(p, o.wave(1)) < 1e-5
    for it in range(1, 6):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    g.AddTDSWaves()
    o.add_tds()
    assert rel_l2(g.GetTDS(), o.tds()) < 1e-4
    # batched STEM in the same mode
    px, py = _raster(N, 3, 17)
    r = g.stem_batched(px, py, dz, nconfig=1, adf=(10.0, 200.0))
    ns = int(np.ceil(H / np.float32(dz)))
    want = o.stem_batch(px, py, dz, ns, 10.0, 200.0)["adf"]
    assert want.min() > 0
    assert np.max(np.abs(r["adf"] - want) / want) < 1e-4
    g.close()


def test_potential_tile_shapes_give_the_same_bits_and_the_tuner_keeps_one(fp):
    """The potential kernel's pixel tiles (16x16, 32x8, 64x4, 8x32) stage
    different bin windows but sum each pixel's atoms in the same order: the
    same bits for every shape, equal to the oracle within 2e-5; the tuner
    keeps the fastest shape."""
    Z, x, y, z = S.srtio3(3, 3, 4)
    g, o = _both(Z, x, y, z, fp)
    g.UploadBinnedAtoms()
    o.bin()
    _setup(g, o, 300.0, 0.15, 96, 20.0, off=(2.0, 2.0))
    outs = []
    for v in range(4):
        g.set_potential_tile(v)
        outs.append(np.stack([g.potential_slice(it, 2.0, real=True) for it in (1, 3, 5)]))
    for v in range(1, 4):
        assert np.array_equal(outs[v], outs[0]), v
    e = o.extents()
    H = np.float32(e[4] - e[5])
    want = o.potential(3, float(H - np.float32(6.0)), 2.0).real
    assert np.max(np.abs(outs[0][1] - want)) < 2e-5 * np.max(want)
    best, ms = g.autotune_potential(2.0, nslices=4)
    assert best in range(4) and ms[best] > 0 and all(m > 0 for m in ms)
    assert np.array_equal(np.stack([g.potential_slice(it, 2.0, real=True) for it in (1, 3, 5)]), outs[0])
