"""GPU parity on grids whose sides are not powers of two, which the
reference's cuFFT plan accepts (U.cu:1961) and its GUI offers (1002x668,
2004x1336, 4008x2672; MW.cs:147-149): the engine's mixed-radix transforms
against the oracle's direct DFT, through every single-handle path
(conventional and finite-difference steps, STEM probe and ADF, CBED
accumulator, CTF image, exit-wave readbacks)."""
import numpy as np
import pytest

from cudamultislice_amd import structures as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _pair(Z, x, y, z, fp):
    from cudamultislice_amd import MultisliceSimulation
    g = MultisliceSimulation()
    g.set_atoms(Z, x, y, z)
    g.GetParameterisation()
    g.UploadBinnedAtoms()
    o = O.OracleSim()
    o.set_atoms(Z, x, y, z)
    o.set_fparams(fp)
    o.bin()
    return g, o


@pytest.mark.parametrize("nx,ny", [(96, 80), (90, 150), (1002, 668)])
def test_tem_exit_wave_and_ctf_on_generic_grids(fp, nx, ny):
    """Plane-wave conventional steps (U.cu:2085-2124) and the CTF image
    (U.cu:1044-1080) on rectangular grids with factors 3, 5 and 167."""
    Z, x, y, z = S.silicon(2, 2, 3)
    g, o = _pair(Z, x, y, z, fp)
    ps = 0.2 if nx < 500 else 0.03
    df, Cs, ap = -300.0, 1.2e7, 20.0
    g.ApplyMicroscopeParameters(200.0, df, 0.0, 0.0, Cs, 0.0, 0.0, ap)
    o.apply_microscope(200.0, df, 0.0, 0.0, Cs, 0.0, 0.0, ap)
    g.SetCalculationVariables(ps, df, nx, ny, 1, 1, 20.0, 2.0, 2.0)
    g.PreCalculateFrequencies()
    o.set_calc(ps, df, nx, ny, 1, 1, 20.0, 2.0, 2.0)
    assert g.info()["band_tiles"] == 0
    g.InitialiseWavefunctions(1.0)
    o.init_plane(1.0)
    n = 6 if nx < 500 else 3
    for it in range(1, n + 1):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    ew, want = g.GetExitWave(), o.wave()
    assert ew.shape == (ny, nx)
    assert np.max(np.abs(want - 1.0)) > 1e-3  # the specimen acted
    assert rel_l2(ew, want) < 1e-4
    assert rel_l2(g.SimulateImage(1.0), o.ctf_image(df)) < 1e-4


def test_stem_probe_steps_adf_and_cbed_on_a_generic_grid(fp):
    """STEM on a 90x72 grid: the probe (U.cu:2672-2724), conventional steps,
    the ADF annulus, the CBED accumulator (U.cu:2734-2766) and
    GetSTEMExitWave's centred spectrum (U.cu:2180-2229)."""
    Z, x, y, z = S.srtio3(2, 2, 3)
    g, o = _pair(Z, x, y, z, fp)
    nx, ny, ps, dz = 90, 72, 0.25, 2.0
    g.ApplyMicroscopeParameters(300.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 25.0)
    g.STEMSetCalculationVariables(ps, 0.0, nx, ny, 1, 1, 20.0)
    g.PreCalculateFrequencies()
    o.apply_microscope(300.0, obj=25.0)
    o.set_calc(ps, 0.0, nx, ny, 1, 1, 20.0, 0, 0, True)
    g.AllocTDS()
    g.STEMInitialiseWavefunctions(31, 17)
    g.set_conventional_dz(dz)
    o.init_stem(31, 17, dz)
    assert rel_l2(g.get_wave(1), o.wave(1)) < 1e-5
    for it in range(1, 6):
        g.MultisliceStepConv(it)
        o.step_conv(it)
    assert rel_l2(g.GetExitWave(), o.wave()) < 1e-4
    assert o.adf(5.0, 200.0) > 0
    assert g.adf(5.0, 200.0) == pytest.approx(o.adf(5.0, 200.0), rel=1e-4)
    g.AddTDSWaves()
    o.add_tds()
    assert rel_l2(g.GetTDS(), o.tds()) < 1e-4
    F = O.fft2(o.wave())
    centred = np.roll(F, (ny // 2, nx // 2), axis=(0, 1))
    assert rel_l2(g.GetSTEMExitWave(), centred) < 1e-4
    # the batched path keeps power-of-two sides
    from cudamultislice_amd._lib import MslError
    with pytest.raises(MslError) as e:
        g.stem_batched(np.array([3], np.int32), np.array([4], np.int32), dz)
    assert e.value.code == 5


def test_fd_steps_on_a_generic_grid(fp):
    """Finite-difference steps (U.cu:2040-2083) on a 60x48 grid."""
    Z, x, y, z = S.silicon(2, 2, 3)
    g, o = _pair(Z, x, y, z, fp)
    nx, ny = 60, 48
    g.ApplyMicroscopeParameters(200.0)
    g.SetCalculationVariables(0.2, 0.0, nx, ny, 1, 1, 20.0, 0.0, 0.0)
    g.PreCalculateFrequencies()
    o.apply_microscope(200.0)
    o.set_calc(0.2, 0.0, nx, ny, 1, 1, 20.0)
    g.InitialiseWavefunctions()
    o.init_plane_fd()
    e = o.extents()
    first = int(0.5 * (e[4] - e[5]) / o.scalars()["dzFD"])
    for it in range(first, first + 20):
        g.MultisliceStep(it)
        o.step_fd(it)
    assert rel_l2(g.get_wave(0), o.wave(0)) < 1e-4
    assert np.max(np.abs(o.wave(0) - 1.0)) > 1e-3
