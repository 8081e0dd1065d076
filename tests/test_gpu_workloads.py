"""GPU parity at the workloads' own shapes: the HIP engine (through the
C-ABI) against the FP64 oracle on the batched STEM + frozen-phonon workload
(SrTiO3 10x10x50 cells, 512^2 over 40 A, 99 slices of 2 A, 300 kV, 25 mrad
probes, ADF 60-200 mrad).  Exit waves within 1e-4 relative L2 and ADF
signals within 1e-4 relative of the oracle; at the full raster, properties
that hold for any correct run plus a sample against the oracle."""
import numpy as np
import pytest

from cudamultislice_amd import structures as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = S.SEED


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _stem512(fp):
    """The batched STEM workload's handle: extents from the atoms,
    STEMSetCalculationVariables at 40 A / 512 px, 300 kV, 25 mrad."""
    from cudamultislice_amd import MultisliceSimulation
    Z, x, y, z = S.srtio3(10, 10, 50)
    g = MultisliceSimulation()
    g.set_atoms(Z, x, y, z)
    g.GetParameterisation()
    g.ApplyMicroscopeParameters(300.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 25.0)
    inf = g.info()
    g.STEMSetCalculationVariables(40.0 / 512, 0.0, 512, 512, inf["MaximumX"] - inf["MinimumX"],
                                  inf["MaximumY"] - inf["MinimumY"], inf["MaximumZ"] - inf["MinimumZ"])
    g.PreCalculateFrequencies()
    o = O.OracleSim()
    o.set_atoms(Z, x, y, z)
    o.set_fparams(fp)
    o.apply_microscope(300.0, obj=25.0)
    e = o.extents()
    o.set_calc(40.0 / 512, 0.0, 512, 512, 1, 1, float(e[4] - e[5]), stem=True)
    ns = int(np.ceil(np.float32(e[4] - e[5]) / np.float32(2.0)))
    return g, o, (Z, x, y, z), ns


def _raster(N=512, n=64, step=8):
    xs = (np.arange(n) * step) % N
    px, py = np.meshgrid(xs, xs)
    return px.ravel().astype(np.int32), py.ravel().astype(np.int32)


# eight probes of the 64 x 64 raster: corners, edges and the interior
SAMPLE = np.array([0, 63, 455, 1234, 2047, 2600, 3333, 4095])


@pytest.fixture(scope="module")
def stem512_oracle(fp):
    """FP64 oracle on the eight sampled raster probes for frozen-phonon
    configurations 0 and 1 (seeded Philox displacements, sigma 0.08 A)."""
    g, o, atoms, ns = _stem512(fp)
    g.close()
    px, py = _raster()
    out = []
    for cfg in range(2):
        o.bin(O.tds_displacements(SEED, cfg, atoms[0].size, 0.08))
        out.append(o.stem_batch(px[SAMPLE], py[SAMPLE], 2.0, ns, 60.0, 200.0, want_exit=True))
    return {"ns": ns, "cfg": out}


@pytest.mark.timeout(900)
def test_stem512_adf_ring_is_the_band_limited_annulus(fp):
    """At 40 A / 512 px and 300 kV the band limit (about half Nyquist,
    U.cu:1929-1953) sits at 62.9 mrad, so the 60-200 mrad annulus integrates
    60 mrad up to the band edge: the ring is non-empty and narrow."""
    g, _, _, ns = _stem512(fp)
    inf = g.info()
    assert ns == 99 and g.GetSlices(True) > 0
    edge = inf["kmax"] * inf["wavelength"] * 1e3
    assert 60.0 < edge < 63.5, edge
    g.close()


@pytest.mark.timeout(900)
def test_stem512_batched_exit_waves_and_adf_match_oracle(fp, stem512_oracle):
    """Eight raster probes x two frozen-phonon configurations through all 99
    slices: every exit wave within 1e-4 relative L2 and the configuration
    mean ADF within 1e-4 relative of the FP64 oracle (U.cu:2085-2124 steps
    from the probe of U.cu:2672-2724, ADF over the annulus of the band
    limited spectrum); the fused-detector hipGraph replay gives the same ADF
    as the eager run that reads the exit waves back."""
    g, _, _, ns = _stem512(fp)
    px, py = _raster()
    px, py = px[SAMPLE], py[SAMPLE]
    r = g.stem_batched(px, py, 2.0, nconfig=2, seed=SEED, sigma=0.08, adf=(60.0, 200.0), exit_waves=True)
    adf_o = np.zeros(len(SAMPLE))
    for cfg in range(2):
        ro = stem512_oracle["cfg"][cfg]
        adf_o += ro["adf"] / 2
        for p in range(len(SAMPLE)):
            err = rel_l2(r["exit"][cfg, p], ro["exit"][p])
            assert err < 1e-4, (cfg, p, err)
    assert adf_o.min() > 0
    err = np.abs(r["adf"] - adf_o) / adf_o
    assert err.max() < 1e-4, err
    fused = g.stem_batched(px, py, 2.0, nconfig=2, seed=SEED, sigma=0.08, adf=(60.0, 200.0), use_graph=True,
                           pacbed=True)
    assert np.max(np.abs(fused["adf"] - adf_o) / adf_o) < 1e-4
    assert np.max(np.abs(fused["adf"] - r["adf"]) / r["adf"]) < 1e-5
    g.close()


@pytest.mark.timeout(900)
def test_stem512_full_raster_one_configuration(fp, stem512_oracle):
    """All 4096 probes of one configuration in one call (hipGraph replay,
    fused detector and PACBED): every ADF finite and positive, the sampled
    probes equal to the oracle's configuration-0 ADF within 1e-4, and the
    PACBED (mean |Psi_hat|^2 over probes) summing to no more than the probe
    intensity (|t| = 1 and |P| <= 1 cannot add intensity)."""
    g, o, atoms, ns = _stem512(fp)
    px, py = _raster()
    r = g.stem_batched(px, py, 2.0, nconfig=1, seed=SEED, sigma=0.08, adf=(60.0, 200.0), use_graph=True,
                       pacbed=True)
    adf = r["adf"]
    assert np.all(np.isfinite(adf)) and adf.min() > 0
    want = stem512_oracle["cfg"][0]["adf"]
    err = np.abs(adf[SAMPLE] - want) / want
    assert err.max() < 1e-4, err
    probe_I = float(np.sum(np.abs(o.probe(int(px[0]), int(py[0]))) ** 2))
    # |Psi_hat|^2 of the unnormalised transform: Parseval gives N^2 sum |psi|^2
    pac = float(r["pacbed"].astype(np.float64).sum()) / (512 * 512)
    assert 0 < pac <= probe_I * (1 + 1e-5), (pac, probe_I)
    # the specimen scatters: the mean exit intensity is below the probe's
    assert pac < probe_I
    g.close()
