"""Serving a mix of independent multislice jobs of uneven size through
msl_server_run (include/msl.h).

Each job is the reference's per-simulation call sequence on its own atoms
(LoadAtomFile extents and UploadBinnedAtoms, SetCalculationVariables with a
zero block offset or STEMSetCalculationVariables, PreCalculateFrequencies,
then MultisliceStepConv or MultisliceStep, "GUI Simulation/
unmanagedGUISimulation.cu":1293-2124); the engine batches every job of one
grid size into shared row / column launches per slice.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import structures as S
from ._lib import MslError, lib, msl_job, msl_serve_params

CONVENTIONAL, FD = 0, 1


@dataclass
class Job:
    n: int
    Z: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    nslices: int
    method: int = CONVENTIONAL
    posx: np.ndarray | None = None  # probe positions (pixels); None: plane wave
    posy: np.ndarray | None = None
    want_exit: bool = True
    result: dict = field(default_factory=dict)

    @property
    def nprobe(self) -> int:
        return 0 if self.posx is None else int(np.asarray(self.posx).size)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def serve(jobs: list[Job], kV: float = 300.0, ps: float = 0.1, dz: float = 2.0, aperture: float = 25.0,
          adf=(60.0, 200.0), potential_mode: int = 0, budget_bytes: int = 0, device: int = 0,
          group_sizes: bool = False) -> dict:
    """Runs every job; fills job.result ("exit" for plane waves, "adf" for
    probes) and returns the engine's statistics.  group_sizes: conventional
    jobs of grid 256..2048 share one launch per pass and slice (half the
    launches; on MI355X the per-size launches are faster for large jobs)."""
    L = lib()
    keep = []
    arr = (msl_job * len(jobs))()
    for k, j in enumerate(jobs):
        Z = np.ascontiguousarray(j.Z, np.int32)
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (j.x, j.y, j.z))
        d = arr[k]
        d.n, d.method, d.nslices, d.natoms = j.n, j.method, j.nslices, Z.size
        d.Z, d.x, d.y, d.z = _p(Z), _p(x), _p(y), _p(z)
        keep += [Z, x, y, z]
        j.result = {}
        if j.nprobe > 0:
            px = np.ascontiguousarray(j.posx, np.int32)
            py = np.ascontiguousarray(j.posy, np.int32)
            out = np.zeros(px.size, np.float32)
            d.nprobe, d.posx, d.posy, d.adf_out = px.size, _p(px), _p(py), _p(out)
            j.result["adf"] = out
            keep += [px, py]
        elif j.want_exit:
            ew = np.zeros((j.n, j.n), np.complex64)
            d.exit_out = _p(ew)
            j.result["exit"] = ew
    prm = msl_serve_params(kV, ps, dz, aperture, adf[0], adf[1], potential_mode, budget_bytes, int(group_sizes))
    h = C.c_void_p()
    st = L.msl_server_create(C.byref(h), device)
    if st != 0:
        raise MslError(st, "msl_server_create failed")
    try:
        stats = np.zeros(8, np.float64)
        st = L.msl_server_run(h, C.byref(prm), arr, len(jobs), _p(stats))
        if st != 0:
            raise MslError(st, L.msl_server_last_error(h).decode())
    finally:
        L.msl_server_destroy(h)
    names = ("total_ms", "setup_ms", "propagation_ms", "batches", "waves", "wave_slices", "launches",
             "largest_batch_bytes")
    return dict(zip(names, (float(v) for v in stats)))


def serving_mix(njobs: int = 256, seed: int = S.SEED, ps: float = 0.1, dz: float = 2.0,
                sizes=(256, 512, 1024, 2048), slices=(20, 200), atoms=(1000, 50000), probes=(16, 256),
                density_cap: float = 0.05) -> list[Job]:
    """The uneven serving mix: per job a grid from `sizes`, a slice count in
    `slices`, half conventional and half finite-difference, half plane wave
    and half STEM with a probe count in `probes` at random pixels, and a
    random atom count in `atoms` (Z uniform in 6..79, minimum distance
    1.4 A, S.jittered_atoms), capped at `density_cap` atoms per A^3 of a box
    that spans the grid (2 A margins) and the slices' depth (finite
    differences: nslices steps of at most 0.06 A)."""
    rng = np.random.Generator(np.random.Philox(seed))
    jobs = []
    for k in range(njobs):
        n = int(rng.choice(sizes))
        ns = int(rng.integers(slices[0], slices[1] + 1))
        method = FD if rng.random() < 0.5 else CONVENTIONAL
        stem = rng.random() < 0.5
        lateral = n * ps - 4.0
        thick = max(3.0, ns * (dz if method == CONVENTIONAL else 0.06))
        nat = int(min(rng.integers(atoms[0], atoms[1] + 1), density_cap * lateral * lateral * thick))
        Z, x, y, z = S.jittered_atoms(max(nat, 1), (lateral, lateral, thick), seed=seed + 1000 + k)
        posx = posy = None
        if stem:
            npr = int(rng.integers(probes[0], probes[1] + 1))
            posx = rng.integers(0, n, npr).astype(np.int32)
            posy = rng.integers(0, n, npr).astype(np.int32)
        jobs.append(Job(n, Z, x, y, z, ns, method, posx, posy))
    return jobs
