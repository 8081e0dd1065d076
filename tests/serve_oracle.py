"""The oracle's result for one serving job (test helper, not collected):
the reference's call sequence on the job's own atoms -- extents and binning,
SetCalculationVariables with a zero block offset (plane wave) or
STEMSetCalculationVariables (probes), then conventional steps
(U.cu:2085-2124) or finite-difference steps (U.cu:2040-2083) from
InitialiseWavefunctions / STEMInitialiseWavefunctions."""
import numpy as np

from cudamultislice_amd.serve import CONVENTIONAL
from oracle import oracle as O


def job_reference(job, fp, kV, ps, dz, aperture, adf):
    """{"exit": wave} for a plane-wave job, {"adf": per probe} for probes"""
    o = O.OracleSim()
    o.set_atoms(job.Z, job.x, job.y, job.z)
    o.set_fparams(fp)
    o.bin()
    e = o.extents()
    stem = job.nprobe > 0
    o.apply_microscope(kV, obj=aperture if stem else 0.0)
    o.set_calc(ps, 0.0, job.n, job.n, 1, 1, float(e[4] - e[5]), 0.0, 0.0, stem)
    if not stem:
        if job.method == CONVENTIONAL:
            o.init_plane(dz)
            for it in range(1, job.nslices + 1):
                o.step_conv(it)
        else:
            o.init_plane_fd()
            for it in range(1, job.nslices + 1):
                o.step_fd(it)
        return {"exit": o.wave()}
    if job.method == CONVENTIONAL:
        return {"adf": o.stem_batch(job.posx, job.posy, dz, job.nslices, adf[0], adf[1])["adf"]}
    return {"adf": o.fd_batch(job.posx, job.posy, job.nslices, adf[0], adf[1])}
