"""The oracle's result for one serving job (test helper, not collected)."""
from oracle import oracle as O


def job_reference(job, fp, kV, ps, dz, aperture, adf):
    """{"exit": wave} for a plane-wave job, {"adf": per probe} for probes"""
    return O.job_reference(job.Z, job.x, job.y, job.z, job.n, job.method, job.nslices,
                           job.posx if job.nprobe > 0 else None, job.posy, fp, kV, ps, dz, aperture, adf)
