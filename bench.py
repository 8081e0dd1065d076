#!/usr/bin/env python3
"""Benchmark of the batched STEM + frozen-phonon multislice path at 512^2.

Workload (the configuration the headline metric is quoted on): SrTiO3
10x10x50 cells (25 000 atoms), 512x512 grid over 40 A, 300 kV, 2 A slices
(99 slices), 64x64 probe raster (4096 probes, 8 px step, 25 mrad), 16
frozen-phonon configurations with sigma 0.08 A (Philox seed 0x4D534C), ADF
60-200 mrad, complex64 waves.  One step = one frozen-phonon configuration:
binning, potentials for every slice and the propagation of all 4096 probes
through all slices, ending in the ADF reduction.  The default run times all
16 configurations after one warm-up configuration.

Alongside the headline value the JSON line reports the per-probe launch path
(the reference API driven one probe at a time, potential rebuilt per slice),
fwd+bwd/fwd of the differentiable path on 1024^2 x 100 slices x 64 probes,
two plane-wave TEM cases (Si 256^2 on GPU and CPU; Si 1024^2 x 252 slices
with an 8-defocus CTF sweep), torch.fft (rocFFT) on the same 512^2 waves as
a reference point, the roofline of the dominant kernel and the CPU
restatement's throughput on a bounded sample.  Diagnostics go to stderr;
stdout carries one JSON line.  --diff-only and --serve JOBS print the line
of one secondary workload instead (the differentiable path; the uneven
serving mix).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def host_cores():
    """Every CPU this process may run on (its affinity set): the CPU
    baseline's thread count."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup_cfg3(MultisliceSimulation, S, N=512, L=40.0):
    Z, x, y, z = S.srtio3(10, 10, 50)
    sim = MultisliceSimulation(0)
    sim.set_atoms(Z, x, y, z)
    sim.GetParameterisation()
    sim.ApplyMicroscopeParameters(300.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 25.0)
    inf = sim.info()
    sim.STEMSetCalculationVariables(L / N, 0.0, N, N, inf["MaximumX"] - inf["MinimumX"],
                                    inf["MaximumY"] - inf["MinimumY"], inf["MaximumZ"] - inf["MinimumZ"])
    sim.PreCalculateFrequencies()
    return sim, (Z, x, y, z)


def raster(N=512, n=64, step=8):
    import numpy as np
    xs = (np.arange(n) * step) % N
    px, py = np.meshgrid(xs, xs)
    return px.ravel().astype(np.int32), py.ravel().astype(np.int32)


def per_probe_path(sim, px, py, dz, nslices, nprobe):
    """The reference's call sequence (MW.cs:366-399) one probe at a time:
    probe, then per slice the potential rebuild and the step, then the
    detector.  Returns probe*slice/s."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in range(nprobe):
        sim.STEMInitialiseWavefunctions(int(px[p]), int(py[p]))
        sim.set_conventional_dz(dz)
        for it in range(1, nslices + 1):
            sim.MultisliceStepConv(it)
        sim.adf(60.0, 200.0)
    torch.cuda.synchronize()
    return nprobe * nslices / (time.perf_counter() - t0)


def diff_ratio(reps=3):
    """fwd+bwd / fwd of the differentiable path: 1024^2, 100 slices of 2 A
    projected potential of amorphous carbon (1.2 g/cm^3, 102.4x102.4x200 A,
    min distance 1.4 A), 64 probes (8x8 raster, 25 mrad, 300 kV), loss
    sum (|psi|^2 - I_target)^2 with I_target from a seeded +-2 % perturbed
    potential.  Timed twice on the same inputs: with checkpoints every 10
    slices (11.8 GiB with every slice's transmission, each segment
    recomputed once in the backward; the headline ratio, the workload's own
    setting) and with every slice's wave kept in HBM (checkpoint_every = 0:
    52.3 GiB of workspace, no recomputation)."""
    import numpy as np
    import torch
    from cudamultislice_amd import MultisliceSimulation, POT_PROJECTED
    from cudamultislice_amd import structures as S
    from cudamultislice_amd.diff import DiffEngine, MultisliceIntensity
    N, ns, B, dz, ps = 1024, 100, 64, 2.0, 0.1
    t0 = time.perf_counter()
    Z, x, y, z = S.amorphous_carbon()
    log(f"[diff] amorphous carbon: {Z.size} atoms in {time.perf_counter() - t0:.1f} s")
    sim = MultisliceSimulation(0)
    sim.set_atoms(Z, x, y, z)
    sim.GetParameterisation()
    sim.UploadBinnedAtoms()
    sim.set_potential_mode(POT_PROJECTED)
    sim.ApplyMicroscopeParameters(300.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 25.0)
    inf = sim.info()
    off = -float(inf["MinimumX"])  # pixel 0 at x = 0 of the specimen
    sim.SetCalculationVariables(ps, 0.0, N, N, 1, 1, inf["MaximumZ"] - inf["MinimumZ"], off, off)
    sim.PreCalculateFrequencies()
    V = np.stack([sim.potential_slice(i, dz, real=True) for i in range(1, ns + 1)])
    psi0 = np.zeros((B, N, N), np.complex64)
    sim.STEMSetCalculationVariables(ps, 0.0, N, N, 1, 1, 1)
    sim.PreCalculateFrequencies()
    xs = (np.arange(8) * N // 8 + N // 16)
    for i, (a, b) in enumerate([(a, b) for b in xs for a in xs]):
        sim.STEMInitialiseWavefunctions(int(a), int(b))
        psi0[i] = sim.get_wave(1)
    sim.close()
    rng = np.random.Generator(np.random.Philox(S.SEED))
    Vt = torch.from_numpy(V).cuda()
    Vp = torch.from_numpy((V * (1 + 0.02 * rng.uniform(-1, 1, V.shape))).astype(np.float32)).cuda()
    p0 = torch.from_numpy(psi0).cuda()
    with torch.no_grad():
        eng = DiffEngine(N, N, ps, 300.0, dz, ns, B, checkpoint_every=10)
        target = eng.forward(Vp, p0).clone()
        eng.close()
    res = {}
    for K in (0, 10):
        eng = DiffEngine(N, N, ps, 300.0, dz, ns, B, checkpoint_every=K)
        res[K] = _diff_time(eng, Vt, p0, target, reps)
        eng.close()
        log(f"[diff] checkpoint_every={K}: fwd {res[K][0]:.1f} ms, fwd+bwd {res[K][1]:.1f} ms")
    tf, tfb, ws = res[10]
    tf0, tfb0, ws0 = res[0]
    return {"fwd_ms": tf, "fwd_bwd_ms": tfb, "ratio": tfb / tf, "grid": N, "slices": ns, "probes": B,
            "checkpoint_every": 10, "workspace_GiB": ws / 2**30,
            "keep_every_slice": {"fwd_ms": tf0, "fwd_bwd_ms": tfb0, "ratio": tfb0 / tf0, "checkpoint_every": 0,
                                 "workspace_GiB": ws0 / 2**30}}


def _diff_time(eng, Vt, p0, target, reps):
    """(forward ms, forward+backward ms, workspace bytes) of one engine"""
    import torch
    from cudamultislice_amd.diff import MultisliceIntensity

    def fwd():  # inference forward: nothing kept for a backward
        with torch.no_grad():
            I = eng.forward(Vt, p0, keep=False)
            return ((I - target) ** 2).sum()

    def fwd_bwd():
        Vr = Vt.detach().requires_grad_(True)
        I = MultisliceIntensity.apply(Vr, p0, eng)
        loss = ((I - target) ** 2).sum()
        loss.backward()
        return Vr.grad

    fwd(), fwd_bwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fwd()
    torch.cuda.synchronize()
    tf = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        g = fwd_bwd()
    torch.cuda.synchronize()
    tfb = (time.perf_counter() - t0) / reps
    assert torch.isfinite(g).all()
    return tf * 1e3, tfb * 1e3, eng.info()["workspace_bytes"]


def copy_rates_gbs(nbytes=4 << 30, reps=5):
    """bytes read + written per second by two device-to-device copies of
    nbytes: torch's copy_ and libmsl's streaming copy kernel (16-byte vector
    loads and stores, grid-stride), each timed with HIP events on torch's
    current stream, median of reps.  msl_copy_device waits for its kernel, so
    its interval also holds the host's return (microseconds against ~1 ms)."""
    import ctypes
    import torch
    from cudamultislice_amd._lib import lib
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ops = {"torch_copy": lambda: dst.copy_(src),
           "k_copy16": lambda: lib().msl_copy_device(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()),
                                                    ctypes.c_longlong(nbytes), st)}
    out = {}
    for name, op in ops.items():
        op()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            r = op()
            e1.record()
            e1.synchronize()
            assert name != "k_copy16" or r == 0
            ts.append(e0.elapsed_time(e1))
        out[name] = 2 * nbytes / (sorted(ts)[len(ts) // 2] * 1e-3) / 1e9
    del src, dst
    return out


def inplace_rate_gbs(nbytes, reps=10):
    """bytes read + written per second by an in-place elementwise pass
    (torch mul_ by 1) over nbytes of float32, timed with HIP events over reps
    back-to-back passes on torch's current stream"""
    import torch
    x = torch.ones(max(1, nbytes // 4), dtype=torch.float32, device="cuda")
    x.mul_(1.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        x.mul_(1.0)
    e1.record()
    e1.synchronize()
    rate = 2 * x.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del x
    return rate


def tem_cases(cpu_threads):
    """Plane-wave TEM through the reference-shaped API: (1) Si 4x4x10 cells,
    256^2 over 21.72 A, 1 A slices, 200 kV, exit wave + CTF image (df -500 A,
    Cs 1.2e7 A, 20 mrad), on the GPU and, in full, in the CPU restatement;
    (2) Si 14x14x92 cells (144 256 atoms), 1024^2 over 76.02 A, 2 A slices,
    300 kV, exit wave + CTF images for 8 defoci -800..+600 A (Cs 1e4 A,
    20 mrad) on the GPU."""
    import numpy as np
    import torch
    from cudamultislice_amd import MultisliceSimulation
    from cudamultislice_amd import structures as S
    from cudamultislice_amd.fparams import fparams
    from oracle import oracle as O
    out = {}
    for name, cells, N, ps, kV, dz, df, Cs in (("si_4x4x10_256", (4, 4, 10), 256, 0.084844, 200.0, 1.0, -500.0, 1.2e7),
                                               ("si_14x14x92_1024", (14, 14, 92), 1024, 0.07424, 300.0, 2.0, -800.0, 1e4)):
        Z, x, y, z = S.silicon(*cells)
        sim = MultisliceSimulation(0)
        sim.set_atoms(Z, x, y, z)
        sim.GetParameterisation()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.UploadBinnedAtoms()
        sim.ApplyMicroscopeParameters(kV, df, 0.0, 0.0, Cs, 0.0, 0.0, 20.0)
        inf = sim.info()
        sim.SetCalculationVariables(ps, df, N, N, inf["MaximumX"] - inf["MinimumX"], inf["MaximumY"] - inf["MinimumY"],
                                    inf["MaximumZ"] - inf["MinimumZ"], 2.0, 2.0)
        sim.PreCalculateFrequencies()
        sim.InitialiseWavefunctions(dz)
        n = sim.GetSlices(True)
        for it in range(1, n + 1):
            sim.MultisliceStepConv(it)
        ew = sim.GetExitWave()
        t1 = time.perf_counter()
        if N == 256:
            img = sim.SimulateImage(1.0)
            ndf = 1
        else:
            img = sim.ctf_sweep(np.arange(-800, 601, 200, dtype=np.float32))
            ndf = 8
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        norm = float(np.mean(np.abs(ew) ** 2))
        assert np.all(np.isfinite(ew)) and np.all(np.isfinite(img)) and 0.0 < norm <= 1.0 + 1e-4
        rec = {"atoms": int(Z.size), "grid": N, "slices": n, "gpu_ms": (t1 - t0) * 1e3, "ctf_images": ndf,
               "ctf_ms": (t2 - t1) * 1e3, "exit_mean_intensity": norm}
        sim.close()
        if N == 256:
            o = O.OracleSim(use_float=True)
            o.set_atoms(Z, x, y, z)
            o.set_fparams(fparams())
            O.set_threads(cpu_threads)
            t0 = time.perf_counter()
            o.bin()
            e = o.extents()
            o.apply_microscope(kV, df, 0.0, 0.0, Cs, 0.0, 0.0, 20.0)
            o.set_calc(ps, df, N, N, float(e[0] - e[1]), float(e[2] - e[3]), float(e[4] - e[5]), 2.0, 2.0, False)
            o.init_plane(dz)
            for it in range(1, n + 1):
                o.step_conv(it)
            o.ctf_image(df)
            rec["cpu_ms"] = (time.perf_counter() - t0) * 1e3
            rec["cpu_threads"] = cpu_threads
        out[name] = rec
        log(f"[tem] {name}: {rec}")
    return out


def rocfft_reference(nbatch=512, N=512, reps=5):
    """torch.fft.fft2 (rocFFT) forward + inverse on nbatch N^2 complex64
    waves, as a reference point for the two propagation passes."""
    import torch
    x = torch.randn(nbatch, N, N, dtype=torch.complex64, device="cuda")
    y = torch.fft.ifft2(torch.fft.fft2(x))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y = torch.fft.ifft2(torch.fft.fft2(x))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    del x, y
    return {"waves": nbatch, "grid": N, "fft2_plus_ifft2_ms": ms, "ms_per_wave": ms / nbatch}


def serving_mix(njobs):
    """The uneven serving mix: njobs jobs with grids in {256, 512, 1024,
    2048}^2 (0.1 A pixels), 20-200 slices of 2 A or finite-difference steps,
    half conventional / half finite differences, half plane wave / half STEM
    with 16-256 probes (25 mrad, 300 kV, ADF 20-200 mrad), 1k-50k atoms with
    Z in 6..79 at least 1.4 A apart (capped at 0.05 atoms per A^3).  Returns
    (record, jobs with their results)."""
    import numpy as np
    from cudamultislice_amd.serve import FD, serve, serving_mix as mix
    t0 = time.perf_counter()
    jobs = mix(njobs)
    log(f"[serve] {njobs} jobs generated in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    # at 0.1 A pixels the band limit sits at 49 mrad: the detector reads 20-200 mrad
    # (a 60-200 mrad annulus would read exactly zero)
    st = serve(jobs, kV=300.0, ps=0.1, dz=2.0, aperture=25.0, adf=(20.0, 200.0))
    wall = time.perf_counter() - t0
    for j in jobs:
        r = j.result.get("adf", j.result.get("exit"))
        assert r is not None and np.all(np.isfinite(r))
    px_slices = sum(max(1, j.nprobe) * j.nslices * j.n * j.n for j in jobs)
    by = {}
    for j in jobs:
        k = f"{j.n}^2 {'fd' if j.method == FD else 'conv'} {'stem' if j.nprobe else 'plane'}"
        by[k] = by.get(k, 0) + 1
    log(f"[serve] {st}")
    return {"jobs": njobs, "wall_s": wall, "jobs_per_s": njobs / wall, "waves": int(st["waves"]),
            "wave_slices": int(st["wave_slices"]), "pixel_slices_per_s": px_slices / wall,
            "wave_slices_per_s_512_equivalent": px_slices / (512 * 512) / wall,
            "engine": st, "classes": by,
            "atoms": int(sum(j.Z.size for j in jobs)), "probes": int(sum(j.nprobe for j in jobs))}, jobs


def serving_cpu_baseline(jobs, kV, ps, dz, aperture, adf):
    """The oracle (FP64 restatement, multithreaded) on the serving mix's
    sample of 5.5 % of its jobs (every method, kind and grid size), timed as
    the CPU baseline; the same run checks the engine's results for those jobs
    (largest relative error: exit waves in L2, ADF per probe)."""
    import numpy as np
    from cudamultislice_amd.fparams import fparams
    from oracle import oracle as O
    sample = O.serving_sample(jobs)
    fp = fparams()
    worst = 0.0
    px_slices = 0
    t0 = time.perf_counter()
    for k in sample:
        j = jobs[k]
        want = O.job_reference(j.Z, j.x, j.y, j.z, j.n, j.method, j.nslices, j.posx if j.nprobe else None, j.posy,
                               fp, kV, ps, dz, aperture, adf)
        if "exit" in want:
            e = float(np.linalg.norm(j.result["exit"] - want["exit"]) / np.linalg.norm(want["exit"]))
        else:
            e = float(np.max(np.abs(j.result["adf"] - want["adf"]) / np.maximum(want["adf"], 1e-300)))
        worst = max(worst, e)
        px_slices += max(1, j.nprobe) * j.nslices * j.n * j.n
    dt = time.perf_counter() - t0
    return {"value": px_slices / dt, "unit": "pixel*slices/s", "cores": host_cores(), "kind": "port",
            "sample": f"{len(sample)} of {len(jobs)} jobs (oracle.serving_sample), FP64 restatement, {dt:.1f} s",
            "max_rel_err_vs_engine": worst}


def cpu_baseline(atoms, px, py, nprobe, nslices, threads):
    """The CPU restatement (oracle/, templated to float, its own radix-2 FFT)
    on the same workload, multithreaded over probes; bounded sample."""
    import numpy as np
    from cudamultislice_amd.fparams import fparams
    from oracle import oracle as O
    Z, x, y, z = atoms
    o = O.OracleSim(use_float=True)
    o.set_atoms(Z, x, y, z)
    o.set_fparams(fparams())
    o.bin(O.tds_displacements(0x4D534C, 0, Z.size, 0.08))
    e = o.extents()
    o.apply_microscope(300.0, obj=25.0)
    o.set_calc(40.0 / 512, 0.0, 512, 512, 1, 1, float(e[4] - e[5]), stem=True)
    O.set_threads(threads)
    t0 = time.perf_counter()
    o.stem_batch(px[:nprobe], py[:nprobe], 2.0, nslices, 60.0, 200.0, nthreads=threads)
    dt = time.perf_counter() - t0
    return nprobe * nslices / dt, dt


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="frozen-phonon configurations timed")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--probes", type=int, default=4096)
    ap.add_argument("--batch", type=int, default=0, help="probes per launch (0: the batch-width autotuner's choice)")
    ap.add_argument("--no-tune-batch", action="store_true", help="with --batch 0: all probes in one launch")
    ap.add_argument("--graph", choices=("on", "off"), default="on",
                    help="timed region: replay each chunk from a hipGraph (on) or launch eagerly (off)")
    ap.add_argument("--profile-configs", type=int, default=2,
                    help="configurations run after the timed region with per-kernel HIP-event timing")
    ap.add_argument("--per-probe", type=int, default=8, help="probes timed on the per-probe path (0: skip)")
    ap.add_argument("--cpu-probes", type=int, default=128)
    ap.add_argument("--cpu-slices", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-tem", action="store_true", help="skip the plane-wave TEM cases and the rocFFT reference")
    ap.add_argument("--diff-only", action="store_true",
                    help="time only the differentiable path (fwd, fwd+bwd) and print its line")
    ap.add_argument("--serve", type=int, default=0, metavar="JOBS",
                    help="time only the uneven serving mix of JOBS jobs and print its line")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied once by the plain copy kernel (traffic-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the probe/configuration "
              "batch shards across GPUs with one ADF reduction at the end, not built here)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cudamultislice_amd import MultisliceSimulation
    from cudamultislice_amd import structures as S

    torch.cuda.set_device(0)
    if args.diff_only:
        print(json.dumps({"fwd_bwd_over_fwd": diff_ratio()}), flush=True)
        return
    if args.serve > 0:
        rec, jobs = serving_mix(args.serve)
        out = {"serving_mix": rec}
        if not args.no_cpu:
            out["cpu_baseline"] = serving_cpu_baseline(jobs, 300.0, 0.1, 2.0, 25.0, (20.0, 200.0))
            log(f"[serve] cpu baseline {out['cpu_baseline']}")
        print(json.dumps(out), flush=True)
        return
    if args.calibrate_copy > 0:
        import ctypes
        from cudamultislice_amd._lib import lib
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        assert lib().msl_copy_device(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nbytes,
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)) == 0
        log(f"[bench] copy kernel moved {nbytes} B read + {nbytes} B write")
        del src, dst
    sim, atoms = setup_cfg3(MultisliceSimulation, S)
    inf = sim.info()
    dz = 2.0
    nslices = int(math.ceil((inf["MaximumZ"] - inf["MinimumZ"]) / dz))
    px, py = raster()
    px, py = px[:args.probes], py[:args.probes]
    nprobe = px.size
    log(f"[bench] {atoms[0].size} atoms, {nslices} slices, {nprobe} probes, 512^2")

    # batch width: the autotuner's choice (one chunk per candidate width over
    # 16 slices, before the timed region) unless --batch fixes it
    tuned = None
    if args.batch == 0 and not args.no_tune_batch:
        best, cost = sim.autotune_batch(px, py, dz, nslices=16, adf=(60.0, 200.0))
        tuned = {"best": best, "us_per_probe_slice": {str(k): round(v, 4) for k, v in cost.items()}}
        log(f"[bench] batch width autotuner: {tuned}")
    # warm-up configurations use seeds distinct from the timed ones (and
    # capture the chunk graph when it is on)
    use_graph = args.graph == "on"
    sim.stem_batched(px, py, dz, nconfig=args.warmup, seed=S.SEED + 1000, sigma=0.08, batch=args.batch,
                     use_graph=use_graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = sim.stem_batched(px, py, dz, nconfig=args.steps, seed=S.SEED, sigma=0.08, adf=(60.0, 200.0),
                         batch=args.batch, use_graph=use_graph)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tm = sim.last_timings()
    # per-kernel durations for the roofline: HIP events around each launch on
    # the handle's stream (eager launches), after the timed region so the
    # events cost it nothing
    sim.set_profiling(True)
    sim.stem_batched(px, py, dz, nconfig=max(1, args.profile_configs), seed=S.SEED + 2000, sigma=0.08,
                     batch=args.batch)
    torch.cuda.synchronize()
    ks = sim.kernel_stats()
    sim.set_profiling(False)
    ps_per_s = args.steps * nprobe * nslices / dt
    adf = r["adf"]
    log(f"[bench] batched: {ps_per_s:.4g} probe*slice/s ({dt:.2f} s), stages {tm}, kernels {ks}")
    log(f"[bench] ADF mean {adf.mean():.4g} min {adf.min():.4g} max {adf.max():.4g}")

    # roofline of the dominant propagation kernel.  Algorithmic bytes: the
    # wave batch read + written once per pass (complex64); with band pruning
    # only the kx columns inside the propagator band (half) move; the row
    # pass also reads the shared transmission slice once.
    npix = 512 * 512
    B = (tuned["best"] if tuned else nprobe) if args.batch <= 0 else min(args.batch, nprobe)
    kern = max(("row", "column"), key=lambda k: ks[k]["ms_total"])
    band = 0.5 if inf["band_tiles"] > 0 else 1.0
    bytes_launch = int(16 * B * npix * band) + (8 * npix if kern == "row" else 0)
    achieved = bytes_launch / (ks[kern]["ms_avg"] * 1e-3) / 1e9
    # measured HBM traffic: FETCH_SIZE / WRITE_SIZE per wave of the same pass
    # (rocprofv3, calibrated on a plain copy kernel), scaled to this batch
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_stem512.json")
    if os.path.exists(tpath):
        tk = json.load(open(tpath))["kernels"].get(kern)
        if tk and band < 1.0:
            traffic = int((tk["fetch_bytes_per_wave"] + tk["write_bytes_per_wave"]) * B)
    roof = {"bound": "hbm", "kernel": f"{kern} pass (512, batch {B})", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": "profiles/traffic_stem512.json (rocprofv3 FETCH_SIZE x 2.0, WRITE_SIZE x 1.0, "
                              "per wave, measured on launches of the same width)" if traffic else None,
            "algorithmic_bytes_per_launch": bytes_launch, "avg_launch_ms": ks[kern]["ms_avg"],
            "launch_timing": f"HIP events around each launch on the handle's stream, {ks[kern]['launches']} "
                             f"eager launches over {max(1, args.profile_configs)} configurations after the timed region",
            "other_passes": {k: ks[k]["ms_avg"] for k in ("row", "column", "detector", "potential")}}
    # what plain device copies reach on this GPU (read + write, 4 GiB each
    # way): the practical HBM rate next to the 8 TB/s peak.  The passes'
    # algorithmic rates can exceed it where the Infinity Cache serves part of
    # their bytes (256-probe launches move 256 MiB of in-band waves).
    cps = copy_rates_gbs()
    cp = max(cps.values())
    roof["copy_GBps"] = {k: round(v, 1) for k, v in cps.items()}
    roof["frac_of_copy"] = round(achieved / cp, 4)
    col_bytes = int(16 * B * npix * band)
    col_rate = col_bytes / (ks["column"]["ms_avg"] * 1e-3) / 1e9
    roof["column_GBps"] = round(col_rate, 1)
    log(f"[bench] device copies {roof['copy_GBps']} GB/s (read + write); {kern} pass at {achieved / cp:.3f} "
        f"of the faster, column pass {col_rate:.0f} GB/s")
    # the passes work in place on B in-band waves (8 B * npix * band each):
    # an in-place streaming pass over a buffer of that size (read and write
    # every element once, the working set the Infinity Cache then holds) is
    # their closest plain-memory reference; both passes of a slice against it
    ip = inplace_rate_gbs(int(8 * B * npix * band))
    both = (bytes_launch + col_bytes) / ((ks[kern]["ms_avg"] + ks["column"]["ms_avg"]) * 1e-3) / 1e9
    roof["inplace_GBps"] = round(ip, 1)
    roof["slice_passes_GBps"] = round(both, 1)
    roof["slice_passes_frac_of_inplace"] = round(both / ip, 4)
    log(f"[bench] in-place stream over {8 * B * npix * band / 2**20:.0f} MiB {ip:.0f} GB/s; row + column passes "
        f"{both:.0f} GB/s ({both / ip:.3f} of it)")
    # the row pass is bound by vector-ALU issue: instructions per launch from
    # the committed counter run (per wave, scaled to this batch) against one
    # wave64 instruction per SIMD every 4 cycles at 2.4 GHz
    vpath = os.path.join(ROOT, "profiles", "valu_stem512.json")
    if os.path.exists(vpath) and band < 1.0:
        vk = json.load(open(vpath))["kernels"].get(kern)
        if vk:
            instr = vk["valu_instr_per_wave"] * B
            floor_ms = instr * 4 / (1024 * 2.4e9) * 1e3
            roof["valu"] = {"instr_per_launch": instr, "min_ms_at_issue_peak": round(floor_ms, 5),
                            "issue_frac": round(floor_ms / ks[kern]["ms_avg"], 4),
                            "source": "profiles/valu_stem512.json (rocprofv3 SQ_INSTS_VALU per wave)"}
            log(f"[bench] {kern} pass VALU issue {floor_ms / ks[kern]['ms_avg']:.3f} of the 2.4 GHz rate")

    per_probe = None
    if args.per_probe > 0:
        sim.set_tds(S.SEED, 0.08)
        sim.UploadBinnedAtomsTDS()
        v = per_probe_path(sim, px, py, dz, nslices, args.per_probe)
        # same path with the potential cached: one probe per launch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.stem_batched(px[:args.per_probe], py[:args.per_probe], dz, nconfig=1, seed=S.SEED, sigma=0.08,
                         batch=1)
        torch.cuda.synchronize()
        v_cached = args.per_probe * nslices / (time.perf_counter() - t0 - sim.last_timings()["potential_ms"] * 1e-3
                                               - sim.last_timings()["binning_ms"] * 1e-3)
        per_probe = {"value": v, "cached_potential_value": v_cached, "probes_timed": args.per_probe,
                     "speedup_batched_over_per_probe": ps_per_s / v,
                     "speedup_batched_over_cached": ps_per_s / v_cached}
        log(f"[bench] per-probe {v:.4g}, cached-potential per-probe {v_cached:.4g} probe*slice/s")
    sim.close()

    diff = None
    if not args.no_diff:
        diff = diff_ratio()
        log(f"[bench] diff: {diff}")

    tem = rocfft = None
    if not args.no_tem:
        tem = tem_cases(host_cores())
        rocfft = rocfft_reference()
        # the library's two passes for the same work (one slice: the same two
        # 2D transforms plus the pointwise factors), per wave
        rocfft["msl_row_plus_column_ms_per_wave"] = (ks["row"]["ms_avg"] + ks["column"]["ms_avg"]) / B
        log(f"[bench] rocfft reference {rocfft}")

    cpu = None
    if not args.no_cpu:
        threads = host_cores()
        v, t = cpu_baseline(atoms, px, py, args.cpu_probes, args.cpu_slices, threads)
        model = None
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
        except OSError:
            pass
        cpu = {"value": v, "unit": "probe*slice/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_probes} probes x {args.cpu_slices} slices x 1 configuration of the same "
                         f"workload (potentials for those slices included), float restatement, {t:.1f} s",
               "host_cpus": os.cpu_count(), "cpus_usable": threads, "cpu_model": model}
        log(f"[bench] cpu baseline {v:.4g} probe*slice/s on {threads} threads ({t:.1f} s)")

    out = {
        "metric": "probe*slice/s at 512^2 FP32 (batched STEM + TDS)",
        "value": ps_per_s,
        "unit": "probe*slice/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64 waves)",
        "data": "synthetic (SrTiO3 supercell, seeded Philox frozen-phonon displacements)",
        "config": {"workload": "batched STEM + TDS, SrTiO3 10x10x50, 512^2 over 40 A, 300 kV, 2 A slices",
                   "probes": nprobe, "slices": nslices, "configs_per_step": 1, "batch": B,
                   "batch_autotuner": tuned, "graph_replay": use_graph,
                   "adf_mrad": [60, 200], "step": "one frozen-phonon configuration"},
        "stage_ms_total": tm,
        "per_probe_path": per_probe,
        "fwd_bwd_over_fwd": diff,
        "tem": tem,
        "rocfft_reference": rocfft,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
