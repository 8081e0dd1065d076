"""Differentiable multislice behind torch.autograd.

`MultisliceIntensity.apply(V, psi0, engine)` returns the exit-wave
intensities |psi_n|^2 of a batch of probes propagated through per-slice
potentials V (float32, [nslices, ny, nx], V*A); its backward returns dL/dV
through the hand-written adjoint kernels of libmsl (waves checkpointed every
`checkpoint_every` slices, or all kept when they fit).  The engine is the
torch extension class torch.classes.msl.DiffEngine (libmsl_torch.so, over
the msl_diff_* C-ABI); torch supplies device memory and the stream, all
arithmetic runs in libmsl.
"""
from __future__ import annotations

import os
import re

import torch

from ._lib import MslError, lib

EXT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmsl_torch.so")
_loaded = False


def load_extension() -> None:
    """Registers torch.classes.msl (raises if the extension is not built)."""
    global _loaded
    if _loaded:
        return
    if not os.path.exists(EXT_PATH):
        raise RuntimeError(f"{EXT_PATH} missing: build it (make -C cudamultislice_amd/csrc)")
    lib()  # libmsl.so first, bound to the HIP runtime torch already loaded
    torch.ops.load_library(EXT_PATH)
    _loaded = True


def _call(f, *args):
    """f(*args), with the extension's failing msl status raised as MslError"""
    try:
        return f(*args)
    except RuntimeError as e:
        m = re.search(r"msl status (\d+)", str(e))
        if m:
            raise MslError(int(m.group(1)), str(e)) from None
        raise


class DiffEngine:
    def __init__(self, nx, ny, ps, kV, dz, nslices, nprobe, checkpoint_every=-1, device=0):
        """checkpoint_every: K > 0 keeps every K-th wave and recomputes the rest
        in the backward; 0 keeps every slice's wave; -1 (default) the longest
        segment whose waves fit in half the free device memory"""
        load_extension()
        self.shape = (nprobe, ny, nx)
        self.nslices = nslices
        self.device = device
        self._e = _call(torch.classes.msl.DiffEngine, device, nx, ny, float(ps), float(kV), float(dz), nslices,
                        nprobe, checkpoint_every)
        self._gen = 0

    def close(self):
        """releases the engine's device memory (once no graph holds it)"""
        self._e = None

    def __del__(self):
        self.close()

    def info(self):
        wl, sg, km, ws = _call(self._e.info)
        return {"wavelength": wl, "sigma": sg, "kmax": km, "workspace_bytes": int(ws)}

    def set_chunk(self, probes: int) -> None:
        """probes per adjoint chunk of the backward (0: automatic)"""
        _call(self._e.set_chunk, int(probes))

    def forward(self, V, psi0, keep=True):
        """intensities |psi_n|^2; keep=False stores nothing for a backward
        (inference), after which backward raises until the next forward"""
        s = torch.cuda.current_stream(V.device).cuda_stream
        out = _call(self._e.forward, V, psi0, bool(keep), s)
        self._gen += 1
        return out

    def backward(self, dLdI, gen=None):
        if gen is not None and gen != self._gen:
            raise RuntimeError("DiffEngine: another forward ran since this graph was built")
        s = torch.cuda.current_stream(dLdI.device).cuda_stream
        return _call(self._e.backward, dLdI, s)


class MultisliceIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, psi0, engine: DiffEngine):
        I = engine.forward(V, psi0)
        ctx.save_for_backward(V, psi0)
        ctx.engine = engine
        ctx.gen = engine._gen
        return I

    @staticmethod
    def backward(ctx, gI):
        V, psi0 = ctx.saved_tensors  # keeps the inputs the engine points at alive
        return ctx.engine.backward(gI, ctx.gen), None, None
