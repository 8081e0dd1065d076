"""Differentiable multislice behind torch.autograd.

`MultisliceIntensity.apply(V, psi0, engine)` returns the exit-wave
intensities |psi_n|^2 of a batch of probes propagated through per-slice
potentials V (float32, [nslices, ny, nx], V*A); its backward returns dL/dV
through the hand-written adjoint kernels of libmsl (waves checkpointed every
`checkpoint_every` slices, or all kept when they fit).  torch supplies device memory and the stream;
all arithmetic runs in libmsl.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import MslError, lib


class DiffEngine:
    def __init__(self, nx, ny, ps, kV, dz, nslices, nprobe, checkpoint_every=-1, device=0):
        """checkpoint_every: K > 0 keeps every K-th wave and recomputes the rest
        in the backward; 0 keeps every slice's wave; -1 (default) the longest
        segment whose waves fit in half the free device memory"""
        self.shape = (nprobe, ny, nx)
        self.nslices = nslices
        self.device = device
        h = C.c_void_p()
        st = lib().msl_diff_create(C.byref(h), device, nx, ny, ps, kV, dz, nslices, nprobe, checkpoint_every)
        if st != 0:
            raise MslError(st, "msl_diff_create")
        self._h = h
        self._gen = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().msl_diff_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def info(self):
        wl, sg, km, ws = C.c_float(), C.c_float(), C.c_float(), C.c_longlong()
        lib().msl_diff_info(self._h, C.byref(wl), C.byref(sg), C.byref(km), C.byref(ws))
        return {"wavelength": wl.value, "sigma": sg.value, "kmax": km.value, "workspace_bytes": ws.value}

    def set_chunk(self, probes: int) -> None:
        """probes per adjoint chunk of the backward (0: automatic)"""
        self._check(lib().msl_diff_set_chunk(self._h, int(probes)))

    def _check(self, st):
        if st != 0:
            raise MslError(st, lib().msl_diff_last_error(self._h).decode())

    def _validate(self, V, psi0):
        B, ny, nx = self.shape
        if V.dtype != torch.float32 or tuple(V.shape) != (self.nslices, ny, nx) or not V.is_contiguous():
            raise ValueError(f"V must be contiguous float32 {(self.nslices, ny, nx)}")
        if psi0.dtype != torch.complex64 or tuple(psi0.shape) != (B, ny, nx) or not psi0.is_contiguous():
            raise ValueError(f"psi0 must be contiguous complex64 {(B, ny, nx)}")
        if V.device.type != "cuda" or psi0.device.type != "cuda":
            raise ValueError("V and psi0 must be on the GPU")

    def forward(self, V, psi0, keep=True):
        """intensities |psi_n|^2; keep=False stores nothing for a backward
        (inference), after which backward raises until the next forward"""
        self._validate(V, psi0)
        out = torch.empty(self.shape, dtype=torch.float32, device=V.device)
        s = torch.cuda.current_stream(V.device).cuda_stream
        f = lib().msl_diff_forward if keep else lib().msl_diff_forward_only
        self._check(f(self._h, C.c_void_p(V.data_ptr()), C.c_void_p(psi0.data_ptr()), C.c_void_p(out.data_ptr()),
                      C.c_void_p(s)))
        self._gen += 1
        return out

    def backward(self, dLdI, gen=None):
        if gen is not None and gen != self._gen:
            raise RuntimeError("DiffEngine: another forward ran since this graph was built")
        dLdI = dLdI.contiguous().to(torch.float32)
        g = torch.empty((self.nslices,) + self.shape[1:], dtype=torch.float32, device=dLdI.device)
        s = torch.cuda.current_stream(dLdI.device).cuda_stream
        self._check(lib().msl_diff_backward(self._h, C.c_void_p(dLdI.data_ptr()), C.c_void_p(g.data_ptr()),
                                            C.c_void_p(s)))
        return g


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
