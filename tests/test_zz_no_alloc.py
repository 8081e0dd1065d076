"""msl_fft2_device_noalloc through the C-ABI: the same transforms as
msl_fft2_device (which allocates its twiddle tables per call), with the
tables in the library's static device array; two calls in a row, the second
smaller, each against msl_fft2_device on the same input."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_noalloc_fft_matches_the_allocating_entry_point():
    import torch
    from cudamultislice_amd._lib import lib
    L = lib()
    rng = np.random.default_rng(21)
    stream = torch.cuda.current_stream().cuda_stream
    for ny, nx, nb, inv in ((1024, 2048, 3, False), (256, 512, 2, True), (64, 64, 5, False)):
        a = (rng.standard_normal((nb, ny, nx)) + 1j * rng.standard_normal((nb, ny, nx))).astype(np.complex64)
        want = torch.from_numpy(a.copy()).cuda()
        got = torch.from_numpy(a.copy()).cuda()
        torch.cuda.synchronize()
        assert L.msl_fft2_device(nx, ny, nb, C.c_void_p(want.data_ptr()), int(inv), C.c_void_p(stream)) == 0
        assert L.msl_fft2_device_noalloc(nx, ny, nb, C.c_void_p(got.data_ptr()), int(inv), C.c_void_p(stream)) == 0
        torch.cuda.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        err = rel_l2(g, w)
        assert err < 1e-6, (ny, nx, inv, err)
        ref = np.fft.ifft2(a) * (nx * ny) if inv else np.fft.fft2(a)
        assert rel_l2(g, ref) < 3e-6
    # sizes beyond the static tables are refused before any work
    x = torch.zeros(4, dtype=torch.complex64, device="cuda")
    assert L.msl_fft2_device_noalloc(4096, 16, 1, C.c_void_p(x.data_ptr()), 0, C.c_void_p(stream)) == 5
    assert L.msl_fft2_device_noalloc(16, 16, 70000, C.c_void_p(x.data_ptr()), 0, C.c_void_p(stream)) == 5
    assert L.msl_fft2_device_noalloc(16, 16, 1, None, 0, C.c_void_p(stream)) == 1
