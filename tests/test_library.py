"""CPU checks of the C-ABI library: it loads, exports every function that
include/msl.h declares, and its host-only entry points behave."""
import ctypes as C
import os

import numpy as np
import pytest

from cudamultislice_amd import _lib
from oracle import oracle as O

needs_lib = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libmsl.so not built")


@needs_lib
def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    declared = _lib.declared_symbols()
    assert len(declared) > 40
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding declares signatures for all of them
    assert set(declared) <= set(_lib._SIGS), set(declared) - set(_lib._SIGS)


@needs_lib
def test_host_only_entry_points():
    L = _lib.lib()
    assert L.msl_status_string(0) == b"ok"
    assert L.msl_status_string(8) == b"cancelled"
    # Philox displacements: host code, identical to the oracle's
    n = 257
    a = [np.zeros(n, np.float32) for _ in range(3)]
    st = L.msl_tds_displacements(0x4D534C, 3, n, 0.075, *(v.ctypes.data_as(C.c_void_p) for v in a))
    assert st == 0
    for u, v in zip(a, O.tds_displacements(0x4D534C, 3, n, 0.075)):
        assert np.array_equal(u, v)
    # argument validation happens before any device work
    assert L.msl_tds_displacements(1, 0, -1, 0.1, None, None, None) == 1
    assert L.msl_create(None, 0) == 1
    assert L.msl_diff_create(None, 0, 64, 64, 0.1, 200.0, 1.0, 4, 1, 2) == 1


def test_embedded_table_is_what_the_oracle_gets(fp):
    assert fp.shape == (103, 12) and fp.dtype == np.float32
    assert np.all(fp[:, 1::2] > 0)  # b_i, d_i widths are positive


@needs_lib
def test_tiff_writer_round_trips_through_an_independent_reader(tmp_path):
    from PIL import Image
    from cudamultislice_amd import write_tiff
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((48, 80)) * 1e3).astype(np.float32)
    a[0, 0], a[-1, -1] = np.float32(-0.0), np.float32(1e-30)
    p = str(tmp_path / "x.tiff")
    write_tiff(p, a)
    im = Image.open(p)
    tags = {k: (v[0] if isinstance(v, tuple) else v) for k, v in im.tag_v2.items()}  # before decoding
    assert tags[339] == 3 and tags[258] == 32 and tags[259] == 1 and tags[274] == 1 and tags[262] == 1
    assert tags[278] == 48  # one strip
    assert im.mode == "F" and im.size == (80, 48)
    assert np.array_equal(np.asarray(im, np.float32), a)
    L = _lib.lib()
    assert L.msl_write_tiff(p.encode(), 0, 4, a.ctypes.data_as(C.c_void_p)) == 1
    assert L.msl_write_tiff(str(tmp_path / "no" / "such.tiff").encode(), 4, 4, a.ctypes.data_as(C.c_void_p)) == 3


@needs_lib
def test_torch_extension_registers_the_engine_class():
    """libmsl_torch.so registers torch.classes.msl.DiffEngine with the
    forward / backward / set_chunk / info methods the autograd binding calls;
    without a GPU its constructor reports the HIP failure as an msl status."""
    import torch
    from cudamultislice_amd import diff
    diff.load_extension()
    cls = torch.classes.msl.DiffEngine
    assert cls is not None
    if not torch.cuda.is_available():
        from cudamultislice_amd._lib import MslError
        with pytest.raises(MslError) as e:
            diff.DiffEngine(64, 64, 0.25, 300.0, 2.0, 4, 2, checkpoint_every=2)
        assert e.value.code == 4
