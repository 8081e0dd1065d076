"""MI355X-native multislice electron-scattering engine.

The simulation runs in libmsl.so (HIP kernels for gfx950 behind the C-ABI
in include/msl.h); this package is its Python host mirror.
"""
from ._lib import LIB_PATH, MslError, lib  # noqa: F401
from .fparams import fparams  # noqa: F401
from .simulation import MultisliceSimulation, fft2_device, tds_displacements, write_tiff  # noqa: F401

POT_INTEGRATED = 0
POT_PROJECTED = 1
