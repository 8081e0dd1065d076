"""Kirkland parameterisation table as embedded in the engine.

Reads csrc/fparams_table.h (103 x 12 hex-float literals) so Python sees the
exact float values the HIP library uses (the reference reads the same
numbers from fparams.dat, U.cu:1718-1747).
"""
import os
import re

import numpy as np

_HDR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "fparams_table.h")
_cache = None


def fparams() -> np.ndarray:
    """(103, 12) float32; row Z-1 = (a1 b1 a2 b2 a3 b3 c1 d1 c2 d2 c3 d3)."""
    global _cache
    if _cache is None:
        txt = open(_HDR).read()
        vals = [float.fromhex(m) for m in re.findall(r"(-?0x[0-9a-fA-F.]+p[-+]?\d+)f", txt)]
        assert len(vals) == 103 * 12, len(vals)
        _cache = np.array(vals, dtype=np.float32).reshape(103, 12)
    return _cache.copy()
