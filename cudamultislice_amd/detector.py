"""Detector curves as embedded in the engine.

Reads csrc/mtf_tables.h (hex-float literals) so Python sees the exact float
values the HIP library uses for the reference's K2 DQE and NTF curves
(k2DQE / k2NTF, used by DQEKernel / NTFKernel, U.cu:1120-1186).
"""
import os
import re

import numpy as np

_HDR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "mtf_tables.h")
_cache = None


def tables():
    """(dqe, ntf): two float32 arrays of 725 values."""
    global _cache
    if _cache is None:
        txt = open(_HDR).read()
        out = []
        for name in ("kDetectorDQE", "kDetectorNTF"):
            body = re.search(name + r"\[725\] = \{([^}]*)\}", txt).group(1)
            vals = [float.fromhex(m) for m in re.findall(r"(-?0x[0-9a-fA-F.]+p[-+]?\d+)f", body)]
            assert len(vals) == 725, (name, len(vals))
            out.append(np.array(vals, dtype=np.float32))
        _cache = tuple(out)
    return _cache[0].copy(), _cache[1].copy()
