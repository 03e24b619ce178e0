"""Random material parameters drawn as the reference draws them (through the
C++ API's akantu_random_values, libakantu_amd.so): std::default_random_engine
seeded with the run's seed, values base + distribution(engine) in order, the
input file's parameter order (`uniform [min, max]`, `weibull [scale, shape]`;
src/io/parser/parser_random.cc:71-83, src/common/aka_random_generator.hh:72-96,
230-234).  The draws of one seed are the thresholds a `Yd = base weibull
[scale, shape]` material section gets for its points, in element then point
order, when it is the run's first random parameter."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libakantu_amd.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError("libakantu_amd.so is not built (akantu_amd/build.py)")
        L = ctypes.CDLL(_LIB)
        L.akantu_random_values.restype = ctypes.c_int
        L.akantu_random_values.argtypes = [ctypes.c_long, ctypes.c_char_p, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                           ctypes.c_void_p]
        _lib = L
    return _lib


def random_values(n, dist="weibull", base=0.0, p0=1e6, p1=5.0, seed=5489):
    """n draws of base + dist(p0, p1) from a fresh engine seeded with `seed`."""
    out = np.zeros(int(n), dtype=np.float64)
    rc = _load().akantu_random_values(int(seed), dist.encode(), float(base), float(p0),
                                      float(p1), out.size,
                                      out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError("unknown distribution %r" % dist)
    return out


def weibull_thresholds(n, scale, shape, seed=5489, base=0.0):
    """`Yd = base weibull [scale, shape]` for n points."""
    return random_values(n, "weibull", base, scale, shape, seed)
