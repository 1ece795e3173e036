"""ctypes access to ``oracle/_build/liboracle.so`` (built by ``oracle/build.py``)."""
import ctypes

import numpy as np

from .build import LIBRARY, build

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            build()
        _lib = ctypes.CDLL(str(LIBRARY))
    return _lib


def arr(a, dtype):
    """Contiguous copy/view of ``a`` as ``dtype`` plus its address."""
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(ctypes.c_void_p)
