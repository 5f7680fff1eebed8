"""frovedis_amd: MI355X (gfx950) dataframe execution core.

Python front end over the C ABI of ``libfrovedis_hip.so``
(``include/frovedis_hip.h``).  Names and argument meanings follow the
reference's interfaces (src/frovedis/core/radix_sort.hpp, prefix_sum.hpp,
set_operations.hpp, lower_bound.hpp; src/frovedis/dataframe/hashtable.hpp,
join.hpp, dfcolumn_groupby.cc, dfcolumn_grouped_agg.cc).  Device memory comes
from torch tensors; every operation runs in the HIP library, and importing
this package fails if that library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfrovedis_hip.so")

I32, U32, I64, U64, F32, F64 = range(6)
TYPE_CODE = {np.dtype(np.int32): I32, np.dtype(np.uint32): U32,
             np.dtype(np.int64): I64, np.dtype(np.uint64): U64,
             np.dtype(np.float32): F32, np.dtype(np.float64): F64}
CODE_DTYPE = {v: k for k, v in TYPE_CODE.items()}


class FhipError(RuntimeError):
    """Raised when an entry point returns a non-zero status (the reference
    throws std::runtime_error in the same places)."""


_lib = None


def lib():
    """Loads libfrovedis_hip.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FhipError(
                f"{LIB_PATH} is not built; run `python -m frovedis_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        # every entry point gets its prototype from include/frovedis_hip.h:
        # sizes and pointers then cross as 64-bit values, never as C ints
        from . import abi
        abi.apply(L)
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise FhipError(f"status {rc}: {lib().fhip_last_error().decode()}")


from .core import *  # noqa: E402,F401,F403
from .dataframe import *  # noqa: E402,F401,F403
from .learned import *  # noqa: E402,F401,F403
from .batch import *  # noqa: E402,F401,F403
from .dftable import *  # noqa: E402,F401,F403
