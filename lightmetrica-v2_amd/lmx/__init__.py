"""lmx -- Python host side of the MI355X-native Lightmetrica hot path.

The compute lives in liblmx.so (HIP kernels for gfx950 behind the C ABI in
include/lmx.h).  This package loads it with ctypes and fails loudly when it
is missing: there is no CPU fallback on the product path.
"""
from ._lib import lib, LmxError, check, LIB_PATH  # noqa: F401
from . import accel, render, scenes  # noqa: F401
