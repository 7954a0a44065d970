"""MI355X-native FP64 2+1D Teukolsky evolution (s=-2 m-modes of Kerr).

The compute path is libteuk (hand-written HIP kernels for gfx950 behind the
C ABI in include/teuk.h).  This package holds its sources (csrc/), the C++
driver that keeps the reference's run interface (driver/), and a thin ctypes
view used by tests and benchmarks (teuk.py).
"""
from . import teuk  # noqa: F401

__all__ = ["teuk"]
