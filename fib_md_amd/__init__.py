"""fib_md_amd -- MI355X-native force-and-integrate path of FIB_MD.

The product is the C-ABI library ``libfibgpu.so`` (include/fibgpu.h) built from
``csrc/`` for gfx950.  This package holds its ctypes binding (``_lib``) and a
thin numpy-level wrapper (``gpu``) used by tests and benchmarks.  There is no
CPU fallback: without a gfx950 device every compute entry point raises.
"""
from . import _lib  # noqa: F401
from .gpu import FibGPU, GPUError, default_params  # noqa: F401

__all__ = ["FibGPU", "GPUError", "default_params"]
