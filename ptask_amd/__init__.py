"""ptask_amd -- an MI355X-native execution layer for the PTask dataflow runtime.

The product is native: libptask.so (the PTask runtime behind include/ptask/ptaskapi.h, with Host
and HIP backends), libptk_hip.so (the thin HIP C ABI of include/ptk_hip.h) and gfx950 code
objects under lib/kernels/.  This Python package only locates those files and binds the C ABI
for tests and the benchmark driver.
"""
from . import paths  # noqa: F401

__all__ = ["paths"]
