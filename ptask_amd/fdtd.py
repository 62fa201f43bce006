"""FDTD field layout and launch helpers shared by the Python drivers (tests, benchmark).

The graph itself is built in C++ against include/ptask/ptaskapi.h; this module only mirrors the
few facts a driver needs to call the Yee task kernels directly through the C ABI: the staggered
extents, the (own, in1, in2) argument order of each task (ptaskUnitTest/graphfdtd.cpp:448-560)
and the FDTD_PARAMS bytes (include/ptk_fdtd.h).
"""
from __future__ import annotations

import ctypes as C
import math
import struct

FIELDS = ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")
HX, HY, HZ, EX, EY, EZ = range(6)

# task kernel name, own field, in1, in2 -- in step order (graphfdtd.cpp:314-340)
TASKS = (
    ("HxComputation", HX, EY, EZ),
    ("HyComputation", HY, EX, EZ),
    ("HzComputation", HZ, EX, EY),
    ("ExComputation", EX, HY, HZ),
    ("EyComputation", EY, HX, HZ),
    ("EzComputation", EZ, HX, HY),
)


def extent(field: int, nx: int, ny: int, nz: int) -> tuple[int, int, int]:
    """(d1, d2, d3) of a field array, i fastest (graphfdtd.cpp:70-86)."""
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1),
            (nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)][field]


def elems(field: int, nx: int, ny: int, nz: int) -> int:
    a, b, c = extent(field, nx, ny, nz)
    return a * b * c


class Params(C.Structure):
    """FDTD_PARAMS (include/ptk_fdtd.h), 40 bytes."""
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
                ("Cx", C.c_float), ("Cy", C.c_float), ("Cz", C.c_float),
                ("eps0", C.c_float), ("mu0", C.c_float), ("c0", C.c_float), ("Dt", C.c_float)]

    def to_bytes(self) -> bytes:
        return bytes(self)


def make_params(nx: int, ny: int, nz: int, courant: float = 0.5) -> Params:
    """ptk_fdtd_make_params (include/ptk_fdtd.h) computed in binary32."""
    import numpy as np
    f = np.float32
    p = Params()
    p.Nx, p.Ny, p.Nz = nx, ny, nz
    cx, cy, cz = f(nx) / f(0.05), f(ny) / f(0.04), f(nz) / f(0.03)
    p.Cx, p.Cy, p.Cz = float(cx), float(cy), float(cz)
    p.eps0 = float(f(8.8541878e-12))
    p.mu0 = float(f(4e-7) * f(3.1415))
    c0 = f(299792458.0)
    p.c0 = float(c0)
    nrm = np.sqrt(f(cx * cx + cy * cy + cz * cz), dtype=f)
    p.Dt = float(f(courant) / (c0 * nrm))
    return p


def default_geometry(d1: int, d2: int, d3: int, block: int = 256,
                     target_slab: int = 8) -> tuple[tuple[int, int, int], tuple[int, int, int]]:
    """Grid/block for the column-walk kernels: one lane per (i,j) column, planes split into
    slabs of about `target_slab`."""
    ncol = d1 * d2
    gx = max(1, math.ceil(ncol / block))
    slabs = max(1, min(d3, math.ceil(d3 / target_slab)))
    return (gx, slabs, 1), (block, 1, 1)


def reference_geometry(d1: int, d2: int, d3: int) -> tuple[tuple[int, int, int], tuple[int, int, int]]:
    """The one-thread-per-cell CUDA geometry graphfdtd sets (graphfdtd.cpp:439-446)."""
    return (math.ceil(d1 / 32), d2 * d3, 1), (32, 1, 1)


def pack_task_args(hip_mod, own: int, in1: int, in2: int, params: Params, out: int):
    """Formal parameters (own, in1, in2, FDTD_PARAMS, out) as HIPTask lays them out."""
    return (hip_mod.ArgPack().ptr(own).ptr(in1).ptr(in2).raw(params.to_bytes(), 4).ptr(out))


def pack_bytes(own: int, in1: int, in2: int, params: Params, out: int) -> bytes:
    b = struct.pack("<QQQ", own, in1, in2) + params.to_bytes()
    return b + struct.pack("<Q", out)
