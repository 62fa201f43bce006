"""Groupby scratch layout (include/ptk_groupby.h) for host-side sizing and tests."""
from __future__ import annotations

CHUNK = 16384         # rows per chunk (PTK_GB_CHUNK)
SUB = 4096            # rows a GroupbyPartition block ranks at once (PTK_GB_SUB)
TILE = 4096           # partitioned rows per tile (PTK_GB_TILE)
BUCKET = 256          # keys per bucket (PTK_GB_BUCKET)
HIST_THREADS = 512    # PTK_GB_HIST_THREADS
COUNT_THREADS = 256   # PTK_GB_COUNT_THREADS
PART_THREADS = 512    # PTK_GB_PART_THREADS
TILE_THREADS = 512    # PTK_GB_TILE_THREADS
SCAN_THREADS = 1024   # PTK_GB_SCAN_THREADS
KEYSCAN_THREADS = 1024  # PTK_GB_KEYSCAN_THREADS
TOTALS_THREADS = 256    # PTK_GB_TOTALS_THREADS
MAX_BUCKETS = 2048    # PTK_GB_MAX_BUCKETS
MAX_KEYS = BUCKET * MAX_BUCKETS   # numKeys + 1 <= MAX_KEYS


def chunks(n: int) -> int:
    return (n + CHUNK - 1) // CHUNK


def buckets(num_keys: int) -> int:
    return (num_keys + 1 + BUCKET - 1) // BUCKET


def tiles_max(n: int, num_keys: int) -> int:
    return (n + TILE - 1) // TILE + buckets(num_keys)


def _a256(b: int) -> int:
    return (b + 255) // 256 * 256


_ORDER = ["head", "btot", "bstart", "tstart", "tinfo", "nz", "berr", "hist", "kcnt", "c2", "s2", "tval",
          "tlow", "total"]


def scratch_offsets(n: int, num_keys: int) -> dict:
    """Byte offsets of the scratch arrays (include/ptk_groupby.h) and the total size."""
    C, NB, TM = chunks(n), buckets(num_keys), tiles_max(n, num_keys)
    KP = NB * BUCKET
    size = {"head": 256, "btot": 4 * NB, "bstart": 4 * (NB + 1), "tstart": 4 * (NB + 1), "tinfo": 16 * TM,
            "nz": 8 * NB, "berr": 4 * C, "hist": 4 * C * NB, "kcnt": 4 * KP, "c2": 4 * TM * BUCKET,
            "s2": 8 * TM * BUCKET, "tval": 4 * n, "tlow": n + 16}
    o, at = {}, 0
    for name in _ORDER[:-1]:
        o[name] = at
        at = _a256(at + size[name])
    o["total"] = at
    return o


def scratch_bytes(n: int, num_keys: int) -> int:
    return scratch_offsets(n, num_keys)["total"]


# ---- the direct-call library (include/ptk_groupby_api.h, ptask_amd/lib/libptk_groupby.so) ----

OK, EINVAL, ESPACE, EHIP = 0, 1, 2, 3
_lib = None


def lib():
    """ctypes handle of libptk_groupby.so with its argument types bound; fails loudly when the
    library is not built."""
    import ctypes as C
    import os

    from . import paths
    global _lib
    if _lib is None:
        path = paths.lib_path("libptk_groupby.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.ptk_groupby_workspace_bytes.restype = C.c_longlong
        L.ptk_groupby_workspace_bytes.argtypes = [C.c_longlong, C.c_int]
        L.ptk_groupby_run.restype = C.c_int
        L.ptk_groupby_run.argtypes = ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p,
                                       C.c_ulonglong] + [C.c_void_p] * 5)
        L.ptk_groupby_noalloc_capacity.restype = C.c_ulonglong
        L.ptk_groupby_noalloc_capacity.argtypes = []
        L.ptk_groupby_run_noalloc.restype = C.c_int
        L.ptk_groupby_run_noalloc.argtypes = ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_int]
                                              + [C.c_void_p] * 5)
        L.ptk_groupby_last_error.restype = C.c_char_p
        L.ptk_groupby_last_error.argtypes = []
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Names of the functions declared in include/ptk_groupby_api.h."""
    import os
    import re

    from . import paths
    text = open(os.path.join(paths.include_dir(), "ptk_groupby_api.h")).read()
    return re.findall(r"^\s*(?:int|long long|unsigned long long|const char \*)\s*(ptk_groupby_\w+)\s*\(", text, re.M)
