"""ctypes binding of the batched / served executor's C ABI (include/ptk_batch.h).

The executor itself is native (ptask_amd/lib/libptk_batch.so); this module only loads it for the
tests and the benchmark driver.  Loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import paths

PLAIN, BATCHED = 0, 1
FDTD, GROUPBY, GMM = 0, 1, 2
QUEUED, ACTIVE, DONE, RELEASED = 0, 1, 2, 3

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = paths.lib_path("libptk_batch.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run `make -C ptask_amd/csrc`")
        L = C.CDLL(path)
        P, I, U64, LL, D = C.c_void_p, C.c_int, C.c_ulonglong, C.c_longlong, C.c_double
        sig = {
            "ptk_batch_create": [C.c_char_p, U64, I, C.POINTER(P)],
            "ptk_batch_destroy": [P],
            "ptk_batch_submit_fdtd": [P, I, I, I, I, U64, I, C.POINTER(I)],
            "ptk_batch_submit_groupby": [P, LL, I, D, U64, I, C.POINTER(I)],
            "ptk_batch_submit_gmm": [P, I, I, I, I, U64, I, C.POINTER(I)],
            "ptk_batch_footprint": [I, LL, I, I, I],
            "ptk_batch_tick": [P, I, C.POINTER(I), C.POINTER(I)],
            "ptk_batch_run": [P, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)],
            "ptk_batch_drain": [P],
            "ptk_batch_status": [P, I, C.POINTER(I), C.POINTER(D), C.POINTER(D), C.POINTER(D)],
            "ptk_batch_checksums": [P, I, C.POINTER(U64)],
            "ptk_batch_read": [P, I, I, P, C.POINTER(U64)],
            "ptk_batch_release": [P, I],
            "ptk_batch_set_tiles": [P, I, I, I],
            "ptk_batch_set_launch": [P, I],
            "ptk_batch_set_fdtd_steps": [P, I],
            "ptk_batch_get_fdtd_steps": [P, C.POINTER(I), C.POINTER(I)],
            "ptk_batch_get_tiles": [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)],
            "ptk_batch_autotune": [P, C.POINTER(D)],
            "ptk_batch_stats": [P, C.POINTER(U64)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = U64 if name == "ptk_batch_footprint" else I
        L.ptk_batch_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Names of the functions declared in include/ptk_batch.h (parsed from the header text)."""
    import re
    text = open(os.path.join(paths.include_dir(), "ptk_batch.h")).read()
    return re.findall(r"^\s*(?:int|unsigned long long|const char \*)\s*(ptk_batch_\w+)\s*\(", text, re.M)


class BatchError(RuntimeError):
    pass


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise BatchError(f"{what} failed ({rc}): {lib().ptk_batch_last_error().decode()}")


class Executor:
    def __init__(self, arena_bytes: int, mode: int = BATCHED):
        h = C.c_void_p()
        _check(lib().ptk_batch_create(paths.KERNEL_DIR.encode(), arena_bytes, mode, C.byref(h)), "create")
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().ptk_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit_fdtd(self, nx, ny, nz, steps, seed, keep=False) -> int:
        i = C.c_int()
        _check(lib().ptk_batch_submit_fdtd(self.h, nx, ny, nz, steps, seed, int(keep), C.byref(i)), "submit_fdtd")
        return i.value

    def submit_groupby(self, rows, num_keys, zipf_s, seed, keep=False) -> int:
        i = C.c_int()
        _check(lib().ptk_batch_submit_groupby(self.h, rows, num_keys, zipf_s, seed, int(keep), C.byref(i)),
               "submit_groupby")
        return i.value

    def submit_gmm(self, n, d, k, steps, seed, keep=False) -> int:
        i = C.c_int()
        _check(lib().ptk_batch_submit_gmm(self.h, n, d, k, steps, seed, int(keep), C.byref(i)), "submit_gmm")
        return i.value

    def tick(self, max_admit: int = -1) -> tuple[int, int]:
        a, c = C.c_int(), C.c_int()
        _check(lib().ptk_batch_tick(self.h, max_admit, C.byref(a), C.byref(c)), "tick")
        return a.value, c.value

    def run(self, max_iters: int, max_admit: int = -1) -> tuple[int, int, int]:
        n, a, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().ptk_batch_run(self.h, max_admit, max_iters, C.byref(n), C.byref(a), C.byref(c)), "run")
        return n.value, a.value, c.value

    def drain(self) -> None:
        _check(lib().ptk_batch_drain(self.h), "drain")

    def status(self, i: int):
        st, a, b, c = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().ptk_batch_status(self.h, i, C.byref(st), C.byref(a), C.byref(b), C.byref(c)), "status")
        return st.value, a.value, b.value, c.value

    def checksums(self, i: int) -> list[int]:
        out = (C.c_ulonglong * 8)()
        _check(lib().ptk_batch_checksums(self.h, i, out), "checksums")
        return list(out)

    def read(self, i: int, which: int, dtype) -> np.ndarray:
        n = C.c_ulonglong()
        _check(lib().ptk_batch_read(self.h, i, which, None, C.byref(n)), "read size")
        a = np.empty(n.value // np.dtype(dtype).itemsize, dtype=dtype)
        _check(lib().ptk_batch_read(self.h, i, which, a.ctypes.data, C.byref(n)), "read")
        return a

    def release(self, i: int) -> None:
        _check(lib().ptk_batch_release(self.h, i), "release")

    def set_tiles(self, cols, depth, gb_rows) -> None:
        _check(lib().ptk_batch_set_tiles(self.h, cols, depth, gb_rows), "set_tiles")

    def set_fdtd_steps(self, steps: int) -> None:
        """FDTD steps per batched launch with the whole-step form: 3 (default) or 1."""
        _check(lib().ptk_batch_set_fdtd_steps(self.h, steps), "set_fdtd_steps")

    def fdtd_steps(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().ptk_batch_get_fdtd_steps(self.h, C.byref(a), C.byref(b)), "get_fdtd_steps")
        return a.value, b.value

    def set_launch(self, graphs: bool) -> None:
        _check(lib().ptk_batch_set_launch(self.h, int(graphs)), "set_launch")

    def tiles(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().ptk_batch_get_tiles(self.h, C.byref(a), C.byref(b), C.byref(c)), "get_tiles")
        return a.value, b.value, c.value

    def autotune(self) -> float:
        ms = C.c_double()
        _check(lib().ptk_batch_autotune(self.h, C.byref(ms)), "autotune")
        return ms.value

    def stats(self) -> dict:
        out = (C.c_ulonglong * 7)()
        _check(lib().ptk_batch_stats(self.h, out), "stats")
        keys = ("ticks", "launches", "graph_launches", "arena_used", "arena_peak", "arena_size", "exec_us")
        return dict(zip(keys, list(out)))


def footprint(kind: int, a: int, b: int = 0, c: int = 0, steps: int = 1) -> int:
    return lib().ptk_batch_footprint(kind, a, b, c, steps)
