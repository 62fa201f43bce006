"""ctypes binding of the thin HIP C ABI (include/ptk_hip.h, libptk_hip.so).

This is how Python tests and the benchmark reach the device: the same entry points the C++
HIPAccelerator/HIPTask/PHIPBuffer/HIPAsyncContext classes call.  Every call is checked; a non-zero
status raises HipError with the library's own error text.  Nothing here falls back to the CPU: if
the library or a code object is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
from typing import Iterable, Sequence

from . import paths

PTK_HIP_OK = 0
PTK_HIP_NOT_READY = 1

_STATUS_NAMES = {
    0: "ok", 1: "not ready", 2: "invalid argument", 3: "no HIP device", 4: "out of memory",
    5: "code object or symbol not found", 6: "launch failed", 7: "HIP runtime error",
}


class HipError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str):
        super().__init__(f"{what} failed: {_STATUS_NAMES.get(status, status)} ({detail})")
        self.status = status


class DeviceProps(C.Structure):
    _fields_ = [
        ("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int),
        ("clock_khz", C.c_int), ("mem_clock_khz", C.c_int), ("mem_bus_width_bits", C.c_int),
        ("l2_bytes", C.c_int), ("max_threads_per_block", C.c_int), ("warp_size", C.c_int),
        ("lds_bytes_per_block", C.c_int), ("concurrent_kernels", C.c_int),
        ("runtime_version", C.c_int), ("total_global_mem", C.c_size_t),
    ]


# Every symbol include/ptk_hip.h declares, with its ctypes signature (restype, argtypes).
_vp, _sz, _i, _u = C.c_void_p, C.c_size_t, C.c_int, C.c_uint
SIGNATURES = {
    "ptk_hip_get_device_count": (_i, [C.POINTER(_i)]),
    "ptk_hip_get_device_props": (_i, [_i, C.POINTER(DeviceProps)]),
    "ptk_hip_set_device": (_i, [_i]),
    "ptk_hip_get_device": (_i, [C.POINTER(_i)]),
    "ptk_hip_device_synchronize": (_i, []),
    "ptk_hip_mem_get_info": (_i, [C.POINTER(_sz), C.POINTER(_sz)]),
    "ptk_hip_malloc": (_i, [C.POINTER(_vp), _sz]),
    "ptk_hip_free": (_i, [_vp]),
    "ptk_hip_host_malloc": (_i, [C.POINTER(_vp), _sz]),
    "ptk_hip_host_free": (_i, [_vp]),
    "ptk_hip_memcpy_htod_async": (_i, [_vp, _vp, _sz, _vp]),
    "ptk_hip_memcpy_dtoh_async": (_i, [_vp, _vp, _sz, _vp]),
    "ptk_hip_memcpy_dtod_async": (_i, [_vp, _vp, _sz, _vp]),
    "ptk_hip_memcpy_htod": (_i, [_vp, _vp, _sz]),
    "ptk_hip_memcpy_dtoh": (_i, [_vp, _vp, _sz]),
    "ptk_hip_stream_create": (_i, [C.POINTER(_vp)]),
    "ptk_hip_stream_destroy": (_i, [_vp]),
    "ptk_hip_stream_synchronize": (_i, [_vp]),
    "ptk_hip_stream_query": (_i, [_vp]),
    "ptk_hip_stream_wait_event": (_i, [_vp, _vp]),
    "ptk_hip_event_create": (_i, [C.POINTER(_vp), _i]),
    "ptk_hip_event_destroy": (_i, [_vp]),
    "ptk_hip_event_record": (_i, [_vp, _vp]),
    "ptk_hip_event_query": (_i, [_vp]),
    "ptk_hip_event_synchronize": (_i, [_vp]),
    "ptk_hip_event_elapsed_ms": (_i, [C.POINTER(C.c_float), _vp, _vp]),
    "ptk_hip_module_load": (_i, [C.POINTER(_vp), C.c_char_p]),
    "ptk_hip_module_load_data": (_i, [C.POINTER(_vp), _vp]),
    "ptk_hip_module_unload": (_i, [_vp]),
    "ptk_hip_module_get_function": (_i, [C.POINTER(_vp), _vp, C.c_char_p]),
    "ptk_hip_function_get_attrs": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                                        C.POINTER(_i)]),
    "ptk_hip_launch": (_i, [_vp, _u, _u, _u, _u, _u, _u, _u, _vp, _vp, _sz]),
    "ptk_hip_stream_begin_capture": (_i, [_vp]),
    "ptk_hip_stream_end_capture": (_i, [_vp, C.POINTER(_vp)]),
    "ptk_hip_graph_launch": (_i, [_vp, _vp]),
    "ptk_hip_graph_destroy": (_i, [_vp]),
    "ptk_hip_status_string": (C.c_char_p, [_i]),
    "ptk_hip_last_error_string": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """Load libptk_hip.so (built in-tree by `make -C ptask_amd/csrc`)."""
    global _lib
    if _lib is None:
        path = paths.lib_path("libptk_hip.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def check(status: int, what: str) -> int:
    if status not in (PTK_HIP_OK, PTK_HIP_NOT_READY):
        detail = lib().ptk_hip_last_error_string().decode(errors="replace")
        raise HipError(status, what, detail)
    return status


def device_count() -> int:
    n = C.c_int(0)
    check(lib().ptk_hip_get_device_count(C.byref(n)), "get_device_count")
    return n.value


def device_props(dev: int = 0) -> DeviceProps:
    p = DeviceProps()
    check(lib().ptk_hip_get_device_props(dev, C.byref(p)), "get_device_props")
    return p


def set_device(dev: int) -> None:
    check(lib().ptk_hip_set_device(dev), "set_device")


def synchronize() -> None:
    check(lib().ptk_hip_device_synchronize(), "device_synchronize")


class DeviceBuffer:
    """A device allocation; freed explicitly or when garbage collected."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().ptk_hip_malloc(C.byref(p), self.nbytes), f"malloc({self.nbytes})")
        self.ptr = p.value or 0

    def free(self) -> None:
        if self.ptr:
            check(lib().ptk_hip_free(C.c_void_p(self.ptr)), "free")
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, host, stream: "Stream | None" = None) -> None:
        import numpy as np
        a = np.ascontiguousarray(host)
        if a.nbytes > self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte device buffer")
        if stream is None:
            check(lib().ptk_hip_memcpy_htod(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                            a.nbytes), "memcpy_htod")
        else:
            check(lib().ptk_hip_memcpy_htod_async(C.c_void_p(self.ptr),
                                                  a.ctypes.data_as(C.c_void_p), a.nbytes,
                                                  stream.handle), "memcpy_htod_async")
            stream.synchronize()

    def download(self, out) -> None:
        import numpy as np
        if not out.flags["C_CONTIGUOUS"]:
            raise ValueError("download needs a C-contiguous host array")
        if out.nbytes > self.nbytes:
            raise ValueError(f"download of {out.nbytes} bytes from a {self.nbytes}-byte device buffer")
        check(lib().ptk_hip_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                        out.nbytes), "memcpy_dtoh")

    def copy_from(self, other: "DeviceBuffer", nbytes: int, stream: "Stream") -> None:
        nbytes = int(nbytes)
        if nbytes < 0 or nbytes > self.nbytes or nbytes > other.nbytes:
            raise ValueError(f"copy of {nbytes} bytes between device buffers of {other.nbytes} "
                             f"(source) and {self.nbytes} (destination) bytes")
        check(lib().ptk_hip_memcpy_dtod_async(C.c_void_p(self.ptr), C.c_void_p(other.ptr), nbytes,
                                              stream.handle), "memcpy_dtod_async")


class Stream:
    def __init__(self):
        h = C.c_void_p()
        check(lib().ptk_hip_stream_create(C.byref(h)), "stream_create")
        self.handle = h

    def synchronize(self) -> None:
        check(lib().ptk_hip_stream_synchronize(self.handle), "stream_synchronize")

    def destroy(self) -> None:
        if self.handle:
            check(lib().ptk_hip_stream_destroy(self.handle), "stream_destroy")
            self.handle = None


class Event:
    def __init__(self, timing: bool = True):
        h = C.c_void_p()
        check(lib().ptk_hip_event_create(C.byref(h), 1 if timing else 0), "event_create")
        self.handle = h

    def record(self, stream: Stream) -> None:
        check(lib().ptk_hip_event_record(self.handle, stream.handle), "event_record")

    def synchronize(self) -> None:
        check(lib().ptk_hip_event_synchronize(self.handle), "event_synchronize")

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        check(lib().ptk_hip_event_elapsed_ms(C.byref(ms), self.handle, later.handle),
              "event_elapsed_ms")
        return ms.value

    def destroy(self) -> None:
        if self.handle:
            check(lib().ptk_hip_event_destroy(self.handle), "event_destroy")
            self.handle = None


class Module:
    def __init__(self, path: str):
        h = C.c_void_p()
        check(lib().ptk_hip_module_load(C.byref(h), path.encode()), f"module_load({path})")
        self.handle = h
        self._fns: dict[str, C.c_void_p] = {}

    def function(self, name: str) -> C.c_void_p:
        if name not in self._fns:
            f = C.c_void_p()
            check(lib().ptk_hip_module_get_function(C.byref(f), self.handle, name.encode()),
                  f"module_get_function({name})")
            self._fns[name] = f
        return self._fns[name]


class ArgPack:
    """Packs kernel arguments the way HIPTask lays out formal parameters: each at its natural
    alignment, device pointers as 8 bytes, by-value structs as raw bytes (4-byte aligned)."""

    def __init__(self):
        self.buf = bytearray()

    def _align(self, a: int) -> None:
        while len(self.buf) % a:
            self.buf.append(0)

    def ptr(self, p: int) -> "ArgPack":
        self._align(8)
        self.buf += struct.pack("<Q", p)
        return self

    def i32(self, v: int) -> "ArgPack":
        self._align(4)
        self.buf += struct.pack("<i", v)
        return self

    def u32(self, v: int) -> "ArgPack":
        self._align(4)
        self.buf += struct.pack("<I", v)
        return self

    def u64(self, v: int) -> "ArgPack":
        self._align(8)
        self.buf += struct.pack("<Q", v)
        return self

    def f32(self, v: float) -> "ArgPack":
        self._align(4)
        self.buf += struct.pack("<f", v)
        return self

    def raw(self, b: bytes, align: int = 4) -> "ArgPack":
        self._align(align)
        self.buf += b
        return self

    def finish(self):
        self._align(8)
        n = len(self.buf)
        cbuf = (C.c_char * max(n, 1)).from_buffer_copy(bytes(self.buf) or b"\0")
        return cbuf, n


def launch(fn: C.c_void_p, grid: Sequence[int], block: Sequence[int], args: ArgPack,
           stream: Stream | None = None, lds: int = 0) -> None:
    g = list(grid) + [1] * (3 - len(grid))
    b = list(block) + [1] * (3 - len(block))
    cbuf, n = args.finish()
    check(lib().ptk_hip_launch(fn, g[0], g[1], g[2], b[0], b[1], b[2], lds,
                               stream.handle if stream else None, C.cast(cbuf, C.c_void_p), n),
          "launch")


def declared_symbols(header: str | None = None) -> list[str]:
    """Names of the functions declared in include/ptk_hip.h (parsed from the header text)."""
    import re
    text = open(header or os.path.join(paths.include_dir(), "ptk_hip.h")).read()
    return re.findall(r"^\s*(?:int|const char \*)\s*(ptk_hip_\w+)\s*\(", text, re.M)
