"""The oracle: CPU restatements of the reference algorithms on the PTask hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as the checker and by
the CPU-baseline leg of bench.py -- never by the product (ptask_amd/).  The C code lives beside
this file and is built into oracle/_build/libptk_oracle.so by oracle/Makefile.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libptk_oracle.so")

FIELDS = ("Hx", "Hy", "Hz", "Ex", "Ey", "Ez")


class FdtdParams(C.Structure):
    """FDTD_PARAMS, 40 bytes (ptaskUnitTest/graphfdtd.h:6-19)."""
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
                ("Cx", C.c_float), ("Cy", C.c_float), ("Cz", C.c_float),
                ("eps0", C.c_float), ("mu0", C.c_float), ("c0", C.c_float), ("Dt", C.c_float)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def make_params(nx: int, ny: int, nz: int, courant: float = 0.5) -> FdtdParams:
    """Restates ptk_fdtd_make_params (include/ptk_fdtd.h) in binary32 via numpy, following
    run_graph_fdtd_task (ptaskUnitTest/graphfdtd.cpp:1357-1381)."""
    f = np.float32
    lx, ly, lz = f(0.05), f(0.04), f(0.03)
    p = FdtdParams()
    p.Nx, p.Ny, p.Nz = nx, ny, nz
    cx, cy, cz = f(nx) / lx, f(ny) / ly, f(nz) / lz
    p.Cx, p.Cy, p.Cz = float(cx), float(cy), float(cz)
    p.eps0 = float(f(8.8541878e-12))
    p.mu0 = float(f(4e-7) * f(3.1415))
    c0 = f(299792458.0)
    p.c0 = float(c0)
    nrm = np.sqrt(f(cx * cx + cy * cy + cz * cz), dtype=f)
    p.Dt = float(f(courant) / (c0 * nrm))
    return p


def extent(field: int, nx: int, ny: int, nz: int) -> tuple[int, int, int]:
    """(d1, d2, d3) of a field array (graphfdtd.cpp:70-86)."""
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1),
            (nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)][field]


def init_fields(nx: int, ny: int, nz: int, seed: int) -> list[np.ndarray]:
    """The seeded U[-1,1) fields (flat, i fastest) from the shared counter-based generator."""
    L = lib()
    L.ptk_oracle_fdtd_init_field.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_uint64]
    out = []
    for f in range(6):
        d1, d2, d3 = extent(f, nx, ny, nz)
        a = np.empty(d1 * d2 * d3, dtype=np.float32)
        L.ptk_oracle_fdtd_init_field(a.ctypes.data, f, nx, ny, nz, seed)
        out.append(a)
    return out


def fdtd_run(fields: list[np.ndarray], params: FdtdParams, steps: int,
             dtype=np.float32) -> list[np.ndarray]:
    """Runs `steps` full Yee steps (fdtd_cpu, graphfdtd.cpp:314-340) on copies of `fields`."""
    L = lib()
    sfx = "f32" if dtype == np.float32 else "f64"
    fn = getattr(L, f"ptk_oracle_fdtd_run_ptrs_{sfx}")
    fn.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(FdtdParams), C.c_int]
    cur = [np.array(a, dtype=dtype, copy=True) for a in fields]
    tmp = [np.empty_like(a) for a in cur]
    keep = {a.ctypes.data: a for a in cur + tmp}
    fp = (C.c_void_p * 6)(*[a.ctypes.data for a in cur])
    sp = (C.c_void_p * 6)(*[a.ctypes.data for a in tmp])
    fn(fp, sp, C.byref(params), steps)
    return [keep[fp[i]] for i in range(6)]


def pfxsum(a: np.ndarray) -> np.ndarray:
    """Exclusive scan with wrap-around (host_pfxsum, ptaskUnitTest/pfxsum.cpp:56-69)."""
    L = lib()
    if a.dtype == np.int64:
        fn = L.ptk_oracle_pfxsum_i64
    else:
        a = a.astype(np.int32)
        fn = L.ptk_oracle_pfxsum_i32
    a = np.ascontiguousarray(a)
    out = np.empty_like(a)
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong]
    fn(a.ctypes.data, out.ctypes.data, a.size)
    return out


def groupby(keys: np.ndarray, vals: np.ndarray, num_keys: int):
    """The groupby pipeline (oracle/groupby.c): returns (groups[G,3] of key,count,sum int64,
    header (G, rows, errors), offsets[num_keys+1], out_keys, out_vals) with groups in ascending
    key order and rows of a group in input order."""
    L = lib()
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.int32)
    n = keys.size
    groups = np.zeros(3 + 3 * (num_keys + 1), dtype=np.int64)
    offsets = np.zeros(num_keys + 1, dtype=np.int64)
    ok, ov = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    fn = L.ptk_oracle_groupby
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_int] + [C.c_void_p] * 4
    fn(keys.ctypes.data, vals.ctypes.data, n, num_keys, groups.ctypes.data, offsets.ctypes.data,
       ok.ctypes.data, ov.ctypes.data)
    G = int(groups[0])
    return groups[3:3 + 3 * G].reshape(G, 3), groups[:3].copy(), offsets, ok, ov


def gmm_em(X: np.ndarray, means_init: np.ndarray, iters: int, threads: int = 0):
    """Diagonal-covariance EM as CPU_EM drives NR3's Gaumixmod (oracle/gmm.c): the constructor's
    E+M step, then `iters` more.  Returns (mu[K,D], var[K,D], frac[K], loglike[iters+1]).
    threads > 0 runs the multi-threaded form (oracle/parallel.c), bitwise equal to the serial one."""
    L = lib()
    X = np.ascontiguousarray(X, dtype=np.float64)
    N, D = X.shape
    mu = np.array(means_init, dtype=np.float64, copy=True).reshape(-1, D)
    K = mu.shape[0]
    var = np.empty((K, D), np.float64)
    frac = np.empty(K, np.float64)
    ll = np.empty(iters + 1, np.float64)
    args = [X.ctypes.data, N, D, K, iters, mu.ctypes.data, var.ctypes.data, frac.ctypes.data,
            ll.ctypes.data]
    if threads > 0:
        fn = L.ptk_oracle_gmm_em_mt
        fn.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_int]
        fn(*args, threads)
    else:
        fn = L.ptk_oracle_gmm_em
        fn.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4
        fn(*args)
    return mu, var, frac, ll


def pca_gs(X: np.ndarray, K: int, J: int, center: bool = True, threads: int = 0):
    """Gram-Schmidt PCA as pca_cpu runs it (oracle/pca.c).  X is M x N (any layout; it is passed
    column-major).  Returns (T[M,K], P[N,K], R[M,N], L[K]).  threads > 0 runs the multi-threaded
    form (oracle/parallel.c), bitwise equal to the serial one."""
    L_ = lib()
    X = np.asfortranarray(X, dtype=np.float64)
    M, N = X.shape
    T = np.zeros((M, K), order="F")
    P = np.zeros((N, K), order="F")
    R = np.zeros((M, N), order="F")
    Lv = np.zeros(K)
    args = [X.ctypes.data, M, N, K, J, int(center), T.ctypes.data, P.ctypes.data, R.ctypes.data,
            Lv.ctypes.data]
    if threads > 0:
        fn = L_.ptk_oracle_pca_gs_mt
        fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_int]
        fn(*args, threads)
    else:
        fn = L_.ptk_oracle_pca_gs
        fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4
        fn(*args)
    return T, P, R, Lv


def default_threads() -> int:
    """Threads for the oracle's multi-threaded forms: OMP_NUM_THREADS when set, else the CPUs
    this process may run on, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, len(os.sched_getaffinity(0))))


def fdtd_update(field: int, own: np.ndarray, a: np.ndarray, b: np.ndarray,
                params: FdtdParams, dtype=np.float32) -> np.ndarray:
    """One of the six half-step updates (graphfdtd.cpp:185-312) with the reference's argument
    order (own, in1, in2)."""
    L = lib()
    name = ["hx", "hy", "hz", "ex", "ey", "ez"][field]
    sfx = "f32" if dtype == np.float32 else "f64"
    fn = getattr(L, f"ptk_oracle_fdtd_{name}_{sfx}")
    fn.argtypes = [C.c_void_p] * 3 + [C.POINTER(FdtdParams), C.c_void_p]
    own, a, b = (np.ascontiguousarray(x, dtype=dtype) for x in (own, a, b))
    out = np.empty_like(own)
    fn(own.ctypes.data, a.ctypes.data, b.ctypes.data, C.byref(params), out.ctypes.data)
    return out
