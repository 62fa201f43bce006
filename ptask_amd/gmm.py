"""EM-GMM model layout and the graph application's dump format, for the Python drivers.

Mirrors include/ptk_gmm.h: the GMM_DIMS constants block and the binary64 model block
([0] steps done, [1] N, [2] D, [3] K, mu[K*D], var[K*D], frac[K], loglike[maxIter+1]).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

MAX_K = 64
MAX_D = 32


class Dims(C.Structure):
    """GMM_DIMS (include/ptk_gmm.h), 32 bytes."""
    _fields_ = [("N", C.c_int), ("D", C.c_int), ("K", C.c_int), ("blocks", C.c_int),
                ("maxIter", C.c_int), ("pad", C.c_int * 3)]


def dpad(D: int) -> int:
    return (D + 3) & ~3


def model_offsets(D: int, K: int, max_iter: int = 0) -> dict:
    mu = 4
    var = mu + K * D
    frac = var + K * D
    ll = frac + K
    table = (ll + max_iter + 2) & ~1
    return {"mu": mu, "var": var, "frac": frac, "ll": ll, "table": table}


def table_floats(D: int, K: int) -> int:
    return 2 * K * dpad(D) + K


def model_elems(D: int, K: int, max_iter: int) -> int:
    return model_offsets(D, K, max_iter)["table"] + (table_floats(D, K) + 1) // 2


def fill_table(model: np.ndarray, D: int, K: int, max_iter: int) -> None:
    """The binary32 E-step table of a model from its mu, var and frac, as ptk_gmm_fill_table
    (include/ptk_gmm.h) builds it."""
    o = model_offsets(D, K, max_iter)
    Dp = dpad(D)
    mu = model[o["mu"]:o["var"]].reshape(K, D)
    var = model[o["var"]:o["frac"]].reshape(K, D)
    frac = model[o["frac"]:o["ll"]]
    tab = np.zeros(table_floats(D, K), np.float32)
    rows = tab[:2 * K * Dp].reshape(K, 2, Dp)
    rows[:, 0, :D] = mu.astype(np.float32)
    rows[:, 1, :D] = (1.0 / var).astype(np.float32)
    with np.errstate(divide="ignore"):
        tab[2 * K * Dp:] = (-0.5 * np.log(var).sum(axis=1) + np.log(frac)).astype(np.float32)
    raw = np.zeros(2 * (model.size - o["table"]), np.float32)
    raw[:tab.size] = tab
    model[o["table"]:] = raw.view(np.float64)


def partial_stride(D: int, K: int) -> int:
    return K + 2 * K * D + 1


@dataclass
class Model:
    steps: int
    mu: np.ndarray
    var: np.ndarray
    frac: np.ndarray
    loglike: np.ndarray


def parse_model(a: np.ndarray, D: int, K: int, steps: int) -> Model:
    o = model_offsets(D, K)
    return Model(int(a[0]), a[o["mu"]:o["var"]].reshape(K, D), a[o["var"]:o["frac"]].reshape(K, D),
                 a[o["frac"]:o["ll"]].copy(), a[o["ll"]:o["ll"] + steps].copy())


def read_dump(path: str, N: int, D: int, K: int, steps: int):
    """(X[N,D] binary32, initial means[K,D], final Model) from `ptk_gmm --dump`."""
    raw = open(path, "rb").read()
    nx = N * D * 4
    X = np.frombuffer(raw[:nx], dtype=np.float32).reshape(N, D)
    mu0 = np.frombuffer(raw[nx:nx + K * D * 8], dtype=np.float64).reshape(K, D)
    model = np.frombuffer(raw[nx + K * D * 8:], dtype=np.float64)
    if model.size != model_elems(D, K, steps):
        raise ValueError("dump size does not match N, D, K, steps")
    return X, mu0, parse_model(model, D, K, steps)
