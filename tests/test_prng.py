"""The numpy restatement of the shared generator (tests/prng.py) against the C one."""
import ctypes as C

import numpy as np

import oracle
from prng import rand_u64, rand_unit_f64


def test_numpy_generator_matches_c():
    f = oracle.lib().ptk_oracle_fdtd_init_field
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]
    a = np.empty(3 * 4 * 5, np.float32)        # field 2 (Hz) of a 3 x 4 x 4 grid has 3*4*5 cells
    f(a.ctypes.data, 2, 3, 4, 4, 1234)
    k = (rand_u64(1234, 2, np.arange(a.size)) >> np.uint64(40)).astype(np.float32)
    assert np.array_equal(a, k * np.float32(1.0 / 8388608.0) - np.float32(1.0))
    u = rand_unit_f64(7, 1, np.arange(5))
    assert np.all((u >= 0) & (u < 1))
