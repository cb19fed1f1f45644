"""bd_cholesky_noalloc: the factor of a caller's device matrix with its
workspace in the library's static device array (no allocation), against
bd_cholesky, which allocates its workspace: the same kernels on the same
input, so the factors and counts must be identical.  Two calls in a row,
the second on a smaller matrix (the array keeps the first call's contents),
the negative-pivot tensor, and the size limit."""
import os

import numpy as np
import pytest

from brownian_dynamics_cuda_amd import engine as E
from brownian_dynamics_cuda_amd import fiber as Fb

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MAX_M = 2112


def _lib():
    return E.load_library()


def _noalloc(D, mode=E.CHOL_FP64):
    return E.cholesky_device(D, mode, entry="bd_cholesky_noalloc")


def spd(m, seed):
    g = np.random.default_rng(seed)
    X = g.standard_normal((m, m)) / np.sqrt(m)
    return X @ X.T + np.diag(g.uniform(1.0, 2.0, m))


def test_two_calls_in_a_row_match_the_allocating_entry_point():
    print("case: bd_cholesky_noalloc, 100-nucleosome tensor then a 600-row matrix", flush=True)
    fib, st = Fb.make_config(2)
    with E.BDEngine(fib, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        D1 = eng.get_tensor()
    D2 = spd(600, 5)
    for D in (D1, D2):
        for mode in (E.CHOL_FP64, E.CHOL_FP32, E.CHOL_BF16):
            L, neg = _noalloc(D, mode)
            L0, neg0 = E.cholesky_device(D, mode)
            assert neg == neg0 == 0
            assert np.array_equal(L, L0), (D.shape, mode)


def test_negative_pivots_and_limits():
    print("case: bd_cholesky_noalloc, negative pivots and size limit", flush=True)
    z = np.load(os.path.join(GOLDEN, "negative_pivot.npz"))
    L, neg = _noalloc(z["D"])
    L0, neg0 = E.cholesky_device(z["D"])
    assert neg == neg0 == int(z["neg"]) == 6
    assert np.array_equal(L, L0)
    lib = _lib()
    big = E.DeviceBuffer(8)
    try:
        rc = lib.bd_cholesky_noalloc(MAX_M + 1, big.ptr, MAX_M + 1, big.ptr, MAX_M + 1, 0, None, None)
        assert rc == -4 and b"static workspace" in lib.bd_last_error()
        assert lib.bd_cholesky_noalloc(0, big.ptr, 1, big.ptr, 1, 0, None, None) == -1
    finally:
        big.free()
    # the largest size the array takes
    D = spd(MAX_M, 9)
    L, neg = _noalloc(D)
    L0, neg0 = E.cholesky_device(D)
    assert neg == neg0 == 0 and np.array_equal(L, L0)
