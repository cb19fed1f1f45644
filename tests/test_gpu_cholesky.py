"""bd_cholesky: the library's blocked factor on a caller's device matrix
(the replacement of cusolverDnDpotrf, func_cuda.cu:3513-3534), through the
C-ABI, against the compiled reference's factor (the |f| rule of
func.cpp:825-835), LAPACK, and the factor the step itself takes."""
import os

import numpy as np
import pytest

from brownian_dynamics_cuda_amd import engine as E
from brownian_dynamics_cuda_amd import fiber as Fb
from decks import rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def spd(m, seed):
    """A well-conditioned SPD matrix: RPY-like diagonal dominance."""
    g = np.random.default_rng(seed)
    X = g.standard_normal((m, m)) / np.sqrt(m)
    return X @ X.T + np.diag(g.uniform(1.0, 2.0, m))


def test_negative_pivots_against_reference():
    z = np.load(os.path.join(GOLDEN, "negative_pivot.npz"))
    L, neg = E.cholesky_device(z["D"])
    assert neg == int(z["neg"]) == 6
    assert rel_err(L, z["L"]) <= 1e-10


@pytest.mark.parametrize("m", [1, 3, 64, 65, 200, 1031])
def test_sizes_against_lapack(m):
    D = spd(m, m)
    L, neg = E.cholesky_device(D)
    assert neg == 0
    assert rel_err(L, np.linalg.cholesky(D)) <= 1e-12
    assert np.all(np.triu(L, 1) == 0)


def test_super_panel_size_and_mixed_modes():
    """4500 rows: 71 block columns, factored in super-panels of 8."""
    D = spd(4500, 7)
    L0 = np.linalg.cholesky(D)
    L, neg = E.cholesky_device(D)
    assert neg == 0 and rel_err(L, L0) <= 1e-12
    for mode in (E.CHOL_FP32, E.CHOL_BF16):
        Lm, neg = E.cholesky_device(D, mode)
        err = np.linalg.norm(Lm @ Lm.T - D) / np.linalg.norm(D)
        assert neg == 0 and err <= 1e-6, (mode, err)


def test_same_factor_as_the_step():
    """The tensor of the 100-nucleosome fiber: bd_cholesky on it equals the
    factor the step forms in place (same kernels, same arithmetic)."""
    fib, st = Fb.make_config(2)
    with E.BDEngine(fib, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        D = eng.get_tensor()
        eng.factor()
        L0, neg0 = eng.get_factor()
    L, neg = E.cholesky_device(D)
    assert neg == neg0 == 0
    assert np.array_equal(L, L0)


def test_in_place_and_argument_errors():
    lib = E.load_library()
    D = spd(130, 3)
    buf = E.DeviceBuffer(D.nbytes)
    try:
        buf.upload(D)
        import ctypes as C
        neg = C.c_int(-1)
        assert lib.bd_cholesky(130, buf.ptr, 130, buf.ptr, 130, E.CHOL_FP64, C.byref(neg), None) == 0
        L = buf.download((130, 130))
        assert neg.value == 0 and rel_err(L, np.linalg.cholesky(D)) <= 1e-12
        assert lib.bd_cholesky(0, buf.ptr, 130, buf.ptr, 130, 0, None, None) == -1
        assert lib.bd_cholesky(130, buf.ptr, 100, buf.ptr, 130, 0, None, None) == -1
        assert lib.bd_cholesky(130, buf.ptr, 130, buf.ptr, 130, 7, None, None) == -1
    finally:
        buf.free()
