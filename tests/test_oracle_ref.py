"""The oracle against the compiled reference (oracle/_ref/libbdref.so, built
from the reference's own func.cpp / mt.cpp / utilities.cpp) on seeded
decks.  Skipped where the reference library was not built."""
import numpy as np
import pytest

import oracle_py as O
from brownian_dynamics_cuda_amd import fiber as Fb
from decks import deck, rel_err

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")


@pytest.mark.parametrize("kind", ["tiny", "small", "cfg1"])
def test_forces_equal_reference(kind):
    fib, st = deck(kind)
    fib.groups_per_core = 10  # the reference's literal
    a, b = O.Oracle(fib, "or"), O.Oracle(fib, "ref")
    fa, fb = a.prepare(st), b.prepare(st)
    for k in O.STATE_FIELDS:
        assert np.array_equal(fa.view(k), fb.view(k)), k
    for term in ("mech", "elec", "tail", "lh"):
        x, y = a.forces(fa, (term,)), b.forces(fb, (term,))
        for k in ("F", "T", "tF", "lhF"):
            assert rel_err(x[k], y[k]) <= 1e-13, (term, k)
        assert x["E"] == pytest.approx(y["E"], rel=1e-13)


@pytest.mark.parametrize("hi", [0, 1])
def test_three_steps_equal_reference(hi):
    fib, st = deck("small", hi_all=hi)
    fib.groups_per_core = 10
    a, b = O.Oracle(fib, "or"), O.Oracle(fib, "ref")
    fa, fb = a.prepare(st), b.prepare(st)
    rng = O.Rng(5489)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    for s in range(3):
        rr, p = rng.step_noise(s, fib.r_var(), nD3)
        ea, eb = a.step(fa, rr, p)["E"], b.step(fb, rr, p)["E"]
        assert ea == pytest.approx(eb, rel=1e-12)
        for k in ("r", "a", "b", "c", "r_t", "r_lh"):
            assert rel_err(fa.view(k), fb.view(k)) <= 1e-11, (s, k)
        fa = fb.copy()  # re-synchronise on the reference's state


def test_rng_equals_reference():
    lib = O.ref_lib()
    lib.ref_init_genrand(20261202)
    g = O.Rng(20261202)
    assert all(g.real1() == lib.ref_genrand_real1() for _ in range(2000))
    sds = np.abs(np.random.default_rng(1).normal(size=501)) + 0.1
    assert all(g.normal(0.0, sd) == lib.ref_rand_normal(0.0, sd) for sd in sds)
