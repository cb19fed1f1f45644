"""The oracle against the fixtures written from the compiled reference
(tests/golden/make_golden.py): frame update, every force term, the RPY
tensor and its factor, one or two full steps in both hydrodynamic modes,
and the host RNG.  Runs without the reference tree."""
import numpy as np
import pytest

import oracle_py as O
from decks import load_golden, rel_err

DECKS = ["testdata_deck", "tiny_deck"]


@pytest.mark.parametrize("name", DECKS)
def test_prepare_matches_reference(name):
    fib, st, g = load_golden(name)
    fs = O.Oracle(fib, "or").prepare(st)
    for k in O.STATE_FIELDS:
        a, b = fs.view(k), g[f"prep_{k}"]
        assert rel_err(a, b) <= 1e-13, k


@pytest.mark.parametrize("name", DECKS)
@pytest.mark.parametrize("term", ["mech", "elec", "tail", "lh", "all"])
def test_force_terms_match_reference(name, term):
    fib, st, g = load_golden(name)
    orc = O.Oracle(fib, "or")
    fs = orc.prepare(st)
    f = orc.forces(fs, (term,))
    for k in ("F", "T", "tF", "lhF"):
        assert rel_err(f[k], g[f"force_{term}_{k}"]) <= 1e-12, (term, k)
    assert f["E"] == pytest.approx(float(g[f"force_{term}_E"]), rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("name", DECKS)
def test_tensor_block_matches_reference(name):
    fib, st, g = load_golden(name)
    orc = O.Oracle(fib, "or")
    fs = orc.prepare(st)
    nb = min(20, fib.n)
    D, L, neg = orc.diffusion_tensor(fs.view("r")[:3 * nb], fib.rad[:nb])
    assert rel_err(D, g["D20"]) <= 1e-15
    assert rel_err(L, g["L20"]) <= 1e-13
    assert neg == 0


@pytest.mark.parametrize("name", DECKS)
@pytest.mark.parametrize("hi", [0, 1])
def test_steps_match_reference(name, hi):
    fib, st, g = load_golden(name)
    fib.hi_all = hi
    orc = O.Oracle(fib, "or")
    fs = orc.prepare(st)
    s = 0
    while f"step{hi}_{s}_E" in g:
        rr, p = g[f"noise_rr_{s}"], g[f"noise_p_{s}"]
        e = orc.step(fs, rr, p)["E"]
        assert e == pytest.approx(float(g[f"step{hi}_{s}_E"]), rel=1e-12)
        # positions and frames; Euler angles near acos(+-1) are compared
        # through the frames they are computed from
        for k in ("r", "a", "b", "c", "r_t", "r_lh", "length", "a_dna", "beta_t", "beta_lh"):
            assert rel_err(fs.view(k), g[f"step{hi}_{s}_{k}"]) <= 1e-11, (s, k)
        s += 1
    assert s >= 1


def test_overlap_tensor():
    z = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                           "overlap_tensor.npz"))
    from brownian_dynamics_cuda_amd import fiber as Fb
    fib, _ = Fb.make_fiber(3, seed=1)
    orc = O.Oracle(fib, "or")
    D, L, neg = orc.diffusion_tensor(z["r"], z["rad"])
    assert rel_err(D, z["D"]) <= 1e-15 and rel_err(L, z["L"]) <= 1e-13
    D2, L2, neg2 = orc.diffusion_tensor(z["r2"], z["rad2"])
    assert rel_err(D2, z["D2"]) <= 1e-15
    assert rel_err(L2, z["L2"]) <= 1e-12
    # eight overlapping equal beads: positive definite, no |f| pivot
    assert neg == 0 and neg2 == 0
    assert np.linalg.eigvalsh(D2).min() > 0


def test_negative_pivot_rule():
    """The |f| branch of the reference's Cholesky (func.cpp:825-835) on a
    129 x 129 tensor with six negative pivots in two tiles (two beads with a
    negative radius): the oracle's L and count against the reference's, and
    the closed form of the rule -- the factor is the exact Cholesky factor of
    D + 2 diag(|f_i|) over the flipped pivots i, so L L^T - D vanishes off
    the diagonal and equals 2 L_ii^2 on the flipped rows."""
    z = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                           "negative_pivot.npz"))
    from brownian_dynamics_cuda_amd import fiber as Fb
    fib, _ = Fb.make_fiber(7, seed=5)
    orc = O.Oracle(fib, "or")
    D, L, neg = orc.diffusion_tensor(z["r"], z["rad"])
    assert rel_err(D, z["D"]) <= 1e-15
    assert rel_err(L, z["L"]) <= 1e-13
    assert neg == int(z["neg"]) == 6
    flipped = np.concatenate([np.arange(3 * b, 3 * b + 3) for b in z["neg_beads"]])
    R = L @ L.T - D
    off = R - np.diag(np.diag(R))
    assert np.max(np.abs(off)) <= 1e-13 * np.max(np.abs(D))
    expect = np.zeros(D.shape[0])
    expect[flipped] = 2 * np.diag(L)[flipped] ** 2
    assert rel_err(np.diag(R), expect) <= 1e-12
    # the flipped pivots are exactly the negative diagonal blocks' rows
    assert np.all(np.diag(D)[flipped] < 0) and np.sum(np.diag(D) < 0) == 6


def test_rng_matches_reference():
    z = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                           "rng.npz"))
    g = O.Rng(5489)
    assert np.array_equal(np.array([g.real1() for _ in range(64)]), z["real1"])
    g = O.Rng(5489)
    got = np.array([g.normal(0.25, sd) for sd in z["normal_sd"]])
    assert np.array_equal(got, z["normal"])
