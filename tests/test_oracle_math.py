"""Checks of the oracle that the mathematics fixes independently of the
reference: RPY tensor properties and its closed form, Cholesky against
LAPACK and the |f| pivot rule in closed form, Rodrigues rotation against
scipy, force = -grad(energy) where no clamp acts, conservation of the total
force, and the MT19937 known answer."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_py as O
from brownian_dynamics_cuda_amd import fiber as Fb
from brownian_dynamics_cuda_amd import system as S
from decks import deck, rel_err


def rpy_textbook(ri, rj, ai, aj, kT_over_eta):
    """Rotne-Prager-Yamakawa mobility of two non-overlapping spheres of
    unequal radii (Rotne & Prager 1969; Garcia de la Torre & Bloomfield
    1977): kT/(8 pi eta r) [(1 + (ai^2+aj^2)/(3 r^2)) I + (1 - (ai^2+aj^2)/r^2) rr/r^2]."""
    d = ri - rj
    r = np.linalg.norm(d)
    u = np.outer(d, d) / r ** 2
    q = (ai * ai + aj * aj) / r ** 2
    return kT_over_eta / (8 * math.pi * r) * ((1 + q / 3) * np.eye(3) + (1 - q) * u)


def test_rpy_tensor_properties_and_closed_form():
    fib, st = deck("cfg1")
    orc = O.Oracle(fib)
    nb = 30
    D, L, neg = orc.diffusion_tensor(st.r[:3 * nb], fib.rad[:nb])
    assert np.array_equal(D, D.T)
    kT_eta = fib.phys.a2 * 8 * math.pi  # a2 = kT / (8 pi eta)
    r = st.r[:3 * nb].reshape(-1, 3)
    for i in range(nb):
        assert np.allclose(D[3 * i:3 * i + 3, 3 * i:3 * i + 3], fib.phys.a1 / fib.rad[i] * np.eye(3),
                           rtol=1e-15, atol=0)
        for j in range(i + 1, nb):
            if np.linalg.norm(r[i] - r[j]) >= fib.rad[i] + fib.rad[j]:
                ref = rpy_textbook(r[i], r[j], fib.rad[i], fib.rad[j], kT_eta)
                assert rel_err(D[3 * i:3 * i + 3, 3 * j:3 * j + 3], ref) < 1e-13
    assert np.linalg.eigvalsh(D).min() > 0
    assert neg == 0


def test_rpy_overlap_equal_spheres_textbook():
    """Equal overlapping spheres: the effective radius is the radius and the
    reference formula is the RPY overlap form (1 - 9r/32a) I + 3/(32a) rr/r."""
    fib, _ = deck("tiny")
    orc = O.Oracle(fib)
    a, rr = 1.5, np.array([0.0, 0, 0, 1.1, 0.7, -0.4])
    D, _, _ = orc.diffusion_tensor(rr, np.array([a, a]))
    d = rr[:3] - rr[3:]
    r = np.linalg.norm(d)
    ref = fib.phys.a1 / a * ((1 - 9 * r / (32 * a)) * np.eye(3) + 3 / (32 * a) * np.outer(d, d) / r)
    assert rel_err(D[:3, 3:], ref) < 1e-14


def test_cholesky_matches_lapack():
    fib, st = deck("cfg1")
    orc = O.Oracle(fib)
    D, L, neg = orc.diffusion_tensor(st.r, fib.rad)
    assert neg == 0
    assert rel_err(L, np.linalg.cholesky(D)) < 1e-12
    assert rel_err(L @ L.T, D) < 1e-14


def test_negative_pivot_rule_closed_form():
    """A 2x2 indefinite matrix: the reference takes |f| at the second pivot
    (func.cpp:827-833), so L = [[sqrt a, 0], [b/sqrt a, sqrt(b^2/a - c)]]."""
    lib = O.oracle_lib()
    a, b, c = 4.0, 6.0, 1.0
    D = np.array([[a, b], [b, c]])
    L = np.zeros((2, 2))
    neg = lib.or_cholesky(2, D.ctypes.data_as(C.POINTER(C.c_double)),
                          L.ctypes.data_as(C.POINTER(C.c_double)))
    assert neg == 1
    ref = np.array([[math.sqrt(a), 0], [b / math.sqrt(a), math.sqrt(b * b / a - c)]])
    assert rel_err(L, ref) < 1e-15


def test_rotation_matches_rodrigues():
    """rotate() turns the body frame by the angle |w| about the body-frame
    axis w: new frame vectors are R(w) applied in body coordinates."""
    lib = O.oracle_lib()
    rng = np.random.default_rng(5)
    R0 = Rotation.random(random_state=7).as_matrix()
    a, b, c = R0[:, 0].copy(), R0[:, 1].copy(), R0[:, 2].copy()
    for _ in range(5):
        w = rng.normal(size=3) * 0.3
        out = [np.zeros(3) for _ in range(3)]
        p = lambda x: np.ascontiguousarray(x).ctypes.data_as(C.POINTER(C.c_double))
        wd = np.ascontiguousarray(w)
        lib.or_rotate(1, p(a), p(b), p(c), p(out[0]), p(out[1]), p(out[2]), p(wd), 1.0)
        Rb = Rotation.from_rotvec(w).as_matrix()   # rotation in body coordinates
        F = np.stack([a, b, c], 1)                  # columns a, b, c
        new = F @ Rb                               # a_j' = sum_k Rb[k,j] e_k
        for k in range(3):
            assert np.allclose(out[k], new[:, k], atol=1e-14)
        G = np.stack(out, 1)
        assert np.allclose(G.T @ G, np.eye(3), atol=1e-14)
        assert np.linalg.det(G) == pytest.approx(1.0, abs=1e-14)
    zero = [np.zeros(3) for _ in range(3)]
    lib.or_rotate(1, p(a), p(b), p(c), p(zero[0]), p(zero[1]), p(zero[2]), p(np.zeros(3)), 1.0)
    assert np.array_equal(zero[0], a) and np.array_equal(zero[2], c)


def _elec_energy_force(fib, st_r, st_a, st_b, st_c):
    orc = O.Oracle(fib)
    fs = O.FullState(fib, S.State(st_r, st_a, st_b, st_c, np.zeros(0), np.zeros(0)))
    orc.update_euler(fs)
    f = orc.forces(fs, ("elec",))
    return f["E"], f["F"]


def test_core_terms_force_is_minus_energy_gradient():
    """The core/linker electrostatics (func.cpp:1291-1605) have no clamp, so
    F = -dE/dr for every bead position (central differences, h = 1e-4)."""
    fib, st = Fb.make_fiber(3, seed=4, with_tails=False)
    E0, F0 = _elec_energy_force(fib, st.r, st.a, st.b, st.c)
    assert abs(E0) > 0
    h = 1e-4
    checked = 0
    for i in range(fib.n):
        for k in range(3):
            if abs(F0[3 * i + k]) < 1e-4 * np.max(np.abs(F0)):
                continue
            rp, rm = st.r.copy(), st.r.copy()
            rp[3 * i + k] += h
            rm[3 * i + k] -= h
            Ep, _ = _elec_energy_force(fib, rp, st.a, st.b, st.c)
            Em, _ = _elec_energy_force(fib, rm, st.a, st.b, st.c)
            g = -(Ep - Em) / (2 * h)
            assert g == pytest.approx(F0[3 * i + k], rel=2e-5), (i, k)
            checked += 1
    assert checked > 10


@pytest.mark.parametrize("kind", ["tiny", "cfg1"])
def test_total_force_is_conserved(kind):
    """Every term applies equal and opposite forces (the +-0.9 clamp is odd,
    the anchored-tail redirect moves a force to a core, the x10 LH anchor
    acts on both ends): the sum over all beads vanishes."""
    fib, st = deck(kind)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    f = orc.forces(fs)
    tot = f["F"].reshape(-1, 3).sum(0) + f["tF"].reshape(-1, 3).sum(0) + f["lhF"].reshape(-1, 3).sum(0)
    scale = np.abs(f["F"]).sum() + np.abs(f["tF"]).sum() + np.abs(f["lhF"]).sum()
    assert np.max(np.abs(tot)) < 1e-13 * scale


def test_twist_angle_table_and_formula():
    lib = O.oracle_lib()
    assert lib.or_twist_angle_phi(6) == 0.0215
    for w in (12, 15, 20):
        nbp = (w + 1) * 3 / 0.34
        turns = round(nbp / 10.3)
        expect = math.pi * (nbp * (turns * 360 / nbp - 34.95) / (w + 1)) / 180.0
        assert lib.or_twist_angle_phi(w) == pytest.approx(expect, rel=1e-15)


def test_mt19937_known_answer():
    """The 10000th output of MT19937 seeded with 5489 is 4123659995
    (C++ standard, [rand.predef])."""
    g = O.Rng(5489)
    for _ in range(9999):
        g.u32()
    assert g.u32() == 4123659995


def test_polar_normals_statistics():
    g = O.Rng(1234)
    z = np.array([g.normal(0.0, 1.0) for _ in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.02
