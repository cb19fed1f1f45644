"""The two largest benchmark fibers on the GPU: the 1000-nucleosome fiber
with linker histones (D 20982 x 20982, 50000 tail beads, 22000 free LH
beads) and the 500-nucleosome fiber (D 10482 x 10482), through the C-ABI.

  * frames and the mechanics, core/linker and tail force terms of the
    1000-nucleosome fiber against the oracle (each of those finishes in
    under half a minute on one host core; the oracle's linker-histone term,
    all pairs over 22000 LH beads, takes about nine minutes there, so the LH
    term is checked on a 300-nucleosome fiber with LH built by the same
    recipe);
  * the FP64 factor through L (L^T v) = D v on random vectors (an n^3
    product on the host is out of reach);
  * graph replay equal to eager launches bit for bit;
  * the precision forms (FP32 and three-slice BF16 trailing updates): the
    displacement of one step within 1e-6 of the FP64 step from the same
    state and noise, and ||L L^T - D|| / ||D|| (probed) within 1e-6;
  * the FP64 refinement of B xi at 500 nucleosomes against its formula and
    its covariance bound.
Reference: main_cuda (func_cuda.cu:3434-3680), func.cpp force terms."""
import numpy as np
import pytest

import oracle_py as O
from brownian_dynamics_cuda_amd import engine as E
from brownian_dynamics_cuda_amd import fiber as Fb
from decks import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def noise(fib, step=0, seed=5489):
    g = O.Rng(seed)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    for s in range(step + 1):
        rr, p = g.step_noise(s, fib.r_var(), nD3)
    return rr, p


def sync_engine(eng, fs):
    from brownian_dynamics_cuda_amd.system import State
    eng.set_state(State(fs.view("r"), fs.view("a"), fs.view("b"), fs.view("c"), fs.view("r_t"),
                        fs.view("r_lh")), rebuild=False)
    eng.set_derived({k: fs.view(k) for k in E.DERIVED_KEYS})


def probe_err(L, D, k=3, seed=3):
    g = np.random.default_rng(seed)
    e = 0.0
    for _ in range(k):
        v = g.standard_normal(D.shape[0])
        Dv = D @ v
        e = max(e, float(np.linalg.norm(L @ (L.T @ v) - Dv) / np.linalg.norm(Dv)))
    return e


@pytest.fixture(scope="module")
def cfg5():
    print("case: cfg5, 1000 nucleosomes with LH, D 20982 x 20982", flush=True)
    fib, st = Fb.make_config(5)
    orc = O.Oracle(fib)
    return fib, st, orc, orc.prepare(st)


def test_cfg5_frames(cfg5):
    fib, st, orc, fs = cfg5
    with E.BDEngine(fib, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        g, d = eng.get_state(), eng.get_derived()
    for k in ("a", "b", "c"):
        assert rel_err(getattr(g, k), fs.view(k)) <= 1e-14, k
    # the exit frames come through acos/sin of near-parallel directions
    # (measured 1.1e-13 on this 7000-bead chain)
    for k in ("length", "a_dna", "b_dna", "c_dna", "beta_t", "beta_lh"):
        assert rel_err(d[k], fs.view(k)) <= 1e-12, k


@pytest.mark.parametrize("term", ["mech", "elec", "tail"])
def test_cfg5_force_terms(cfg5, term):
    fib, st, orc, fs = cfg5
    code = {"mech": E.TERM_MECH, "elec": E.TERM_ELEC, "tail": E.TERM_TAIL}[term]
    print(f"case: cfg5 force term {term}", flush=True)
    with E.BDEngine(fib, use_graph=False) as eng:
        sync_engine(eng, fs)
        got = eng.forces(code)
    ref = orc.forces(fs, (term,))
    for k in ("F", "T", "tF", "lhF"):
        assert rel_err(got[k], ref[k]) <= TOL, k
    assert got["E"] == pytest.approx(ref["E"], rel=TOL)


def test_lh_term_300_nucleosomes_with_lh():
    print("case: 300-nucleosome fiber with LH, LH force term", flush=True)
    fib, st = Fb.make_fiber(300, seed=20261202 + 5, with_lh=True)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    with E.BDEngine(fib, use_graph=False) as eng:
        sync_engine(eng, fs)
        got = eng.forces(E.TERM_LH)
    ref = orc.forces(fs, ("lh",))
    for k in ("F", "T", "tF", "lhF"):
        assert rel_err(got[k], ref[k]) <= TOL, k
    assert got["E"] == pytest.approx(ref["E"], rel=TOL)


def test_cfg5_factor_and_graph_replay(cfg5):
    fib, st, _, _ = cfg5
    with E.BDEngine(fib, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        assert eng.chol_schedule() == 8
        D = eng.get_tensor()
        eng.factor()
        L, neg = eng.get_factor()
    assert neg == 0 and D.shape == (20982, 20982)
    e = probe_err(L, D)
    print(f"cfg5 FP64 factor: probe |L L^T v - D v| / |D v| = {e:.2e}", flush=True)
    assert e <= 1e-13
    del D, L
    steps = [noise(fib, s) for s in range(2)]
    out = {}
    for graph in (True, False):
        with E.BDEngine(fib, use_graph=graph) as eng:
            eng.set_state(st, rebuild=True)
            en = [eng.step(rr, p) for rr, p in steps]
            out[graph] = (en, eng.get_state())
    assert out[True][0] == out[False][0] and all(np.isfinite(out[True][0]))
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        np.testing.assert_array_equal(getattr(out[True][1], k), getattr(out[False][1], k), err_msg=k)


@pytest.mark.parametrize("cfg", [3, 5])
def test_precision_forms(cfg):
    print(f"case: precision forms at cfg{cfg}", flush=True)
    fib, st = Fb.make_config(cfg)
    rr, p = noise(fib)
    disp, ferr = {}, {}
    for mode in (E.CHOL_FP64, E.CHOL_FP32, E.CHOL_BF16):
        with E.BDEngine(fib, chol_mode=mode) as eng:
            eng.set_state(st, rebuild=True)
            x0 = eng.get_state()
            eng.step(rr, p)
            x1 = eng.get_state()
            disp[mode] = np.concatenate([x1.r - x0.r, x1.r_t - x0.r_t, x1.r_lh - x0.r_lh])
            if mode != E.CHOL_FP64:
                eng.set_state(st, rebuild=True)
                D = eng.get_tensor()
                eng.factor()
                L, neg = eng.get_factor()
                assert neg == 0
                ferr[mode] = probe_err(L, D)
                del D, L
    for mode in (E.CHOL_FP32, E.CHOL_BF16):
        rel = np.linalg.norm(disp[mode] - disp[E.CHOL_FP64]) / np.linalg.norm(disp[E.CHOL_FP64])
        print(f"cfg{cfg} mode {mode}: displacement {rel:.2e}, factor {ferr[mode]:.2e}", flush=True)
        assert rel <= 1e-6 and ferr[mode] <= 1e-6


@pytest.mark.parametrize("mode", [E.CHOL_FP32, E.CHOL_BF16])
def test_refined_noise_cfg3(mode):
    """The FP64 refinement of B xi at 500 nucleosomes (164 block columns:
    super-panels, the pre-split BF16 update, 82 two-block solve launches):
    the device noise against rd = s2dt (L~ p + D u) / 2, u = inv(L~)^T p, from
    the device factor and tensor, and the covariance of the refined map,
    ||B B^T v - D v|| / ||D v|| on random probes (B = L~ (I + X / 2),
    X = inv(L~) D inv(L~)^T - I, applied through triangular solves)."""
    from scipy.linalg import solve_triangular
    print(f"case: refined noise at cfg3, mode {mode}", flush=True)
    fib, st = Fb.make_config(3)
    with E.BDEngine(fib, chol_mode=mode, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        D = eng.get_tensor()
        eng.factor()
        L, neg = eng.get_factor()
        p = np.random.default_rng(17).standard_normal(3 * (fib.n + fib.n_tail + fib.n_LH))
        rd = eng.noise_displacement(p)
    assert neg == 0
    m = D.shape[0]
    u = solve_triangular(L.T, p[:m], lower=False)
    exp = fib.phys.s2dt * 0.5 * (L @ p[:m] + D @ u)
    e = rel_err(rd[:m], exp)
    g = np.random.default_rng(5)
    cov = 0.0
    for _ in range(2):
        v = g.standard_normal(m)
        y = L.T @ v
        bt = 0.5 * (y + solve_triangular(L, D @ solve_triangular(L.T, y, lower=False), lower=True))
        bbt = 0.5 * (L @ bt + D @ solve_triangular(L.T, bt, lower=False))
        Dv = D @ v
        cov = max(cov, float(np.linalg.norm(bbt - Dv) / np.linalg.norm(Dv)))
    print(f"  rd vs formula {e:.2e}, refined covariance {cov:.2e}, factor {probe_err(L, D):.2e}", flush=True)
    assert e <= 1e-11 and cov <= 1e-12
