"""FP64 refinement of the noise displacement in the reduced-precision factor
modes (bd_set_refine), through the C-ABI.

With L~ the FP32 / three-slice BF16 factor (L~ L~^T = D + E), the library
forms rd = s2dt (L~ p + D u) / 2 for the hydrodynamic rows, u = inv(L~)^T p
(B p for B = L~ (I + X / 2), X = inv(L~) D inv(L~)^T - I, since
L~ X p = D u - L~ p).  Checked here:
  * the device path against the same formula in numpy/scipy (a triangular
    solve and a dense D u from the tensor the library assembled), on the
    one-block-column schedule and on super-panels (k_step, k_panel,
    k_update, k_trail all park their tiles where the solve reads them);
  * what the formula fixes, from the device factor: B = L~ (I + X / 2) with
    X = inv(L~) D inv(L~)^T - I gives ||B B^T - D|| / ||D|| <= 1e-12, against
    ~1e-7 for L~ itself, so the refined noise has D's covariance to FP64
    accuracy;
  * refinement off gives s2dt L~ p exactly as before, and FP64 mode ignores
    the switch.
Reference: the FP64 factor and rd_cal (func_cuda.cu:3513-3552, 634-647),
which this replaces at reduced precision."""
import numpy as np
import pytest
from scipy.linalg import solve_triangular

import oracle_py as O
from brownian_dynamics_cuda_amd import engine as E
from brownian_dynamics_cuda_amd import fiber as Fb
from decks import deck, rel_err

pytestmark = pytest.mark.gpu


def _noise(fib, seed=11):
    g = np.random.default_rng(seed)
    return g.standard_normal(3 * (fib.n + fib.n_tail + fib.n_LH))


def _expected(fib, L, D, p, refine):
    m = D.shape[0]
    out = np.empty_like(p)
    out[:m] = L @ p[:m]
    if refine:
        u = solve_triangular(L.T, p[:m], lower=False)
        out[:m] = 0.5 * (out[:m] + D @ u)
    out[:m] *= fib.phys.s2dt
    rad = np.repeat(fib.rad_all(), 3)
    out[m:] = fib.phys.s2dt * np.sqrt(fib.phys.a1 / rad[m:]) * p[m:]
    return out


def _device(fib, st, mode, refine, panel=0):
    with E.BDEngine(fib, chol_mode=mode, use_graph=False) as eng:
        eng.set_chol_schedule(panel)
        eng.set_refine(refine)
        eng.set_state(st, rebuild=True)
        D = eng.get_tensor()
        eng.factor()
        L, neg = eng.get_factor()
        assert neg == 0
        p = _noise(fib)
        rd = eng.noise_displacement(p)
    return D, L, p, rd


CASES = [("cfg2", E.CHOL_FP32, 0), ("cfg2", E.CHOL_BF16, 0), ("small-hi", E.CHOL_FP32, 2),
         ("small-hi", E.CHOL_BF16, 3)]


def _fiber(kind):
    if kind == "cfg2":
        return Fb.make_config(2)
    return deck("small", hi_all=1)


@pytest.mark.parametrize("kind,mode,panel", CASES)
def test_refined_noise_matches_formula(kind, mode, panel):
    print(f"case: refine {kind} mode {mode} panel {panel}", flush=True)
    fib, st = _fiber(kind)
    D, L, p, rd = _device(fib, st, mode, True, panel)
    exp = _expected(fib, L, D, p, True)
    e = rel_err(rd, exp)
    # a triangular solve and two products in FP64, summed in another order
    print(f"  refined rd vs formula: {e:.2e}")
    assert e <= 1e-11
    # the covariance the refined noise carries: B B^T against D
    m = D.shape[0]
    Li = solve_triangular(L, np.eye(m), lower=True)
    X = Li @ D @ Li.T - np.eye(m)
    B = L @ (np.eye(m) + 0.5 * X)
    err_b = np.linalg.norm(B @ B.T - D) / np.linalg.norm(D)
    err_l = np.linalg.norm(L @ L.T - D) / np.linalg.norm(D)
    print(f"  |B B^T - D| / |D| = {err_b:.2e}, |L L^T - D| / |D| = {err_l:.2e}")
    assert err_b <= 1e-12 and err_l >= 1e3 * err_b
    # and the refined displacement stays within 1e-6 of the FP64 factor's
    L64 = np.linalg.cholesky(D)
    rd64 = _expected(fib, L64, D, p, False)
    assert rel_err(rd, rd64) <= 1e-6


def test_refine_off_and_fp64_mode():
    fib, st = Fb.make_config(2)
    D, L, p, rd = _device(fib, st, E.CHOL_FP32, False)
    assert np.array_equal(rd, _device(fib, st, E.CHOL_FP32, False)[3])
    assert rel_err(rd, _expected(fib, L, D, p, False)) <= 1e-13
    r64 = {on: _device(fib, st, E.CHOL_FP64, on)[3] for on in (False, True)}
    assert np.array_equal(r64[False], r64[True])


def test_refined_step_graph_equals_eager():
    """A step in FP32 mode with refinement, graph replay against eager
    launches, bit for bit."""
    fib, st = Fb.make_config(2)
    rr, p = O.Rng(4).step_noise(0, fib.r_var(), 3 * (fib.n + fib.n_tail + fib.n_LH))
    out = {}
    for graph in (True, False):
        with E.BDEngine(fib, chol_mode=E.CHOL_FP32, use_graph=graph) as eng:
            eng.set_state(st, rebuild=True)
            e = [eng.step(rr, p) for _ in range(2)]
            out[graph] = (e, eng.get_state())
    assert out[True][0] == out[False][0]
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        np.testing.assert_array_equal(getattr(out[True][1], k), getattr(out[False][1], k), err_msg=k)


def test_refine_switched_on_after_the_factor():
    """The positions the refinement reads are kept only by an assembly made
    while a refinement is on: switching the FP32 mode (from FP64) or the
    refinement on after the factor asks for a new factor, which then gives
    the refined noise of a context that had it on from the start."""
    fib, st = Fb.make_config(2)
    ref = _device(fib, st, E.CHOL_FP32, True)[3]
    p = _noise(fib)
    for start_mode, start_refine in ((E.CHOL_FP64, True), (E.CHOL_FP32, False)):
        print(f"case: refine switched on, from mode {start_mode} refine {start_refine}", flush=True)
        with E.BDEngine(fib, chol_mode=start_mode, use_graph=False) as eng:
            eng.set_refine(start_refine)
            eng.set_state(st, rebuild=True)
            eng.factor()
            eng.set_chol_mode(E.CHOL_FP32)
            eng.set_refine(True)
            with pytest.raises(E.BDError, match="bd_factor again"):
                eng.noise_displacement(p)
            eng.factor()
            np.testing.assert_array_equal(eng.noise_displacement(p), ref)


def test_step_after_refine_switched_on_forms_a_new_factor():
    """Under frequency_RP the step after the refinement is switched on forms
    D and L again (bit for bit a step whose factor was forced fresh by a mode
    round trip), in the captured graph."""
    fib, st = Fb.make_config(2)
    g = O.Rng(6)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    noise = [g.step_noise(s, fib.r_var(), nD3) for s in range(3)]
    out = []
    for force in (False, True):
        with E.BDEngine(fib, chol_mode=E.CHOL_FP32, use_graph=True) as eng:
            eng.set_refine(False)
            eng.set_frequency_rp(5)
            eng.set_state(st, rebuild=True)
            for rr, p in noise[:2]:
                eng.step(rr, p)
            eng.set_refine(True)
            if force:
                eng.set_chol_mode(E.CHOL_FP64)
                eng.set_chol_mode(E.CHOL_FP32)
            e = eng.step(*noise[2])
            out.append((e, eng.get_state()))
    assert out[0][0] == out[1][0]
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        np.testing.assert_array_equal(getattr(out[0][1], k), getattr(out[1][1], k), err_msg=k)
