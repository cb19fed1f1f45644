"""The host driver, batched steps and frequency_RP reuse on the GPU.

- bd_run_steps (k steps per call) against k single bd_step calls: same
  kernels, same order, so bit-identical states and energies.
- The Python driver over the reference's input deck: the first energy
  against the compiled reference's energy of the deck state, linker-histone
  moves on the GPU, and a checkpoint/resume that continues bit for bit.
- frequency_RP: 1 reproduces the default step exactly; k > 1 rebuilds D and
  L on steps 0, k, 2k, ... and reuses L in between (graph and eager launches
  agree bit for bit); the displacement error of the reuse is measured.
"""
import dataclasses
import os
import shutil

import numpy as np
import pytest

import oracle_py as O
from brownian_dynamics_cuda_amd import driver as D
from brownian_dynamics_cuda_amd import engine as E
from brownian_dynamics_cuda_amd.system import State
from decks import GOLDEN, deck, load_golden

pytestmark = pytest.mark.gpu

DECK = os.path.join(GOLDEN, "testdata")


def make_deck(tmp, **setup):
    d = os.path.join(str(tmp), "deck")
    os.makedirs(d, exist_ok=True)
    for f in os.listdir(DECK):
        shutil.copy(os.path.join(DECK, f), d)
        os.chmod(os.path.join(d, f), 0o644)
    lines = []
    for line in open(os.path.join(d, "setup.txt")):
        t = line.split()
        if t and t[0] in setup:
            line = f"{t[0]} {setup[t[0]]}\n"
        lines.append(line)
    open(os.path.join(d, "setup.txt"), "w").writelines(lines)
    return d


def noise(fib, k, seed=5489):
    rng = E.HostRng(seed)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    out = [rng.step_noise(s, fib.r_var(), nD3) for s in range(k)]
    return np.array([o[0] for o in out]), np.array([o[1] for o in out])


def test_run_steps_equals_single_steps():
    fib, st = deck("small", hi_all=1)
    rr, p = noise(fib, 7)
    print("case: small fiber, 7 steps, bd_run_steps vs bd_step", flush=True)
    with E.BDEngine(fib) as a, E.BDEngine(fib) as b:
        a.set_state(st)
        b.set_state(st)
        ea = a.run_steps(rr, p)
        eb = np.array([b.step(rr[s], p[s]) for s in range(7)])
        sa, sb = a.get_state(), b.get_state()
    np.testing.assert_array_equal(ea, eb)
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        np.testing.assert_array_equal(getattr(sa, k), getattr(sb, k), err_msg=k)


def test_driver_on_reference_deck(tmp_path):
    d = make_deck(tmp_path, number_of_steps=210, frequency_of_sampling=50, unbind=0.5, Pbind=0.5)
    print("case: Python driver, reference deck, 210 steps, seed 5", flush=True)
    run = D.Run(d, out=str(tmp_path / "out"), seed=5)
    try:
        res = run.run()
    finally:
        run.close()
    En = np.loadtxt(tmp_path / "out" / "Energy.txt")
    assert En.shape == (210,) and np.all(np.isfinite(En))
    # step 0's energy follows the move made before it (main.cpp:386-465):
    # the oracle's energy of the deck state with that binding
    fib0, st0, _, _ = D.load_deck(d)
    flags0 = np.array(fib0.nc_lh_flag, np.int32)
    flags1 = flags0.copy()
    for step, core, bound in res["lh_events"]:
        if step == 0:
            flags1[core] = bound
    r_lh, lh_rad = D.rebind_positions(fib0, flags0, flags1, st0.r_lh,
                                      (st0.r, st0.a, st0.b, st0.c))
    fib1 = dataclasses.replace(fib0, nc_lh_flag=flags1, lh_rad=lh_rad)
    st1 = State(r=st0.r, a=st0.a, b=st0.b, c=st0.c, r_t=st0.r_t, r_lh=r_lh)
    orc = O.Oracle(fib1, "or")
    E0 = orc.forces(orc.prepare(st1))["E"]
    # Energy.txt carries 6 significant digits (main.cpp:494)
    assert En[0] == pytest.approx(E0, rel=5e-6)
    if np.array_equal(flags0, flags1):
        _, _, g = load_golden("testdata_deck")
        assert En[0] == pytest.approx(float(g["force_all_E"]), rel=5e-6)
    # three Monte Carlo moves (steps 0, 100, 200); the final binding is written
    flags = [int(x) for x in open(tmp_path / "out" / "LH.in")]
    assert flags == res["nc_lh_flag"]
    nlh = run.fib.n_lh_n + run.fib.n_lh_c
    r_lh = np.loadtxt(tmp_path / "out" / "restart_lh.txt")
    assert r_lh.shape == (3 * nlh * sum(flags),) and np.all(np.isfinite(r_lh))
    frames = open(tmp_path / "out" / "out.xyz").read().count("comments zl3765")
    assert frames == 5  # steps 0, 50, 100, 150, 200
    print(f"driver: {res['steps_per_s']:.1f} steps/s, moves {res['lh_events']}", flush=True)


def test_driver_checkpoint_resume_is_exact(tmp_path):
    # sampling at 0, 33, 66, 99: the host frame copy at the move of step 100
    # is the state after step 99 in both runs
    d = make_deck(tmp_path, number_of_steps=150, frequency_of_sampling=33, unbind=1.0, Pbind=1.0)
    print("case: Python driver, 150 steps straight vs 100 + resume 50", flush=True)
    for name, kw in (("a", {}), ("b", dict(steps=100))):
        run = D.Run(d, out=str(tmp_path / name), seed=21, **kw)
        try:
            run.run()
        finally:
            run.close()
    run = D.Run(d, out=str(tmp_path / "b"), resume=str(tmp_path / "b" / D.CHECKPOINT_FILE))
    try:
        res = run.run()
    finally:
        run.close()
    assert res["steps"] == 50
    a = np.load(tmp_path / "a" / D.CHECKPOINT_FILE)
    b = np.load(tmp_path / "b" / D.CHECKPOINT_FILE)
    for k in ("r", "a", "b", "c", "r_t", "r_lh", "nc_lh_flag", "rng_words"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert open(tmp_path / "a" / "Energy.txt").read() == open(tmp_path / "b" / "Energy.txt").read()


def test_frequency_rp_one_is_the_default_step():
    fib, st = deck("small", hi_all=1)
    rr, p = noise(fib, 4)
    print("case: small fiber, frequency_RP 1 vs default", flush=True)
    with E.BDEngine(fib) as a, E.BDEngine(fib) as b:
        a.set_state(st)
        b.set_state(st)
        b.set_frequency_rp(1)
        ea, eb = a.run_steps(rr, p), b.run_steps(rr, p)
        sa, sb = a.get_state(), b.get_state()
    np.testing.assert_array_equal(ea, eb)
    np.testing.assert_array_equal(sa.r, sb.r)


@pytest.mark.parametrize("hi_all", [1, 0])
def test_frequency_rp_reuses_the_factor(hi_all):
    fib, st = deck("small", hi_all=hi_all)
    k = 5
    rr, p = noise(fib, 2 * k)
    print(f"case: small fiber hi_all={hi_all}, frequency_RP {k}, graph vs eager vs fresh",
          flush=True)
    with E.BDEngine(fib) as g, E.BDEngine(fib, use_graph=False) as e, E.BDEngine(fib) as f:
        for x in (g, e, f):
            x.set_state(st)
        g.set_frequency_rp(k)
        e.set_frequency_rp(k)
        x0 = f.get_state()
        errs = []
        for s in range(2 * k):
            eg, ee = g.step(rr[s], p[s]), e.step(rr[s], p[s])
            f.step(rr[s], p[s])
            sg, se, sf = g.get_state(), e.get_state(), f.get_state()
            # the captured reuse step and the eager one are the same kernels
            for key in ("r", "a", "r_t", "r_lh"):
                np.testing.assert_array_equal(getattr(sg, key), getattr(se, key), err_msg=key)
            assert eg == ee
            dg = np.concatenate([sg.r - x0.r, sg.r_t - x0.r_t, sg.r_lh - x0.r_lh])
            df = np.concatenate([sf.r - x0.r, sf.r_t - x0.r_t, sf.r_lh - x0.r_lh])
            errs.append(float(np.linalg.norm(dg - df) / np.linalg.norm(df)))
    print("frequency_RP 5, rel displacement from x0 vs rebuilding every step:",
          ["%.2e" % v for v in errs], flush=True)
    assert errs[0] == 0.0          # step 0 builds D and L in both
    assert errs[1] > 0.0           # step 1 reuses the factor of step 0
    assert all(np.isfinite(errs)) and max(errs) < 0.5
