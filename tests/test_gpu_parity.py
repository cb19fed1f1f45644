"""Parity of the HIP step (libbdhip.so, through its C-ABI) with the CPU
oracle and with the fixtures written from the compiled reference.

Tolerances (FP64): forces, torques, D, L, rd and positions within 1e-10
relative (max-norm, relative to the largest entry); the measured agreement
is ~1e-14.  Euler angles are compared through the frames they come from,
because acos() near +-1 amplifies any rounding difference to ~1e-8."""
import numpy as np
import pytest

import oracle_py as O
from brownian_dynamics_cuda_amd import engine as E
from brownian_dynamics_cuda_amd import fiber as Fb
from decks import deck, load_golden, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
TERMS = {"mech": E.TERM_MECH, "elec": E.TERM_ELEC, "tail": E.TERM_TAIL, "lh": E.TERM_LH,
         "all": E.TERM_ALL}
DERIVED = E.DERIVED_KEYS


def sync_engine(eng, fs):
    """Put the oracle's state (positions, frames and derived arrays) on the GPU."""
    from brownian_dynamics_cuda_amd.system import State
    eng.set_state(State(fs.view("r"), fs.view("a"), fs.view("b"), fs.view("c"), fs.view("r_t"),
                        fs.view("r_lh")), rebuild=False)
    eng.set_derived({k: fs.view(k) for k in DERIVED})


def noise(fib, step=0, seed=5489):
    g = O.Rng(seed)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    for s in range(step + 1):
        rr, p = g.step_noise(s, fib.r_var(), nD3)
    return rr, p


@pytest.mark.parametrize("kind", ["tiny", "small", "cfg1"])
def test_frame_update(kind):
    fib, st = deck(kind)
    fs = O.Oracle(fib).prepare(st)
    with E.BDEngine(fib, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        g = eng.get_state()
        d = eng.get_derived()
    for k in ("a", "b", "c"):
        assert rel_err(getattr(g, k), fs.view(k)) <= 1e-14, k
    for k in ("length", "a_dna", "b_dna", "c_dna", "beta_t", "beta_lh"):
        assert rel_err(d[k], fs.view(k)) <= 1e-13, k


@pytest.mark.parametrize("n_c,linker_beads", [(3, 70), (2, 1100)])
def test_frame_update_uncut_chain(n_c, linker_beads):
    """Chains with more than 64 beads between cores cannot be cut into
    frame segments: the frames then take the single-workgroup launch (up to
    1025 beads) or the two pass launches, against the oracle as above."""
    print(f"case: frames of an uncut chain, {n_c} cores, {linker_beads} linker beads", flush=True)
    fib, st = Fb.make_fiber(n_c, seed=7, linker_beads=linker_beads)
    fs = O.Oracle(fib).prepare(st)
    with E.BDEngine(fib, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        g = eng.get_state()
        d = eng.get_derived()
    for k in ("a", "b", "c"):
        assert rel_err(getattr(g, k), fs.view(k)) <= 1e-14, k
    for k in ("length", "a_dna", "b_dna", "c_dna", "beta_t", "beta_lh"):
        assert rel_err(d[k], fs.view(k)) <= 1e-13, k


@pytest.mark.parametrize("kind", ["tiny", "small", "cfg1"])
@pytest.mark.parametrize("term", list(TERMS))
def test_force_terms(kind, term):
    fib, st = deck(kind)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    with E.BDEngine(fib, use_graph=False) as eng:
        sync_engine(eng, fs)
        got = eng.forces(TERMS[term])
    ref = orc.forces(fs, (term,))
    for k in ("F", "T", "tF", "lhF"):
        assert rel_err(got[k], ref[k]) <= TOL, k
    assert got["E"] == pytest.approx(ref["E"], rel=TOL, abs=1e-300)


@pytest.mark.parametrize("name", ["testdata_deck", "tiny_deck"])
def test_against_reference_fixtures(name):
    """The HIP forces and steps against outputs of the reference's own
    func.cpp (tests/golden)."""
    fib, st, g = load_golden(name)
    for hi in (0, 1):
        fib.hi_all = hi
        fs = O.Oracle(fib).prepare(st)
        with E.BDEngine(fib, use_graph=False) as eng:
            eng.set_state(st, rebuild=True)
            d = eng.get_derived()
            for k in ("length", "a_dna", "beta_t", "beta_lh"):
                assert rel_err(d[k], g[f"prep_{k}"]) <= 1e-13, k
            sync_engine(eng, fs)
            if hi == 0:
                for term in ("mech", "elec", "tail", "lh", "all"):
                    f = eng.forces(TERMS[term])
                    for k in ("F", "T", "tF", "lhF"):
                        assert rel_err(f[k], g[f"force_{term}_{k}"]) <= TOL, (term, k)
                    assert f["E"] == pytest.approx(float(g[f"force_{term}_E"]), rel=TOL)
            e = eng.step(g["noise_rr_0"], g["noise_p_0"])
            got = eng.get_state()
        assert e == pytest.approx(float(g[f"step{hi}_0_E"]), rel=TOL)
        for k in ("r", "a", "b", "c", "r_t", "r_lh"):
            assert rel_err(getattr(got, k), g[f"step{hi}_0_{k}"]) <= TOL, (hi, k)
    fib.hi_all = 0


@pytest.mark.parametrize("kind,hi", [("tiny", 0), ("tiny", 1), ("small", 0), ("small", 1),
                                     ("cfg1", 0)])
def test_tensor_factor_noise(kind, hi):
    fib, st = deck(kind, hi_all=hi)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    r_all = np.concatenate([fs.view("r"), fs.view("r_t"), fs.view("r_lh")])[:3 * fib.n_hi]
    D0, L0, neg0 = orc.diffusion_tensor(r_all, fib.rad_all()[:fib.n_hi])
    rr, p = noise(fib)
    with E.BDEngine(fib, use_graph=False) as eng:
        sync_engine(eng, fs)
        D = eng.get_tensor()
        eng.factor()
        L, neg = eng.get_factor()
        rd = eng.noise_displacement(p)
    assert rel_err(D, D0) <= 1e-14
    assert rel_err(L, L0) <= TOL
    assert neg == neg0
    assert rel_err(L @ L.T, D0) <= 1e-13
    out = orc.step(fs.copy(), rr, p, want=True)
    assert rel_err(rd, out["rd"]) <= TOL


@pytest.mark.parametrize("kind,hi", [("tiny", 0), ("tiny", 1), ("small", 0), ("small", 1),
                                     ("cfg1", 0)])
def test_steps_resynchronised(kind, hi):
    """Three steps, each from the oracle's state with the same noise."""
    fib, st = deck(kind, hi_all=hi)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    g = O.Rng(5489)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    with E.BDEngine(fib, use_graph=True) as eng:
        for s in range(3):
            rr, p = g.step_noise(s, fib.r_var(), nD3)
            sync_engine(eng, fs)
            e = eng.step(rr, p)
            got, der = eng.get_state(), eng.get_derived()
            eo = orc.step(fs, rr, p)["E"]
            assert e == pytest.approx(eo, rel=TOL)
            for k in ("r", "a", "b", "c", "r_t", "r_lh"):
                assert rel_err(getattr(got, k), fs.view(k)) <= TOL, (s, k)
            for k in ("length", "a_dna", "b_dna", "c_dna", "beta_t", "beta_lh"):
                assert rel_err(der[k], fs.view(k)) <= TOL, (s, k)


def test_free_running_trajectory_tracks_oracle():
    """Ten steps without re-synchronisation: the chaotic divergence of two
    FP64 implementations stays at rounding level over this horizon."""
    fib, st = deck("tiny")
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    g = O.Rng(99)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    with E.BDEngine(fib) as eng:
        eng.set_state(st, rebuild=True)
        for s in range(10):
            rr, p = g.step_noise(s, fib.r_var(), nD3)
            eng.step(rr, p)
            orc.step(fs, rr, p)
        got = eng.get_state()
    for k in ("r", "r_t", "r_lh"):
        assert rel_err(getattr(got, k), fs.view(k)) <= 1e-9, k


def test_graph_replay_equals_eager_and_is_deterministic():
    fib, st = deck("small")
    rr, p = noise(fib)
    outs = []
    for graph in (False, True, True):
        with E.BDEngine(fib, use_graph=graph) as eng:
            eng.set_state(st, rebuild=True)
            es = [eng.step(rr, p) for _ in range(3)]
            s = eng.get_state()
        outs.append((es, s))
    for es, s in outs[1:]:
        assert es == outs[0][0]
        for k in ("r", "a", "b", "c", "r_t", "r_lh"):
            assert np.array_equal(getattr(s, k), getattr(outs[0][1], k)), k


@pytest.mark.parametrize("kind", ["small", "cfg1", "cfg2"])
def test_fused_launches_equal_timed_separate_launches(kind):
    """The per-kernel timing mode launches every kernel on its own and in
    order on one stream (the charge sites and chunk spheres apart from the
    frame passes, force pass 2 without branches); the step graph fuses and
    branches them.  Both give the same three steps bit for bit."""
    print(f"case: fused against separate launches, {kind}", flush=True)
    fib, st = Fb.make_config(2) if kind == "cfg2" else deck(kind)
    g = O.Rng(9)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    steps = [g.step_noise(s, fib.r_var(), nD3) for s in range(3)]
    outs = []
    for timed in (False, True):
        with E.BDEngine(fib, use_graph=not timed) as eng:
            eng.set_timing(timed)
            eng.set_state(st, rebuild=True)
            es = [eng.step(rr, p) for rr, p in steps]
            outs.append((es, eng.get_state(), eng.get_derived()))
    assert outs[0][0] == outs[1][0]
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        assert np.array_equal(getattr(outs[0][1], k), getattr(outs[1][1], k)), k
    for k, v in outs[0][2].items():
        assert np.array_equal(v, outs[1][2][k]), k


def test_device_noise_path_equals_host_noise_path():
    fib, st = deck("tiny")
    rr, p = noise(fib)
    with E.BDEngine(fib) as eng:
        eng.set_state(st, rebuild=True)
        e1 = eng.step(rr, p)
        s1 = eng.get_state()
    d_rr, d_p, d_e = E.DeviceBuffer(rr.nbytes), E.DeviceBuffer(p.nbytes), E.DeviceBuffer(8)
    d_rr.upload(rr)
    d_p.upload(p)
    with E.BDEngine(fib) as eng:
        eng.set_state(st, rebuild=True)
        eng.step_device(d_rr.ptr, d_p.ptr, d_e.ptr)
        eng.synchronize()
        e2 = float(d_e.download(1)[0])
        s2 = eng.get_state()
    assert e1 == e2 and np.array_equal(s1.r, s2.r) and np.array_equal(s1.r_lh, s2.r_lh)


def test_linker_histone_rebinding():
    """bd_set_lh: unbind one linker histone and continue; the step equals the
    oracle run on the fiber with that histone absent."""
    fib, st = deck("small")
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    flags = fib.nc_lh_flag.copy()
    flags[2] = 0
    nlh = fib.n_lh_n + fib.n_lh_c
    keep = np.concatenate([np.arange(k * nlh, (k + 1) * nlh) for k in range(fib.n_c) if flags[k]])
    fib2, st2 = Fb.make_fiber(6, seed=12, with_lh=True)
    fib2.nc_lh_flag = flags
    fib2.lh_rad = fib.lh_rad[keep]
    st2.r_lh = fs.view("r_lh").reshape(-1, 3)[keep].reshape(-1)
    orc2 = O.Oracle(fib2)
    fs2 = orc2.prepare(st2)
    rr, p = noise(fib2)
    with E.BDEngine(fib, lh_capacity=fib.n_LH) as eng:
        eng.set_state(st, rebuild=True)
        eng.set_lh(flags, st2.r_lh, fib2.lh_rad)
        e = eng.step(rr, p)
        got = eng.get_state()
    eo = orc2.step(fs2, rr, p)["E"]
    assert e == pytest.approx(eo, rel=TOL)
    for k in ("r", "r_t", "r_lh"):
        assert rel_err(getattr(got, k), fs2.view(k)) <= TOL, k


def test_cfg2_full_size_step():
    """One step of the 100-nucleosome benchmark fiber (D 2082 x 2082) from
    the same state and noise as the oracle."""
    fib, st = Fb.make_config(2)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    rr, p = noise(fib)
    with E.BDEngine(fib) as eng:
        sync_engine(eng, fs)
        e = eng.step(rr, p)
        got = eng.get_state()
        w = eng.work_counts()
    out = orc.step(fs, rr, p)
    assert e == pytest.approx(out["E"], rel=TOL)
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        assert rel_err(getattr(got, k), fs.view(k)) <= TOL, k
    assert w["core_pairs"] > 0


def test_mixed_precision_factor_displacement_error():
    """FP32 trailing updates (FP64 panels): displacement within 1e-6 of FP64."""
    fib, st = Fb.make_config(2)
    rr, p = noise(fib)
    res = {}
    for mode in (E.CHOL_FP64, E.CHOL_FP32):
        with E.BDEngine(fib, chol_mode=mode) as eng:
            eng.set_state(st, rebuild=True)
            x0 = eng.get_state()
            eng.step(rr, p)
            x1 = eng.get_state()
            eng.factor()
            L, _ = eng.get_factor()
            D = eng.get_tensor()
        res[mode] = (np.concatenate([x1.r - x0.r, x1.r_t - x0.r_t, x1.r_lh - x0.r_lh]), L, D)
    d64, L64, D = res[E.CHOL_FP64]
    d32, L32, _ = res[E.CHOL_FP32]
    assert np.linalg.norm(d32 - d64) / np.linalg.norm(d64) < 1e-6
    err32 = np.linalg.norm(L32 @ L32.T - D) / np.linalg.norm(D)
    err64 = np.linalg.norm(L64 @ L64.T - D) / np.linalg.norm(D)
    assert err64 < 1e-14 and err32 < 1e-5


def test_cfg3_forces_factor_and_long_chain():
    """The 500-nucleosome fiber (D 10482 x 10482; a chain of 3494 beads,
    longer than the one-workgroup frame launch takes): the frames and the
    full force pass against the oracle, the factor through L (L^T v) = D v
    on random vectors (an n^3 product on the host is out of reach), and two
    steps replayed from a graph equal to two eager steps bit for bit."""
    fib, st = Fb.make_config(3)
    print("case: cfg3, 500 nucleosomes, D 10482 x 10482", flush=True)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    with E.BDEngine(fib, use_graph=False) as eng:
        eng.set_state(st, rebuild=True)
        g = eng.get_state()
        d = eng.get_derived()
        sync_engine(eng, fs)
        got = eng.forces(E.TERM_ALL)
        D = eng.get_tensor()
        eng.factor()
        L, neg = eng.get_factor()
    for k in ("a", "b", "c"):
        assert rel_err(getattr(g, k), fs.view(k)) <= 1e-14, k
    for k in ("length", "a_dna", "b_dna", "c_dna", "beta_t"):
        assert rel_err(d[k], fs.view(k)) <= 1e-13, k
    ref = orc.forces(fs, ("all",))
    for k in ("F", "T", "tF"):
        assert rel_err(got[k], ref[k]) <= TOL, k
    assert got["E"] == pytest.approx(ref["E"], rel=TOL)
    assert neg == 0 and D.shape == (10482, 10482)
    rng = np.random.default_rng(3)
    for _ in range(3):
        v = rng.standard_normal(D.shape[0])
        Dv = D @ v
        assert np.linalg.norm(L @ (L.T @ v) - Dv) / np.linalg.norm(Dv) <= 1e-13
    del D, L
    steps = [noise(fib, s) for s in range(2)]
    out = {}
    for graph in (True, False):
        with E.BDEngine(fib, use_graph=graph) as eng:
            eng.set_state(st, rebuild=True)
            en = [eng.step(rr, p) for rr, p in steps]
            out[graph] = (en, eng.get_state())
    assert out[True][0] == out[False][0] and all(np.isfinite(out[True][0]))
    for k in ("r", "a", "b", "c", "r_t"):
        np.testing.assert_array_equal(getattr(out[True][1], k), getattr(out[False][1], k), err_msg=k)


def test_negative_pivot_factor():
    """The |f| rule (func.cpp:825-835) on the HIP factor: a 43-bead fiber
    (D 129 x 129, three 64-wide tiles) whose beads 5 and 30 carry a negative
    radius, so six pivots are negative -- three in the first tile, whose
    flipped columns then feed the trailing updates of the other two, and
    three in the second.  L and the count against the compiled reference
    (tests/golden/negative_pivot.npz: its L and its wrongD.txt count)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "negative_pivot.npz"))
    fib, st = Fb.make_fiber(7, seed=5)
    assert np.array_equal(st.r, z["r"])
    fib.rad = z["rad"].copy()
    for mode in (E.CHOL_FP64, E.CHOL_FP32):
        with E.BDEngine(fib, chol_mode=mode, use_graph=False) as eng:
            eng.set_state(st, rebuild=True)
            D = eng.get_tensor()
            eng.factor()
            L, neg = eng.get_factor()
        assert rel_err(D, z["D"]) <= 1e-14
        assert neg == int(z["neg"]) == 6
        assert rel_err(L, z["L"]) <= (TOL if mode == E.CHOL_FP64 else 1e-5)


@pytest.mark.parametrize("panel", [-1, 2, 3, 8])
def test_super_panel_schedule_factor(panel):
    """The super-panel schedule (panel columns factored inside the panel,
    one K = 64k trailing pass per panel: k_step restricted to the panel,
    k_panel, k_trail, k_diag) against the oracle's factor, on the 6-core
    deck with the full hydrodynamic set (D 1404 x 1404, 22 block columns:
    several whole panels and a ragged last one), FP64; then the FP32 and
    three-slice BF16 trailing updates against the FP64 factor."""
    fib, st = deck("small", hi_all=1)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    r_all = np.concatenate([fs.view("r"), fs.view("r_t"), fs.view("r_lh")])[:3 * fib.n_hi]
    D0, L0, neg0 = orc.diffusion_tensor(r_all, fib.rad_all()[:fib.n_hi])
    Ls = {}
    for mode in (E.CHOL_FP64, E.CHOL_FP32, E.CHOL_BF16):
        with E.BDEngine(fib, chol_mode=mode, use_graph=False) as eng:
            eng.set_chol_schedule(panel)
            assert eng.chol_schedule() == (0 if panel < 0 else panel)
            sync_engine(eng, fs)
            eng.factor()
            Ls[mode], neg = eng.get_factor()
            assert neg == neg0 == 0
    assert rel_err(Ls[E.CHOL_FP64], L0) <= TOL
    for mode in (E.CHOL_FP32, E.CHOL_BF16):
        e = np.linalg.norm(Ls[mode] @ Ls[mode].T - D0) / np.linalg.norm(D0)
        print(f"mode {mode} panel {panel}: |LL^T - D| / |D| = {e:.2e}, "
              f"|L - L64| / |L64| = {rel_err(Ls[mode], L0):.2e}")
        assert e <= 1e-6 and rel_err(Ls[mode], L0) <= 1e-5


@pytest.mark.parametrize("mode", [E.CHOL_FP64, E.CHOL_FP32, E.CHOL_BF16])
@pytest.mark.parametrize("panel", [2, 3, 8])
def test_packed_panels_equal_direct(mode, panel, monkeypatch):
    """The super-panel trailing update from panel tiles packed once in
    matrix-core operand order (k_pack64 + k_trail_pk in FP64 and FP32,
    k_split3 + k_trail_bf in three-slice BF16; the default) against the form
    that reads the parked tiles directly (k_trail<MODE>, BD_PANEL_PACK=0,
    which splits every BF16 operand fragment in every workgroup): the same
    products in the same order, so the same factor bit for bit (6-core deck,
    full hydrodynamic set, D 1404 x 1404, 22 block columns; panels of 3 give
    odd trailing edges)."""
    fib, st = deck("small", hi_all=1)
    # the packed form for every super-panel (by default FP64 takes it from
    # 192 trailing block rows, FP32 from 96)
    monkeypatch.setenv("BD_PANEL_PACK_MIN_ROWS", "0")
    Ls = {}
    for pack in ("1", "0"):
        monkeypatch.setenv("BD_PANEL_PACK", pack)
        with E.BDEngine(fib, chol_mode=mode, use_graph=False) as eng:
            eng.set_chol_schedule(panel)
            eng.set_state(st, rebuild=True)
            eng.factor()
            Ls[pack], neg = eng.get_factor()
            assert neg == 0
    assert np.array_equal(Ls["1"], Ls["0"])


def _mixed_step_errors(fib, st, rr, p, modes):
    """Displacement of one step and ||L L^T - D|| / ||D|| for each factor
    mode, from the same state and noise."""
    res = {}
    for mode in modes:
        with E.BDEngine(fib, chol_mode=mode) as eng:
            eng.set_state(st, rebuild=True)
            x0 = eng.get_state()
            eng.step(rr, p)
            x1 = eng.get_state()
        res[mode] = np.concatenate([x1.r - x0.r, x1.r_t - x0.r_t, x1.r_lh - x0.r_lh])
    return res


def test_bf16_factor_and_displacement_cfg2():
    """Three-slice BF16 trailing updates at the 100-nucleosome fiber: the
    factor within 1e-6 of D and the step's displacement within 1e-6 of the
    FP64 step (same state and noise)."""
    fib, st = Fb.make_config(2)
    rr, p = noise(fib)
    d = _mixed_step_errors(fib, st, rr, p, (E.CHOL_FP64, E.CHOL_BF16))
    rel = np.linalg.norm(d[E.CHOL_BF16] - d[E.CHOL_FP64]) / np.linalg.norm(d[E.CHOL_FP64])
    with E.BDEngine(fib, chol_mode=E.CHOL_BF16) as eng:
        eng.set_state(st, rebuild=True)
        D = eng.get_tensor()
        eng.factor()
        L, neg = eng.get_factor()
    err = np.linalg.norm(L @ L.T - D) / np.linalg.norm(D)
    print(f"cfg2 bf16: displacement {rel:.2e}, |LL^T - D| / |D| = {err:.2e}")
    assert neg == 0 and err < 1e-6 and rel < 1e-6


@pytest.mark.parametrize("hi", [0, 1])
def test_frequency_rp_against_oracle_reuse(hi):
    """frequency_RP = 3 (bd_set_frequency_rp): steps 0 and 3 rebuild D and
    L, the others reuse the factor of the last rebuild for rd, against the
    oracle's reuse mode (or_step_reuse) over six free-running steps."""
    fib, st = deck("small", hi_all=hi)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    g = O.Rng(41)
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    Lk = None
    with E.BDEngine(fib) as eng:
        eng.set_state(st, rebuild=True)
        eng.set_frequency_rp(3)
        for s in range(6):
            rr, p = g.step_noise(s, fib.r_var(), nD3)
            e = eng.step(rr, p)
            if s % 3 == 0:
                out = orc.step(fs, rr, p, want=True)
                Lk = out["L"]
            else:
                out = orc.step(fs, rr, p, want=True, L_reuse=Lk)
            assert e == pytest.approx(out["E"], rel=1e-9)
            got = eng.get_state()
            for k in ("r", "a", "b", "c", "r_t", "r_lh"):
                assert rel_err(getattr(got, k), fs.view(k)) <= 1e-9, (s, k)


@pytest.mark.parametrize("kind", ["small", "cfg2"])
def test_cell_list_matches_chunk_search(kind):
    """The 7 nm cell grids (bd_set_cell_list(1), forced on here; the default
    from 20000 tail and LH beads up) and the chunk spheres (0) evaluate the
    same pair set: forces within summation-order
    rounding of each other and of the oracle; one graph step each."""
    fib, st = deck(kind) if kind != "cfg2" else Fb.make_config(2)
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    rr, p = noise(fib)
    got = {}
    for cells in (True, False):
        with E.BDEngine(fib) as eng:
            eng.set_cell_list(cells)
            sync_engine(eng, fs)
            f = eng.forces(E.TERM_ALL)
            e = eng.step(rr, p)
            got[cells] = (f, e, eng.get_state())
    ref = orc.forces(fs, ("all",))
    for k in ("F", "T", "tF", "lhF"):
        assert rel_err(got[True][0][k], got[False][0][k]) <= 1e-13, k
        assert rel_err(got[True][0][k], ref[k]) <= TOL, k
    assert got[True][1] == pytest.approx(got[False][1], rel=1e-13)
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        assert rel_err(getattr(got[True][2], k), getattr(got[False][2], k)) <= 1e-12, k
