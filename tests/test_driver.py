"""Host driver (brownian_dynamics_cuda_amd/driver.py) on the CPU.

- The writers and readers against the reference's own readfile.cpp
  (compiled into oracle/_ref/libbdref.so): byte-identical out.xyz frames and
  restart files, identical values read back, and the initial Er.
- The placement of a newly bound linker histone against the reference's
  build_LH (func.cpp:2029-2059).
- The run loop with a recording stand-in for the GPU engine: the noise
  stream and Monte Carlo uniforms in main.cpp's order, replayed with the
  reference's own MT19937 and rand_normal; the stale-frame rule of a new
  histone and its fix; batching; checkpoint/resume of the generator state.
"""
import ctypes as C
import dataclasses
import os
import shutil

import numpy as np
import pytest

import oracle_py as O
from brownian_dynamics_cuda_amd import driver as D
from brownian_dynamics_cuda_amd import system as S
from decks import GOLDEN

DECK = os.path.join(GOLDEN, "testdata")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def pd(a):
    return a.ctypes.data_as(_dp)


def pi(a):
    return a.ctypes.data_as(_ip)


def make_deck(tmp, **setup):
    """Copy of the reference deck with setup.txt keys replaced."""
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


def moved_state(fib, st, seed=3):
    """The deck state with every coordinate perturbed (values of all signs
    and magnitudes for the number formats)."""
    g = np.random.default_rng(seed)
    out = st.copy()
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        v = getattr(out, k)
        v += g.normal(0, 0.3, v.shape)
    out.r[4] = -0.0
    out.r[7] = 1.5e-7
    out.r_t[0] = 123456.789
    return out


# ---------------------------------------------------------------------------
# formats against readfile.cpp
# ---------------------------------------------------------------------------

@needs_ref
def test_xyz_frame_matches_reference_writer(tmp_path):
    fib, st, _, _ = D.load_deck(DECK)
    st = moved_state(fib, st)
    flags = np.array(fib.nc_lh_flag, np.int32)
    flags[3] = 0  # one histone unbound: drop its beads
    nlh = fib.n_lh_n + fib.n_lh_c
    keep = np.repeat(np.array([f == 1 for f in fib.nc_lh_flag]), nlh)
    bound_rows = np.repeat(np.arange(fib.n_c)[np.array(fib.nc_lh_flag) == 1], nlh)
    r_lh = st.r_lh.reshape(-1, 3)[bound_rows != 3].reshape(-1)
    st.r_lh = r_lh
    assert keep.sum() * 3 == len(fib.lh_rad) * 3
    lib = O.ref_lib()
    path = str(tmp_path / "ref.xyz")
    types = np.ascontiguousarray(fib.types, np.int32)
    r, a, b, c = (np.ascontiguousarray(x) for x in (st.r, st.a, st.b, st.c))
    core = np.ascontiguousarray(fib.core_pos.reshape(-1))
    lhg = np.ascontiguousarray(fib.LH_g_pos.reshape(-1))
    for _ in range(2):  # appends, like the reference's sampling steps
        lib.ref_write_xyz_append(fib.n, fib.n_c, fib.n_tail, fib.n_lh_g, int(np.sum(flags == 1)),
                                 pi(flags), len(r_lh) // 3, fib.Nq, pi(types), pd(r), pd(a),
                                 pd(b), pd(c), pd(core), pd(np.ascontiguousarray(st.r_t)), pd(lhg),
                                 pd(np.ascontiguousarray(r_lh)), path.encode())
    mine = str(tmp_path / "mine.xyz")
    D.append_xyz(mine, fib, st, flags)
    D.append_xyz(mine, fib, st, flags)
    ref_txt, my_txt = open(path).read(), open(mine).read()
    assert my_txt == ref_txt
    assert "LHG" in my_txt and "NC" in my_txt and "-0 " in my_txt


@needs_ref
def test_restart_files_match_reference_writer_and_reader(tmp_path):
    fib, st, er, _ = D.load_deck(DECK)
    st = moved_state(fib, st)
    lib = O.ref_lib()
    refdir, mydir = tmp_path / "ref", tmp_path / "mine"
    refdir.mkdir()
    mydir.mkdir()
    n3, nt3, nl3 = len(st.r), len(st.r_t), len(st.r_lh)
    arrs = [np.ascontiguousarray(x, np.float64) for x in (st.r, st.a, st.b, st.c, er, st.r_t, st.r_lh)]
    old = os.getcwd()
    try:
        os.chdir(refdir)
        lib.ref_write_restart(n3, nt3, nl3, *[pd(x) for x in arrs])
    finally:
        os.chdir(old)
    D.write_restart(str(mydir), st, er)
    names = ["restart_r.txt", "restart_a.txt", "restart_b.txt", "restart_c.txt", "restart_Er.txt",
             "restart_tail.txt", "restart_lh.txt"]
    for nm in names:
        assert open(mydir / nm).read() == open(refdir / nm).read(), nm
    # the reference's readers and ours agree on what the files hold
    outs = [np.zeros(n3) for _ in range(5)] + [np.zeros(nt3), np.zeros(nl3)]
    try:
        os.chdir(mydir)
        lib.ref_read_restart(n3, nt3, nl3, *[pd(x) for x in outs])
    finally:
        os.chdir(old)
    mine = D.read_restart(str(mydir), n3, nt3, nl3)
    for x, y in zip(mine, outs):
        np.testing.assert_array_equal(x, y)
    # 15 decimals: within half a unit of the 15th decimal of what was written
    for x, y in zip(mine, arrs):
        assert np.max(np.abs(x - y)) <= 5.1e-16 * max(1.0, np.max(np.abs(y))) + 5e-16


@needs_ref
def test_initial_er_matches_reference_reader():
    lib = O.ref_lib()
    out = np.zeros(3 * 200)
    k = lib.ref_initial_er(os.path.join(DECK, "data_mod").encode(), pd(out), len(out))
    types, *_ = S.read_data_mod(os.path.join(DECK, "data_mod"))
    np.testing.assert_array_equal(D.initial_er(types), out[:k])


@needs_ref
def test_lh_placement_matches_build_lh():
    fib, st, _, _ = D.load_deck(DECK)
    st = moved_state(fib, st, seed=9)
    lib = O.ref_lib()
    nlh = fib.n_lh_n + fib.n_lh_c
    flags = np.array([1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 0, 1], np.int32)[:fib.n_c].copy()
    nb = int(np.sum(flags == 1))
    out = np.zeros(3 * nb * nlh)
    rad = np.zeros(nb * nlh)
    types = np.ascontiguousarray(fib.types, np.int32)
    npos = np.ascontiguousarray(fib.LH_n_pos.reshape(-1))
    cpos = np.ascontiguousarray(fib.LH_c_pos.reshape(-1))
    lrad = np.ascontiguousarray(fib.LH_radius[:nlh], np.float64)
    frames = [np.ascontiguousarray(x) for x in (st.r, st.a, st.b, st.c)]
    lib.ref_build_lh(fib.n, pi(types), nb * nlh, pd(npos), fib.n_lh_n, pd(cpos), fib.n_lh_c,
                     pd(lrad), *[pd(x) for x in frames], pi(flags), pd(out), pd(rad))
    # all newly bound: every bead comes from build_LH
    none = np.zeros_like(flags)
    got, got_rad = D.rebind_positions(fib, none, flags, np.zeros(0), frames)
    np.testing.assert_array_equal(got, out)
    np.testing.assert_array_equal(got_rad, rad)


def test_rebind_keeps_beads_of_histones_that_stay():
    fib, st, _, _ = D.load_deck(DECK)
    nlh = fib.n_lh_n + fib.n_lh_c
    old = np.array(fib.nc_lh_flag, np.int32)
    new = old.copy()
    new[2] = 0
    r_lh_old = np.arange(3 * nlh * int(old.sum()), dtype=np.float64)
    frames = (st.r, st.a, st.b, st.c)
    got, _ = D.rebind_positions(fib, old, new, r_lh_old, frames)
    blocks = r_lh_old.reshape(-1, nlh * 3)
    want = np.concatenate([blocks[k] for k in range(len(old)) if k != 2])
    np.testing.assert_array_equal(got, want)
    # the reverse move rebuilds core 2's histone on its frame
    back, _ = D.rebind_positions(fib, new, old, got, frames)
    fresh = S.build_lh_positions(fib.types, old, fib.n_lh_n, fib.LH_n_pos, fib.LH_c_pos, *frames)
    np.testing.assert_array_equal(back.reshape(-1, 3 * nlh)[2], fresh.reshape(-1, 3 * nlh)[2])
    np.testing.assert_array_equal(np.delete(back.reshape(-1, 3 * nlh), 2, 0),
                                  np.delete(r_lh_old.reshape(-1, 3 * nlh), 2, 0))


def test_lh_move_rule():
    f = np.array([1, 0, 1, 0], np.int32)
    # core int(4 u): u = 0.3 -> core 1 (unbound): binds when u_accept <= Pbind
    assert D.lh_move(f, 0.5, 0.3, 0.0, 0.5).tolist() == [1, 1, 1, 0]
    assert D.lh_move(f, 0.51, 0.3, 1.0, 0.5).tolist() == [1, 0, 1, 0]
    # core 2 (bound): unbinds when u_accept <= unbind
    assert D.lh_move(f, 0.2, 0.5, 0.2, 1.0).tolist() == [1, 0, 0, 0]
    assert D.lh_move(f, 0.21, 0.5, 0.2, 1.0).tolist() == [1, 0, 1, 0]
    # u_core = 1 (genrand_real1 reaches 1) clamps to the last core
    assert D.lh_move(f, 0.0, 1.0, 0.0, 1.0).tolist() == [1, 0, 1, 1]
    assert f.tolist() == [1, 0, 1, 0]


def test_rng_state_file_round_trip(tmp_path):
    words = np.arange(624, dtype=np.uint32) * 7919
    p = str(tmp_path / "rng.txt")
    D.write_rng_state(p, 42, words, 17, 1, -0.123456789012345678)
    seed, w, idx, cached, n2 = D.read_rng_state(p)
    assert (seed, idx, cached, n2) == (42, 17, 1, -0.123456789012345678)
    np.testing.assert_array_equal(w, words)


# ---------------------------------------------------------------------------
# the run loop with a recording stand-in for the engine
# ---------------------------------------------------------------------------

class RecordingEngine:
    """CPU stand-in for BDEngine: every step moves the main beads and the
    histone beads by a fixed shift, returns the step count as the energy and
    records the noise it was given."""

    SHIFT = 1e-3

    def __init__(self, fib):
        self.fib = fib
        self.n, self.n_tail = fib.n, fib.n_tail
        self.n_LH = fib.n_LH
        self.n_D3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
        self.rr, self.p, self.lh_calls, self.rp = [], [], [], None
        self.count = 0

    def set_state(self, st, rebuild=True):
        assert len(st.r_lh) == 3 * self.n_LH
        self.st = st.copy()

    def get_state(self):
        return self.st.copy()

    def set_lh(self, flags, r_lh, lh_rad):
        assert len(r_lh) == 3 * len(lh_rad)
        self.lh_calls.append((self.count, np.array(flags), np.array(r_lh)))
        self.n_LH = len(lh_rad)
        self.n_D3 = 3 * (self.n + self.n_tail + self.n_LH)
        self.st.r_lh = np.array(r_lh, np.float64)

    def set_frequency_rp(self, f):
        self.rp = f

    def get_derived(self):
        return {"alpha": np.full(self.n, float(self.count))}

    def set_derived(self, d):
        self.derived_in = {k: np.array(v) for k, v in d.items()}

    def run_steps(self, rr, p):
        assert rr.shape[1] == 3 * self.n and p.shape[1] == self.n_D3
        k = rr.shape[0]
        self.rr += list(rr)
        self.p += list(p)
        self.st.r = self.st.r + self.SHIFT * k
        self.st.r_lh = self.st.r_lh + self.SHIFT * k
        e = np.arange(self.count, self.count + k, dtype=np.float64)
        self.count += k
        return e


def fresh_ref_rng(tmp_path, seed):
    """The reference's MT19937 and rand_normal from a private copy of the
    library, so that rand_normal's static Box-Muller cache starts empty."""
    src = O.REF_SO
    dst = str(tmp_path / "libbdref_rng.so")
    shutil.copy(src, dst)
    lib = C.CDLL(dst)
    lib.ref_init_genrand.argtypes = [C.c_ulong]
    lib.ref_genrand_real1.restype = C.c_double
    lib.ref_rand_normal.argtypes = [C.c_double, C.c_double]
    lib.ref_rand_normal.restype = C.c_double
    lib.ref_init_genrand(seed)
    return lib


def run_with(deck, out, **kw):
    """Run the driver over `deck` with a RecordingEngine sized like the
    driver's fiber (the checkpoint's binding when resuming)."""
    fib, _, _, _ = D.load_deck(deck)
    if "resume" in kw:
        fl = np.array(np.load(kw["resume"])["nc_lh_flag"], np.int32)
        nlh = fib.n_lh_n + fib.n_lh_c
        fib = dataclasses.replace(fib, nc_lh_flag=fl,
                                  lh_rad=np.tile(fib.LH_radius[:nlh], int(fl.sum())))
    eng = RecordingEngine(fib)
    run = D.Run(deck, out=out, engine=eng, **kw)
    res = run.run()
    return run, eng, res


@needs_ref
def test_noise_stream_and_moves_follow_main_cpp_order(tmp_path):
    deck = make_deck(tmp_path, number_of_steps=230, frequency_of_sampling=40, unbind=0.5,
                     Pbind=0.5)
    run, eng, res = run_with(deck, str(tmp_path / "out"), seed=1234, batch=64)
    fib = run.fib
    nlh = fib.n_lh_n + fib.n_lh_c
    ref = fresh_ref_rng(tmp_path, 1234)
    rv = fib.r_var()
    flags = np.array(fib.nc_lh_flag, np.int32)
    for step in range(230):
        if step % 100 == 0:
            u_acc = ref.ref_genrand_real1()
            u_core = ref.ref_genrand_real1()
            flags = D.lh_move(flags, u_acc, u_core, 0.5, 0.5)
        nD3 = 3 * (fib.n + fib.n_tail + int(flags.sum()) * nlh)
        rr = np.array([ref.ref_rand_normal(0.0, v) for v in rv])
        p = np.array([ref.ref_rand_normal(0.0, 1.0) for _ in range(nD3)])
        np.testing.assert_array_equal(eng.rr[step], rr, err_msg=f"rr step {step}")
        np.testing.assert_array_equal(eng.p[step], p, err_msg=f"p step {step}")
    assert res["nc_lh_flag"] == flags.tolist()
    assert len(eng.rr) == 230
    # outputs: one energy per step; frames at steps 0, 40, ..., 200; restart at the end
    E = np.loadtxt(tmp_path / "out" / "Energy.txt")
    np.testing.assert_array_equal(E, np.arange(230.0))
    xyz = open(tmp_path / "out" / "out.xyz").read().splitlines()
    assert sum(1 for l in xyz if l == "comments zl3765") == 6
    assert [int(x) for x in open(tmp_path / "out" / "LH.in")] == flags.tolist()
    assert len(np.loadtxt(tmp_path / "out" / "restart_lh.txt")) == 3 * nlh * int(flags.sum())


@pytest.mark.parametrize("fix", [False, True])
def test_new_histone_built_on_host_frame_copy(tmp_path, fix):
    deck = make_deck(tmp_path, number_of_steps=201, frequency_of_sampling=30, unbind=0.0,
                     Pbind=1.0)
    open(os.path.join(deck, "LH.in"), "w").write("0\n" * 12)
    run, eng, res = run_with(deck, str(tmp_path / "out"), seed=77, fix_stale_frames=fix)
    fib = run.fib
    nlh = fib.n_lh_n + fib.n_lh_c
    st0 = D.load_deck(deck)[1]
    assert [c[0] for c in eng.lh_calls][0] == 0
    checked = 0
    for count, flags, r_lh in eng.lh_calls:
        if count == 0:
            continue
        # the move at step `count`: frames after the last sampling step
        # before it (step 90 -> 91 steps run), or the current ones
        last_sample = max(s for s in range(count) if s % 30 == 0)
        k_run = count if fix else last_sample + 1
        frames = (st0.r + RecordingEngine.SHIFT * k_run, st0.a, st0.b, st0.c)
        fresh = S.build_lh_positions(fib.types, flags, fib.n_lh_n, fib.LH_n_pos, fib.LH_c_pos,
                                     *frames).reshape(-1, 3 * nlh)
        prev = [c for c in eng.lh_calls if c[0] < count][-1][1]
        new_core = int(np.nonzero(prev != flags)[0][0])
        slot = int(np.sum(flags[:new_core] == 1))
        np.testing.assert_allclose(r_lh.reshape(-1, 3 * nlh)[slot], fresh[slot], rtol=0, atol=1e-12)
        checked += 1
    assert checked >= 1


def test_checkpoint_resume_continues_the_noise_stream(tmp_path):
    deck = make_deck(tmp_path, number_of_steps=250, frequency_of_sampling=30, unbind=0.5,
                     Pbind=0.5)
    _, eng_a, res_a = run_with(deck, str(tmp_path / "a"), seed=99)
    _, eng_b, _ = run_with(deck, str(tmp_path / "b"), seed=99, steps=120)
    ck = str(tmp_path / "b" / D.CHECKPOINT_FILE)
    _, eng_c, res_c = run_with(deck, str(tmp_path / "b"), resume=ck)
    assert len(eng_b.rr) == 120 and len(eng_c.rr) == 130
    # the resumed engine gets the saved angles instead of a rebuild
    np.testing.assert_array_equal(eng_c.derived_in["alpha"], np.full(eng_c.n, 120.0))
    for x, y in zip(eng_a.rr, eng_b.rr + eng_c.rr):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(eng_a.p, eng_b.p + eng_c.p):
        np.testing.assert_array_equal(x, y)
    assert res_a["nc_lh_flag"] == res_c["nc_lh_flag"]
    # the text form of the generator state restores the same stream
    from brownian_dynamics_cuda_amd import engine as E
    seed, words, idx, cached, n2 = D.read_rng_state(str(tmp_path / "b" / D.RNG_FILE))
    ck = np.load(ck)
    np.testing.assert_array_equal(words, ck["rng_words"])
    assert (idx, cached, n2, seed) == (int(ck["rng_idx"]), int(ck["rng_cached"]),
                                       float(ck["rng_n2"]), 99)
    g1, g2 = E.HostRng(1), E.HostRng(2)
    g1.set_state(words, idx, cached, n2)
    g2.set_state(ck["rng_words"], int(ck["rng_idx"]), int(ck["rng_cached"]), float(ck["rng_n2"]))
    assert [g1.normal() for _ in range(5)] == [g2.normal() for _ in range(5)]


def test_batches_end_at_moves_and_samples(tmp_path):
    deck = make_deck(tmp_path, number_of_steps=260, frequency_of_sampling=70, unbind=0, Pbind=0)
    fib, *_ = D.load_deck(deck)
    run = D.Run(deck, out=str(tmp_path / "o"), seed=1, batch=64, engine=RecordingEngine(fib))
    plan = list(run._plan())
    assert plan[0][0] == 0 and plan[-1][1] == 260
    assert all(a < b for a, b in plan) and all(plan[i][1] == plan[i + 1][0] for i in range(len(plan) - 1))
    for s0, s1 in plan:
        assert s1 - s0 <= 64
        # no move inside a batch, and a sampling step only at its end
        assert all(s % 100 != 0 for s in range(s0 + 1, s1))
        assert all(s % 70 != 0 for s in range(s0, s1 - 1))
    starts = {s0 for s0, _ in plan}
    assert {0, 100, 200, 1, 71, 141, 211} <= starts


def test_frequency_rp_is_opt_in(tmp_path):
    deck = make_deck(tmp_path, number_of_steps=3)
    fib, *_ = D.load_deck(deck)
    e1 = RecordingEngine(fib)
    D.Run(deck, out=str(tmp_path / "o1"), seed=1, engine=e1)
    assert e1.rp is None
    e2 = RecordingEngine(fib)
    D.Run(deck, out=str(tmp_path / "o2"), seed=1, engine=e2, use_frequency_rp=True)
    assert e2.rp == 20  # setup.txt of the deck


def test_deck_writer_round_trip(tmp_path):
    """A generated fiber written in the reference formats reads back as the
    same fiber and state (17 significant digits, the readers' unit scalings
    undone)."""
    from brownian_dynamics_cuda_amd import fiber as Fb
    fib, st = Fb.make_fiber(4, seed=5, with_lh=True, hi_all=1)
    d = str(tmp_path / "gen")
    S.write_deck(d, fib, st, number_of_steps=300, frequency_of_sampling=25)
    fib2, st2, er, su = D.load_deck(d, hi_all=1, q_l=fib.phys.q_l,
                                    groups_per_core=fib.groups_per_core)
    assert (su.number_of_steps, su.frequency_of_sampling, su.restart) == (300, 25, 0)
    for k in ("r", "a", "b", "c", "r_t", "r_lh"):
        np.testing.assert_allclose(getattr(st2, k), getattr(st, k), rtol=1e-15, atol=1e-15,
                                   err_msg=k)
    for f in dataclasses.fields(fib):
        x, y = getattr(fib2, f.name), getattr(fib, f.name)
        if f.name == "phys":
            for g in dataclasses.fields(y):
                assert getattr(x, g.name) == pytest.approx(getattr(y, g.name), rel=1e-14), g.name
        elif isinstance(y, np.ndarray):
            assert x.shape == y.shape, f.name
            np.testing.assert_allclose(x, y, rtol=1e-14, atol=1e-15, err_msg=f.name)
        else:
            assert x == y, f.name
    np.testing.assert_array_equal(er, D.initial_er(fib.types))
