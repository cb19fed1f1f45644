"""Host driver of a BD run: the loop of main.cpp around the HIP step.

Reads an input deck in the reference's formats (setup.txt, data_mod, the
core, tail and linker-histone files, LH.in and, for restart = 1, the
restart_*.txt files), runs number_of_steps predictor-corrector steps with the
linker-histone binding Monte Carlo every 100 steps (main.cpp:384-465), and
writes Energy.txt, out.xyz, restart_*.txt and LH.in in the reference's formats
(main.cpp:490-512, readfile.cpp:346-441).

What differs from main.cpp, by design:
- Steps between two Monte Carlo moves or two samples run as one batch
  (bd_run_steps): one noise upload and one energy download per batch instead
  of per step, with the next batch's noise drawn on the host while the GPU
  runs the current one.  The draw order is main.cpp's.
- The seed is an argument (main.cpp:60-62 seeds MT19937 with time(0); the
  driver does the same when no seed is given) and the generator's state is
  checkpointed next to the restart files (restart_rng.txt), so a restarted run
  can continue the same noise stream.
- checkpoint.npz holds the full-precision state with its derived angles, the
  binding flags, the step and the generator state: resuming from it continues
  a run bit for bit.  restart_*.txt keep 15 decimals, as the reference writes
  them, and a restart from them rebuilds the angles, which re-orthonormalises
  the linker frames (update_Euler_Angle, func.cpp:448-505) as main.cpp:164
  does.
- fix_stale_frames: main.cpp places a newly bound linker histone with the host
  copy of the core frames, which is refreshed only on sampling steps
  (func_cuda.cu:3665-3674), so the histone is built on a frame up to
  frequency_of_sampling steps old.  The default keeps that behaviour; the flag
  builds it on the current frame instead.
- use_frequency_rp: honours frequency_RP (D and its factor rebuilt every
  frequency_RP steps), which the reference parses and ignores
  (func_cuda.cu:3504).
"""
from __future__ import annotations

import argparse
import dataclasses
import math
import os
import queue
import threading
import time
from typing import Optional

import numpy as np

from . import system as S
from .system import Fiber, State

LH_MOVE_PERIOD = 100  # main.cpp:386
RNG_FILE = "restart_rng.txt"
CHECKPOINT_FILE = "checkpoint.npz"


# ---------------------------------------------------------------------------
# number formats of the reference's writers
# ---------------------------------------------------------------------------

def fmt_g(x: float) -> str:
    """A double as `ostream << x` prints it by default: %g, 6 significant digits."""
    return "%g" % x


def fmt_f15(x: float) -> str:
    """std::fixed << std::setprecision(15) (readfile.cpp:418)."""
    return "%.15f" % x


def _rows(tag: str, xyz: np.ndarray) -> list:
    return [f"{tag} {fmt_g(p[0])} {fmt_g(p[1])} {fmt_g(p[2])}" for p in np.asarray(xyz).reshape(-1, 3)]


# ---------------------------------------------------------------------------
# writers and readers
# ---------------------------------------------------------------------------

def initial_er(types: np.ndarray) -> np.ndarray:
    """Drag coefficients Er per coordinate, as Read_initial_struc fills them
    (readfile.cpp:75-82): 8 pi eta 125 for the three coordinates of a core,
    (4 pi eta r_h^2 lo, 0, 0) for a linker bead."""
    er = np.zeros(3 * len(types))
    for i, t in enumerate(types):
        if t == 1:
            er[3 * i:3 * i + 3] = 8 * S.PI * S.ETA * 125.0
        else:
            er[3 * i] = 4 * S.PI * S.ETA * S.R_H * S.R_H * S.LO
    return er


def append_energy(path: str, energies) -> None:
    """Energy.txt: one pass-1 energy per step (main.cpp:492-495)."""
    with open(path, "a") as f:
        f.write("".join(fmt_g(float(e)) + "\n" for e in energies))


def xyz_frame(fib: Fiber, st: State, nc_lh_flag) -> str:
    """One out.xyz frame in the layout of write_xyz_append (readfile.cpp:346-405):
    a count line, a comment line, then per main bead its centre (CA linker,
    OC core) and frame vectors (H1-H3) with, after a core, its Nq charge
    sites (NC); the tail beads (HT); the globular-head sites of every bound
    linker histone (LHG); the free linker-histone beads (LH)."""
    types = fib.types
    flags = np.asarray(nc_lh_flag)
    n_LH = len(st.r_lh) // 3
    nc_lh = int(np.sum(flags == 1))
    r, a, b, c = (np.asarray(x).reshape(-1, 3) for x in (st.r, st.a, st.b, st.c))
    count = fib.n * 4 + fib.n_c * fib.Nq + fib.n_tail + n_LH + nc_lh * fib.n_lh_g
    out = [str(count), "comments zl3765"]
    for i in range(fib.n):
        out += _rows("CA" if types[i] == 0 else "OC", r[i])
        out += _rows("H1", a[i]) + _rows("H2", b[i]) + _rows("H3", c[i])
        if types[i] != 0:
            out += _rows("NC", S.body_to_world(r[i], a[i], b[i], c[i], fib.core_pos))
    out += _rows("HT", st.r_t)
    k = 0
    for i in range(fib.n):
        if types[i] == 1:
            if flags[k] == 1 and fib.n_lh_g:
                out += _rows("LHG", S.body_to_world(r[i], a[i], b[i], c[i], fib.LH_g_pos))
            k += 1
    out += _rows("LH", st.r_lh)
    return "\n".join(out) + "\n"


def append_xyz(path: str, fib: Fiber, st: State, nc_lh_flag) -> None:
    with open(path, "a") as f:
        f.write(xyz_frame(fib, st, nc_lh_flag))


def write_restart(directory: str, st: State, Er) -> None:
    """restart_r/a/b/c/Er/tail/lh.txt, one value per line with 15 decimals
    (write_restart, readfile.cpp:407-441)."""
    files = {"restart_r.txt": st.r, "restart_a.txt": st.a, "restart_b.txt": st.b,
             "restart_c.txt": st.c, "restart_Er.txt": Er, "restart_tail.txt": st.r_t,
             "restart_lh.txt": st.r_lh}
    for name, v in files.items():
        with open(os.path.join(directory, name), "w") as f:
            f.write("".join(fmt_f15(float(x)) + "\n" for x in np.asarray(v).reshape(-1)))


def write_lh_in(path: str, nc_lh_flag) -> None:
    """LH.in: one binding flag per core (main.cpp:504-508)."""
    with open(path, "w") as f:
        f.write("".join(f"{int(x)}\n" for x in nc_lh_flag))


def _read_column(path: str, count: int) -> np.ndarray:
    """First token of each of the first `count` lines (Read_restart_*)."""
    vals = []
    with open(path) as f:
        for line in f:
            if len(vals) == count:
                break
            tok = line.split()
            vals.append(float(tok[0]) if tok else 0.0)
    if len(vals) < count:
        raise ValueError(f"{path}: {len(vals)} values, {count} needed")
    return np.array(vals)


def read_restart(directory: str, n3: int, n_tail3: int, n_LH3: int):
    """Read_restart_ini, Read_restart_tail, Read_restart_LH (readfile.cpp:443-561):
    r, a, b, c, Er [n3], r_t [n_tail3], r_lh [n_LH3]."""
    j = lambda name: os.path.join(directory, name)
    r, a, b, c, er = (_read_column(j(f"restart_{k}.txt"), n3) for k in ("r", "a", "b", "c", "Er"))
    r_t = _read_column(j("restart_tail.txt"), n_tail3) if n_tail3 else np.zeros(0)
    r_lh = _read_column(j("restart_lh.txt"), n_LH3) if n_LH3 else np.zeros(0)
    return r, a, b, c, er, r_t, r_lh


def write_rng_state(path: str, seed: int, words, idx: int, cached: int, n2: float) -> None:
    """The generator's state after the last step: seed, position in the
    MT19937 state, the cached second Box-Muller normal and its flag, then the
    624 state words."""
    with open(path, "w") as f:
        f.write(f"{seed} {idx} {cached} {float(n2).hex()}\n")
        f.write("\n".join(str(int(w)) for w in words) + "\n")


def read_rng_state(path: str):
    with open(path) as f:
        head = f.readline().split()
        words = np.array([int(l) for l in f if l.strip()], np.uint32)
    if len(words) != 624:
        raise ValueError(f"{path}: {len(words)} state words, 624 needed")
    return int(head[0]), words, int(head[1]), int(head[2]), float.fromhex(head[3])


# ---------------------------------------------------------------------------
# linker-histone binding Monte Carlo (main.cpp:384-465)
# ---------------------------------------------------------------------------

def lh_move(nc_lh_flag, u_accept: float, u_core: float, unbind: float, Pbind: float) -> np.ndarray:
    """One move: core n_b = int(n_c u_core) (clamped to n_c-1) unbinds its
    histone with probability `unbind` or binds one with probability `Pbind`
    (main.cpp:386-404).  Returns the new flags."""
    flags = np.array(nc_lh_flag, np.int32, copy=True)
    n_c = len(flags)
    n_b = min(int(n_c * u_core), n_c - 1)
    if flags[n_b] > 0.5:
        if u_accept <= unbind:
            flags[n_b] = 0
    else:
        if u_accept <= Pbind:
            flags[n_b] = 1
    return flags


def rebind_positions(fib: Fiber, flags_old, flags_new, r_lh_old, frames):
    """Free linker-histone beads after a move (main.cpp:418-440): histones
    that stay bound keep their beads; a new one is built on its core's frame
    (build_LH, func.cpp:2029-2059) from `frames` = (r, a, b, c)."""
    nlh = fib.n_lh_n + fib.n_lh_c
    built = S.build_lh_positions(fib.types, np.asarray(flags_new), fib.n_lh_n, fib.LH_n_pos,
                                 fib.LH_c_pos, *frames).reshape(-1, nlh, 3)
    old = np.asarray(r_lh_old).reshape(-1, nlh, 3)
    k = k1 = 0
    for fo, fn in zip(flags_old, flags_new):
        if fn == 1:
            if fo == 1:
                built[k] = old[k1]
                k1 += 1
            k += 1
        elif fo == 1:
            k1 += 1
    n_bound = int(np.sum(np.asarray(flags_new) == 1))
    lh_rad = np.tile(np.asarray(fib.LH_radius[:nlh], np.float64), n_bound)
    return built.reshape(-1), lh_rad


# ---------------------------------------------------------------------------
# the run
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Setup:
    """setup.txt (Read_const, readfile.cpp:14-38; main.cpp:42-57)."""

    T: float
    number_of_steps: int
    frequency_of_sampling: int
    frequency_RP: int
    sigmas: tuple
    Mg: int
    Cs: float
    restart: int
    unbind: float
    Pbind: float

    @staticmethod
    def read(path: str) -> "Setup":
        v = S.read_setup(path)
        return Setup(T=v["T"], number_of_steps=int(v["number_of_steps"]),
                     frequency_of_sampling=int(v["frequency_of_sampling"]),
                     frequency_RP=int(v["frequency_RP"]),
                     sigmas=(v["sig_dd"], v["sig_dc"], v["sig_cc"], v["sig_tt"], v["sig_tl"],
                             v["sig_tc"]),
                     Mg=int(v["Mg"]), Cs=v["Cs"], restart=int(v["restart"]),
                     unbind=v["unbind"], Pbind=v["Pbind"])


def load_deck(deck: str, hi_all: int = 1, q_l: Optional[float] = None,
              groups_per_core: int = 10):
    """Fiber, start state, Er and setup of a deck directory, as main.cpp:40-283
    assembles them (restart = 1 takes r, a, b, c, Er, r_t and r_lh from the
    restart files of the same directory).  q_l (None: derived from Cs, as
    main.cpp:78-84) and groups_per_core (the literal 10 of func.cpp:2368)
    override the reference's values for decks written from other fibers."""
    su = Setup.read(os.path.join(deck, "setup.txt"))
    ph = S.Physics.derive(su.T, su.Cs, su.Mg, sigmas=su.sigmas, q_l=q_l)
    types, r, a, b, c, rad = S.read_data_mod(os.path.join(deck, "data_mod"))
    er = initial_er(types)
    core_pos, core_q = S.read_core(os.path.join(deck, "core_data.reg.150mM"))
    tl = S.read_tail(os.path.join(deck, "tail_data.mod.200mM"))
    lh = S.read_lh(os.path.join(deck, "LH_N0G6C22.in"))
    n_c = int(np.sum(types == 1))
    flags = S.read_lh_in(os.path.join(deck, "LH.in"), n_c)
    if su.restart == 1:
        n3 = 3 * len(types)
        r, a, b, c, er, _, _ = read_restart(deck, n3, 0, 0)
    fib, st = S.assemble(types, rad, r, a, b, c, core_pos, core_q, tl, lh, flags, ph,
                         hi_all=hi_all, groups_per_core=groups_per_core)
    if su.restart == 1:
        _, _, _, _, _, r_t, r_lh = read_restart(deck, 3 * fib.n, 3 * fib.n_tail, 3 * fib.n_LH)
        st = State(r=st.r, a=st.a, b=st.b, c=st.c, r_t=r_t, r_lh=r_lh)
    return fib, st, er, su


class Run:
    """One BD run over a deck.

    deck: input directory; out: output directory (created); seed: MT19937
    seed (None = time(0), as main.cpp:60-62); steps: overrides
    number_of_steps; hi_all: 1 = D over every bead (the reference's GPU
    path), 0 = cores and linker beads; resume_rng: with restart = 1, continue
    the noise stream of restart_rng.txt; resume: path of a checkpoint.npz to
    continue from; batch: most steps per bd_run_steps call; engine: an object
    with the BDEngine methods used here (tests pass a recording stand-in).
    """

    def __init__(self, deck: str, out: Optional[str] = None, seed: Optional[int] = None,
                 steps: Optional[int] = None, hi_all: int = 1, chol_mode: Optional[int] = None,
                 use_frequency_rp: bool = False, fix_stale_frames: bool = False,
                 resume_rng: bool = False, resume: Optional[str] = None, batch: int = 100,
                 write_checkpoint: bool = True, engine=None, rng=None, q_l: Optional[float] = None,
                 groups_per_core: int = 10):
        from . import engine as E
        self.deck = deck
        self.out = out or deck
        os.makedirs(self.out, exist_ok=True)
        self.fib, st, self.Er, self.setup = load_deck(deck, hi_all=hi_all, q_l=q_l,
                                                      groups_per_core=groups_per_core)
        self.nlh = self.fib.n_lh_n + self.fib.n_lh_c
        self.step0 = 0
        self.seed = int(time.time()) & 0xFFFFFFFF if seed is None else int(seed)
        self.rng = rng if rng is not None else E.HostRng(self.seed)
        derived = None
        if resume is not None:
            ck = np.load(resume, allow_pickle=False)
            derived = {k[4:]: np.array(ck[k]) for k in ck.files if k.startswith("der_")} or None
            st = State(*(np.array(ck[k], np.float64) for k in ("r", "a", "b", "c", "r_t", "r_lh")))
            flags = np.array(ck["nc_lh_flag"], np.int32)
            self.fib = dataclasses.replace(
                self.fib, nc_lh_flag=flags,
                lh_rad=np.tile(np.asarray(self.fib.LH_radius[:self.nlh], np.float64),
                               int(np.sum(flags == 1))))
            self.step0 = int(ck["step"])
            self.seed = int(ck["seed"])
            self.Er = np.array(ck["Er"], np.float64)
            self.rng.set_state(ck["rng_words"], int(ck["rng_idx"]), int(ck["rng_cached"]),
                               float(ck["rng_n2"]))
        elif resume_rng and self.setup.restart == 1:
            self.seed, words, idx, cached, n2 = read_rng_state(os.path.join(deck, RNG_FILE))
            self.rng.set_state(words, idx, cached, n2)
        fib = self.fib
        self.nsteps = int(steps) if steps is not None else self.setup.number_of_steps
        self.fos = max(1, self.setup.frequency_of_sampling)
        self.fix_stale = fix_stale_frames
        self.batch = max(1, int(batch))
        self.write_checkpoint = write_checkpoint
        self.flags = np.array(fib.nc_lh_flag, np.int32)
        if engine is None:
            mode = E.CHOL_FP64 if chol_mode is None else chol_mode
            engine = E.BDEngine(fib, chol_mode=mode, lh_capacity=fib.n_c * self.nlh)
        self.eng = engine
        if derived:
            self.eng.set_state(st, rebuild=False)
            self.eng.set_derived(derived)
        else:
            self.eng.set_state(st, rebuild=True)
        self.lh_rad = np.array(fib.lh_rad, np.float64)
        if use_frequency_rp and self.setup.frequency_RP > 1:
            self.eng.set_frequency_rp(self.setup.frequency_RP)
        # host copy of the main-bead frames as the reference keeps it: refreshed
        # on sampling steps only
        self.host_frames = (st.r.copy(), st.a.copy(), st.b.copy(), st.c.copy())
        self.r_var = fib.r_var()
        self.n3 = 3 * fib.n
        self.energies_written = 0
        self.lh_events = []  # (step, core, bound) of every accepted move

    # -- batches ---------------------------------------------------------------
    def _is_sample(self, s: int) -> bool:
        return s % self.fos == 0 or s == self.nsteps - 1

    def _plan(self):
        """Batches [s0, s1): each starts at a Monte Carlo step or right after
        a sampling step and holds at most `batch` steps."""
        s = self.step0
        while s < self.nsteps:
            e = s + 1
            while (e < self.nsteps and e - s < self.batch and e % LH_MOVE_PERIOD != 0
                   and not self._is_sample(e - 1)):
                e += 1
            yield s, e
            s = e

    def _producer(self, q: "queue.Queue"):
        """Draws, in main.cpp's order, each batch's Monte Carlo uniforms and
        noise; runs one batch ahead of the GPU."""
        flags = self.flags.copy()
        try:
            for s0, s1 in self._plan():
                move = None
                if s0 % LH_MOVE_PERIOD == 0:
                    u_accept = self.rng.uniform()
                    u_core = self.rng.uniform()
                    new = lh_move(flags, u_accept, u_core, self.setup.unbind, self.setup.Pbind)
                    move = (flags, new)
                    flags = new
                nD3 = 3 * (self.fib.n + self.fib.n_tail + int(np.sum(flags == 1)) * self.nlh)
                k = s1 - s0
                rr = np.empty((k, self.n3))
                p = np.empty((k, nD3))
                for j in range(k):
                    rr[j], p[j] = self.rng.step_noise(1, self.r_var, nD3)
                q.put((s0, s1, move, rr, p))
            q.put(None)
        except BaseException as exc:  # handed to the consumer
            q.put(exc)

    def _apply_move(self, step: int, move):
        old, new = move
        if np.array_equal(old, new):
            return
        changed = int(np.nonzero(old != new)[0][0])
        self.lh_events.append((step, changed, int(new[changed])))
        cur = self.eng.get_state()
        frames = (cur.r, cur.a, cur.b, cur.c) if self.fix_stale else self.host_frames
        r_lh, lh_rad = rebind_positions(self.fib, old, new, cur.r_lh, frames)
        self.eng.set_lh(new, r_lh, lh_rad)
        self.flags = new.copy()
        self.lh_rad = lh_rad

    def run(self) -> dict:
        """Runs steps step0 .. nsteps-1 and writes the output files; returns
        counters of the run."""
        q: "queue.Queue" = queue.Queue(maxsize=2)
        th = threading.Thread(target=self._producer, args=(q,), daemon=True)
        th.start()
        energy_path = os.path.join(self.out, "Energy.txt")
        xyz_path = os.path.join(self.out, "out.xyz")
        t0 = time.perf_counter()
        t_gpu = 0.0
        nb = 0
        while True:
            item = q.get()
            if item is None:
                break
            if isinstance(item, BaseException):
                raise item
            s0, s1, move, rr, p = item
            if move is not None:
                self._apply_move(s0, move)
            t1 = time.perf_counter()
            en = self.eng.run_steps(rr, p)
            t_gpu += time.perf_counter() - t1
            nb += 1
            append_energy(energy_path, en)
            self.energies_written += len(en)
            last = s1 - 1
            if self._is_sample(last):
                st = self.eng.get_state()
                self.host_frames = (st.r, st.a, st.b, st.c)
                if last % self.fos == 0:
                    append_xyz(xyz_path, self.fib, st, self.flags)
                if last == self.nsteps - 1:
                    self._finish(st)
        th.join()
        wall = time.perf_counter() - t0
        k = self.nsteps - self.step0
        return dict(steps=k, batches=nb, wall_s=wall, gpu_s=t_gpu,
                    steps_per_s=k / wall if wall > 0 else math.inf, lh_events=list(self.lh_events),
                    nc_lh_flag=self.flags.tolist(), seed=self.seed)

    def _finish(self, st: State):
        """Restart files, LH.in and the checkpoints after the last step
        (main.cpp:502-509)."""
        write_restart(self.out, st, self.Er)
        write_lh_in(os.path.join(self.out, "LH.in"), self.flags)
        words, idx, cached, n2 = self.rng.get_state()
        write_rng_state(os.path.join(self.out, RNG_FILE), self.seed, words, idx, cached, n2)
        if self.write_checkpoint:
            get_derived = getattr(self.eng, "get_derived", None)
            der = get_derived() if get_derived else {}
            np.savez(os.path.join(self.out, CHECKPOINT_FILE), r=st.r, a=st.a, b=st.b, c=st.c,
                     r_t=st.r_t, r_lh=st.r_lh, nc_lh_flag=self.flags, step=self.nsteps,
                     seed=self.seed, Er=self.Er, rng_words=np.asarray(words, np.uint32),
                     rng_idx=idx, rng_cached=cached, rng_n2=n2,
                     **{"der_" + k: np.asarray(v) for k, v in der.items()})

    def close(self):
        close = getattr(self.eng, "close", None)
        if close:
            close()


def main(argv=None):
    ap = argparse.ArgumentParser(description="Brownian dynamics run over a reference input deck")
    ap.add_argument("deck", help="directory with setup.txt, data_mod, core/tail/LH files, LH.in")
    ap.add_argument("--out", default=None, help="output directory (default: the deck)")
    ap.add_argument("--seed", type=int, default=None, help="MT19937 seed (default: time)")
    ap.add_argument("--steps", type=int, default=None, help="override number_of_steps")
    ap.add_argument("--hi-set", choices=("all", "main"), default="all",
                    help="beads in the hydrodynamic tensor")
    ap.add_argument("--mixed", action="store_true", help="FP32 trailing updates in the factor")
    ap.add_argument("--use-frequency-rp", action="store_true")
    ap.add_argument("--fix-stale-frames", action="store_true")
    ap.add_argument("--resume-rng", action="store_true")
    ap.add_argument("--resume", default=None, help="checkpoint.npz to continue from")
    args = ap.parse_args(argv)
    from . import engine as E
    run = Run(args.deck, out=args.out, seed=args.seed, steps=args.steps,
              hi_all=1 if args.hi_set == "all" else 0,
              chol_mode=E.CHOL_FP32 if args.mixed else E.CHOL_FP64,
              use_frequency_rp=args.use_frequency_rp, fix_stale_frames=args.fix_stale_frames,
              resume_rng=args.resume_rng, resume=args.resume)
    try:
        res = run.run()
    finally:
        run.close()
    print(f"{res['steps']} steps in {res['wall_s']:.2f} s ({res['steps_per_s']:.1f} steps/s), "
          f"{len(res['lh_events'])} linker-histone moves accepted, seed {res['seed']}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
