"""ctypes bindings of the CPU oracle (liboracle.so) and of the compiled
reference library (_ref/libbdref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# BD_ORACLE_SO selects another build of the oracle (the sanitizer builds of
# oracle/Makefile's sanitize target)
ORACLE_SO = os.environ.get("BD_ORACLE_SO") or os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbdref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class CSystem(C.Structure):
    _fields_ = ([(k, C.c_int) for k in
                 ("n", "n_c", "n_tail", "n_t", "n_lh_n", "n_lh_g", "n_lh_c", "n_LH", "Nq",
                  "groups_per_core", "hi_all")]
                + [("type", _ip), ("phi_o", _dp), ("rad", _dp), ("core_pos", _dp),
                   ("core_q", _dp), ("tail_pos", _dp), ("tail_fix_tpl", _ip)]
                + [(k, _dp) for k in ("h_t", "g_t", "lo_t", "beta_o_t", "t_q", "t_rad")]
                + [("t_grp", _ip), ("t_fix", _ip), ("LH_g_pos", _dp), ("LH_conn", _ip)]
                + [(k, _dp) for k in ("LH_q", "LH_vdw_hh", "LH_vdw_hc", "LH_vdw_hl",
                                      "LH_vdw_ht", "LH_kstr", "LH_kben", "LH_streq",
                                      "LH_betaeq")]
                + [("nc_lh_flag", _ip), ("lh_rad", _dp)]
                + [(k, C.c_double) for k in
                   ("h", "g", "s", "debyell", "debye", "q_l", "k_e", "k_ex", "sig_dd",
                    "sig_dc", "sig_cc", "sig_tt", "sig_tl", "sig_tc", "a1", "a2", "s2dt",
                    "del_", "time_step")])


STATE_FIELDS = ("r", "a", "b", "c", "alpha", "beta", "gamma", "length", "a_dna", "b_dna",
                "c_dna", "alpha_p", "beta_p", "gamma_p", "r_t", "beta_t", "r_lh", "beta_lh")


class CState(C.Structure):
    _fields_ = [(k, _dp) for k in STATE_FIELDS]


class CStepOut(C.Structure):
    _fields_ = [(k, _dp) for k in ("D", "L", "rd", "F0", "T0", "tF0", "lhF0", "F1", "T1",
                                   "tF1", "lhF1")] + [("neg_pivots", C.c_int)]


class CRng(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int), ("cached", C.c_int),
                ("n2", C.c_double)]


def _d(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _i(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _ptr(a, ty=_dp):
    if a is None:
        return C.cast(None, ty)
    return a.ctypes.data_as(ty)


class SysBinding:
    """Keeps the numpy buffers behind a CSystem alive."""

    def __init__(self, fib):
        ph = fib.phys
        self.keep = {}

        def d(name, x):
            a = _d(np.asarray(x).reshape(-1))
            if a.size == 0:
                a = np.zeros(1)
            self.keep[name] = a
            return _ptr(a)

        def i(name, x):
            a = _i(np.asarray(x).reshape(-1))
            if a.size == 0:
                a = np.zeros(1, np.int32)
            self.keep[name] = a
            return _ptr(a, _ip)

        s = CSystem()
        s.n, s.n_c, s.n_tail, s.n_t = fib.n, fib.n_c, fib.n_tail, fib.n_t
        s.n_lh_n, s.n_lh_g, s.n_lh_c, s.n_LH = fib.n_lh_n, fib.n_lh_g, fib.n_lh_c, fib.n_LH
        s.Nq, s.groups_per_core, s.hi_all = fib.Nq, fib.groups_per_core, fib.hi_all
        s.type = i("type", fib.types)
        s.phi_o = d("phi_o", fib.phi_o)
        s.rad = d("rad", fib.rad)
        s.core_pos = d("core_pos", fib.core_pos)
        s.core_q = d("core_q", fib.core_q)
        s.tail_pos = d("tail_pos", fib.tail_pos)
        s.tail_fix_tpl = i("tail_fix_tpl", fib.tail_fix_tpl)
        for k in ("h_t", "g_t", "lo_t", "beta_o_t", "t_q", "t_rad"):
            setattr(s, k, d(k, getattr(fib, k)))
        s.t_grp = i("t_grp", fib.t_grp)
        s.t_fix = i("t_fix", fib.t_fix)
        s.LH_g_pos = d("LH_g_pos", fib.LH_g_pos)
        s.LH_conn = i("LH_conn", fib.LH_conn)
        for k in ("LH_q", "LH_vdw_hh", "LH_vdw_hc", "LH_vdw_hl", "LH_vdw_ht", "LH_kstr",
                  "LH_kben", "LH_streq", "LH_betaeq"):
            setattr(s, k, d(k, getattr(fib, k)))
        s.nc_lh_flag = i("nc_lh_flag", fib.nc_lh_flag)
        s.lh_rad = d("lh_rad", fib.lh_rad)
        for k in ("h", "g", "s", "debyell", "debye", "q_l", "k_e", "k_ex", "sig_dd", "sig_dc",
                  "sig_cc", "sig_tt", "sig_tl", "sig_tc", "a1", "a2", "s2dt", "time_step"):
            setattr(s, k, getattr(ph, k))
        s.del_ = ph.dell
        self.c = s


class FullState:
    """Positions, frames and every derived array, as numpy buffers."""

    def __init__(self, fib, state=None):
        n, nc, nt, nl = fib.n, fib.n_c, fib.n_tail, fib.n_LH
        z = lambda k: np.zeros(max(k, 1))
        self.r, self.a, self.b, self.c = z(3 * n), z(3 * n), z(3 * n), z(3 * n)
        self.alpha, self.beta, self.gamma, self.length = z(n), z(n), z(n), z(n)
        self.a_dna, self.b_dna, self.c_dna = z(3 * nc), z(3 * nc), z(3 * nc)
        self.alpha_p, self.beta_p, self.gamma_p = z(nc), z(nc), z(nc)
        self.r_t, self.beta_t = z(3 * nt + 3), z(nt + 2)
        self.r_lh, self.beta_lh = z(3 * nl + 3), z(nl + 2)
        self.sizes = dict(n=n, nc=nc, nt=nt, nl=nl)
        if state is not None:
            self.r[:3 * n] = state.r
            self.a[:3 * n] = state.a
            self.b[:3 * n] = state.b
            self.c[:3 * n] = state.c
            self.r_t[:3 * nt] = state.r_t
            self.r_lh[:3 * nl] = state.r_lh

    def cstruct(self):
        return CState(*[_ptr(getattr(self, k)) for k in STATE_FIELDS])

    def copy(self):
        o = FullState.__new__(FullState)
        for k in STATE_FIELDS:
            setattr(o, k, getattr(self, k).copy())
        o.sizes = dict(self.sizes)
        return o

    def view(self, k):
        s = self.sizes
        lens = dict(r=3 * s["n"], a=3 * s["n"], b=3 * s["n"], c=3 * s["n"], alpha=s["n"],
                    beta=s["n"], gamma=s["n"], length=s["n"], a_dna=3 * s["nc"],
                    b_dna=3 * s["nc"], c_dna=3 * s["nc"], alpha_p=s["nc"], beta_p=s["nc"],
                    gamma_p=s["nc"], r_t=3 * s["nt"], beta_t=s["nt"], r_lh=3 * s["nl"],
                    beta_lh=s["nl"])
        return getattr(self, k)[:lens[k]]


def _setup(lib, prefix):
    sp, stp = C.POINTER(CSystem), C.POINTER(CState)
    if prefix == "or":
        lib.or_update_euler.argtypes = [sp] + [_dp] * 14
        lib.or_mech.argtypes = [sp, stp, _dp, _dp, _dp]
        lib.or_elec.argtypes = [sp, stp, _dp, _dp, _dp]
        lib.or_tail.argtypes = [sp, stp, _dp, _dp, _dp, _dp]
        lib.or_lh.argtypes = [sp, stp, _dp, _dp, _dp, _dp, _dp]
        lib.or_forces.argtypes = [sp, stp, _dp, _dp, _dp, _dp, _dp]
        lib.or_step.argtypes = [sp, stp, _dp, _dp, _dp, C.POINTER(CStepOut)]
        lib.or_step.restype = C.c_int
        lib.or_step_reuse.argtypes = [sp, stp, _dp, _dp, _dp, _dp, C.POINTER(CStepOut)]
        lib.or_step_reuse.restype = C.c_int
        lib.or_diffusion_tensor.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp]
        lib.or_cholesky.argtypes = [C.c_int, _dp, _dp]
        lib.or_cholesky.restype = C.c_int
        lib.or_rotate.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]
        lib.or_update_tail_beta.argtypes = [C.c_int, _dp, _dp, _ip]
        lib.or_update_lh_beta.argtypes = [C.c_int, C.c_int, _dp, _dp]
        lib.or_twist_angle_phi.argtypes = [C.c_int]
        lib.or_twist_angle_phi.restype = C.c_double
        lib.or_rng_seed.argtypes = [C.POINTER(CRng), C.c_uint32]
        lib.or_rng_u32.argtypes = [C.POINTER(CRng)]
        lib.or_rng_u32.restype = C.c_uint32
        lib.or_rng_real1.argtypes = [C.POINTER(CRng)]
        lib.or_rng_real1.restype = C.c_double
        lib.or_rng_normal.argtypes = [C.POINTER(CRng), C.c_double, C.c_double]
        lib.or_rng_normal.restype = C.c_double
        lib.or_rng_step_noise.argtypes = [C.POINTER(CRng), C.c_int, C.c_int, _dp, C.c_int, _dp,
                                          _dp]
    else:
        lib.ref_update_euler.argtypes = [sp, stp]
        lib.ref_mech.argtypes = [sp, stp, _dp, _dp, _dp]
        lib.ref_elec.argtypes = [sp, stp, _dp, _dp, _dp]
        lib.ref_tail.argtypes = [sp, stp, _dp, _dp, _dp, _dp]
        lib.ref_lh.argtypes = [sp, stp, _dp, _dp, _dp, _dp, _dp]
        lib.ref_forces.argtypes = [sp, stp, _dp, _dp, _dp, _dp, _dp]
        lib.ref_step.argtypes = [sp, stp, _dp, _dp, _dp]
        lib.ref_diffusion_tensor.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        lib.ref_rotate.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]
        lib.ref_update_tail_beta.argtypes = [C.c_int, _dp, _dp, _ip]
        lib.ref_update_lh_beta.argtypes = [C.c_int, C.c_int, _dp, _dp]
        lib.ref_init_genrand.argtypes = [C.c_ulong]
        lib.ref_genrand_real1.restype = C.c_double
        lib.ref_rand_normal.argtypes = [C.c_double, C.c_double]
        lib.ref_rand_normal.restype = C.c_double
        lib.ref_twist_angle_phi.argtypes = [C.c_int]
        lib.ref_twist_angle_phi.restype = C.c_double
        lib.ref_build_lh.argtypes = [C.c_int, _ip, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp,
                                     _dp, _dp, _dp, _dp, _ip, _dp, _dp]
        lib.ref_write_xyz_append.argtypes = ([C.c_int] * 5 + [_ip] + [C.c_int] * 2 + [_ip]
                                             + [_dp] * 8 + [C.c_char_p])
        lib.ref_write_restart.argtypes = [C.c_int] * 3 + [_dp] * 7
        lib.ref_read_restart.argtypes = [C.c_int] * 3 + [_dp] * 7
        lib.ref_initial_er.argtypes = [C.c_char_p, _dp, C.c_int]
        lib.ref_initial_er.restype = C.c_int
    return lib


_LIBS = {}


def oracle_lib():
    if "or" not in _LIBS:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        _LIBS["or"] = _setup(C.CDLL(ORACLE_SO), "or")
    return _LIBS["or"]


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    if "ref" not in _LIBS:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} missing (built only where the reference is present)")
        _LIBS["ref"] = _setup(C.CDLL(REF_SO), "ref")
    return _LIBS["ref"]


@contextlib.contextmanager
def in_tempdir():
    """The reference writes wf.txt / wrongD.txt into the working directory."""
    old = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        try:
            yield d
        finally:
            os.chdir(old)


class Oracle:
    """Runs the C restatement (kind='or') or the compiled reference (kind='ref')."""

    def __init__(self, fib, kind="or"):
        self.fib = fib
        self.kind = kind
        self.lib = oracle_lib() if kind == "or" else ref_lib()
        self.sb = SysBinding(fib)

    def _f(self, name):
        return getattr(self.lib, f"{self.kind}_{name}")

    def update_euler(self, fs: FullState):
        cs = fs.cstruct()
        if self.kind == "or":
            self.lib.or_update_euler(C.byref(self.sb.c), *[_ptr(getattr(fs, k)) for k in
                                     ("r", "a", "b", "c", "alpha", "beta", "gamma", "length",
                                      "a_dna", "b_dna", "c_dna", "alpha_p", "beta_p",
                                      "gamma_p")])
        else:
            self.lib.ref_update_euler(C.byref(self.sb.c), C.byref(cs))

    def update_betas(self, fs: FullState):
        f = self.fib
        if f.n_tail:
            self._f("update_tail_beta")(f.n_tail, _ptr(fs.r_t), _ptr(fs.beta_t),
                                        _ptr(self.sb.keep["t_grp"], _ip))
        if f.n_LH:
            self._f("update_lh_beta")(f.n_LH, f.n_lh_c, _ptr(fs.r_lh), _ptr(fs.beta_lh))

    def prepare(self, state):
        """FullState with the derived arrays of the initial state (main.cpp:164,223,283)."""
        fs = FullState(self.fib, state)
        self.update_euler(fs)
        self.update_betas(fs)
        return fs

    def forces(self, fs: FullState, terms=("all",)):
        f = self.fib
        F = np.zeros(3 * f.n)
        T = np.zeros(3 * f.n)
        tF = np.zeros(3 * f.n_tail + 3)
        lF = np.zeros(3 * f.n_LH + 3)
        E = C.c_double(0.0)
        cs = fs.cstruct()
        S = C.byref(self.sb.c)
        with in_tempdir():
            if terms == ("all",):
                self._f("forces")(S, C.byref(cs), _ptr(F), _ptr(T), _ptr(tF), _ptr(lF),
                                  C.byref(E))
            else:
                for t in terms:
                    if t == "mech":
                        self._f("mech")(S, C.byref(cs), _ptr(F), _ptr(T), C.byref(E))
                    elif t == "elec":
                        self._f("elec")(S, C.byref(cs), _ptr(F), _ptr(T), C.byref(E))
                    elif t == "tail":
                        self._f("tail")(S, C.byref(cs), _ptr(F), _ptr(T), _ptr(tF), C.byref(E))
                    elif t == "lh":
                        self._f("lh")(S, C.byref(cs), _ptr(F), _ptr(T), _ptr(tF), _ptr(lF),
                                      C.byref(E))
        return dict(F=F, T=T, tF=tF[:3 * f.n_tail], lhF=lF[:3 * f.n_LH], E=E.value)

    def step(self, fs: FullState, rr, p, want=False, L_reuse=None):
        """One step; with L_reuse (oracle only) the noise uses that factor
        of an earlier step instead of the factor of the current D."""
        rr = _d(rr)
        p = _d(p)
        E = C.c_double(0.0)
        out = {}
        with in_tempdir():
            if self.kind == "or":
                o = CStepOut()
                if want:
                    f = self.fib
                    m = 3 * f.n_hi
                    n_D3 = 3 * (f.n + f.n_tail + f.n_LH)
                    for k, sz in (("D", m * m), ("L", m * m), ("rd", n_D3), ("F0", 3 * f.n),
                                  ("T0", 3 * f.n), ("F1", 3 * f.n), ("T1", 3 * f.n),
                                  ("tF0", 3 * f.n_tail + 3), ("tF1", 3 * f.n_tail + 3),
                                  ("lhF0", 3 * f.n_LH + 3), ("lhF1", 3 * f.n_LH + 3)):
                        out[k] = np.zeros(max(sz, 1))
                        setattr(o, k, _ptr(out[k]))
                if L_reuse is None:
                    self.lib.or_step(C.byref(self.sb.c), C.byref(fs.cstruct()), _ptr(rr), _ptr(p),
                                     C.byref(E), C.byref(o))
                else:
                    Lr = _d(L_reuse).reshape(-1)
                    self.lib.or_step_reuse(C.byref(self.sb.c), C.byref(fs.cstruct()), _ptr(rr),
                                           _ptr(p), _ptr(Lr), C.byref(E), C.byref(o))
                out["neg_pivots"] = o.neg_pivots
            else:
                if L_reuse is not None:
                    raise ValueError("L_reuse is an oracle-only mode")
                self.lib.ref_step(C.byref(self.sb.c), C.byref(fs.cstruct()), _ptr(rr), _ptr(p),
                                  C.byref(E))
        out["E"] = E.value
        return out

    def diffusion_tensor(self, r, rad):
        r = _d(r)
        rad = _d(rad)
        nb = len(rad)
        m = 3 * nb
        D = np.zeros((m, m))
        L = np.zeros((m, m))
        f = self.fib.phys
        if self.kind == "or":
            self.lib.or_diffusion_tensor(nb, _ptr(r), _ptr(rad), f.a1, f.a2, _ptr(D))
            neg = self.lib.or_cholesky(m, _ptr(D), _ptr(L))
        else:
            with in_tempdir():
                self.lib.ref_diffusion_tensor(nb, _ptr(r), _ptr(rad), f.a1, f.a2, _ptr(D), _ptr(L))
            neg = None
        return D, L, neg


class Rng:
    """The oracle's MT19937 + polar sampler."""

    def __init__(self, seed=5489):
        self.lib = oracle_lib()
        self.g = CRng()
        self.lib.or_rng_seed(C.byref(self.g), seed)

    def u32(self):
        return self.lib.or_rng_u32(C.byref(self.g))

    def real1(self):
        return self.lib.or_rng_real1(C.byref(self.g))

    def normal(self, mean=0.0, sd=1.0):
        return self.lib.or_rng_normal(C.byref(self.g), mean, sd)

    def step_noise(self, step, r_var, n_D3):
        rr = np.zeros(len(r_var))
        p = np.zeros(n_D3)
        rv = _d(r_var)
        self.lib.or_rng_step_noise(C.byref(self.g), step, len(rv), _ptr(rv), n_D3, _ptr(rr),
                                   _ptr(p))
        return rr, p
