"""Python host mirror of the step boundary: ctypes over libbdhip.so.

``BDEngine`` is the object-oriented face of the C-ABI in include/bd_hip.h,
which itself stands where the reference's func_cuda.h entry points stand
(``cuda_application_init_data`` -> ``BDEngine(...)``, ``main_cuda`` ->
``step``, ``free_all`` -> ``close``).  Everything runs in the HIP library;
if the library is missing this module raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .system import Fiber, State

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbdhip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

CHOL_FP64, CHOL_FP32, CHOL_BF16 = 0, 1, 2
TERM_MECH, TERM_ELEC, TERM_TAIL, TERM_LH, TERM_ALL = 1, 2, 4, 8, 15


class BdSystem(C.Structure):
    """Mirror of bd_system (include/bd_hip.h)."""
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


class BdOptions(C.Structure):
    _fields_ = [("device", C.c_int), ("chol_mode", C.c_int), ("use_graph", C.c_int),
                ("lh_capacity", C.c_int)]


class BdDerived(C.Structure):
    _fields_ = [(k, _dp) for k in ("alpha", "beta", "gamma", "length", "a_dna", "b_dna",
                                   "c_dna", "alpha_p", "beta_p", "gamma_p", "beta_t",
                                   "beta_lh")]


class BdKernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("total_ms", C.c_double), ("launches", C.c_longlong)]


class BdWorkCounts(C.Structure):
    _fields_ = [("core_pairs", C.c_longlong), ("core_charge_pairs", C.c_longlong),
                ("core_linker_pairs", C.c_longlong)]


# Every extern "C" symbol of include/bd_hip.h with its argument types.
API = {
    "bd_create": ([C.POINTER(BdSystem), C.POINTER(BdOptions), C.POINTER(C.c_void_p)], C.c_int),
    "bd_destroy": ([C.c_void_p], C.c_int),
    "bd_set_state": ([C.c_void_p] + [_dp] * 6 + [C.c_int], C.c_int),
    "bd_get_state": ([C.c_void_p] + [_dp] * 6, C.c_int),
    "bd_set_lh": ([C.c_void_p, _ip, C.c_int, _dp, _dp], C.c_int),
    "bd_set_integrator": ([C.c_void_p] + [C.c_double] * 5, C.c_int),
    "bd_set_derived": ([C.c_void_p, C.POINTER(BdDerived)], C.c_int),
    "bd_get_derived": ([C.c_void_p, C.POINTER(BdDerived)], C.c_int),
    "bd_step": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "bd_step_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "bd_forces": ([C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp], C.c_int),
    "bd_factor": ([C.c_void_p], C.c_int),
    "bd_get_factor": ([C.c_void_p, _dp, _ip], C.c_int),
    "bd_get_tensor": ([C.c_void_p, _dp], C.c_int),
    "bd_noise_displacement": ([C.c_void_p, _dp, _dp], C.c_int),
    "bd_set_chol_mode": ([C.c_void_p, C.c_int], C.c_int),
    "bd_set_chol_schedule": ([C.c_void_p, C.c_int], C.c_int),
    "bd_set_cell_list": ([C.c_void_p, C.c_int], C.c_int),
    "bd_set_refine": ([C.c_void_p, C.c_int], C.c_int),
    "bd_cc_forward": ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
                      + [C.c_double] * 4 + [C.c_void_p] * 3, C.c_int),
    "bd_cc_backward": ([C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
                       + [C.c_double] * 4 + [C.c_void_p] * 5, C.c_int),
    "bd_cholesky": ([C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, _ip, C.c_void_p],
                    C.c_int),
    "bd_cholesky_noalloc": ([C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, _ip,
                             C.c_void_p], C.c_int),
    "bd_get_chol_schedule": ([C.c_void_p, _ip], C.c_int),
    "bd_set_frequency_rp": ([C.c_void_p, C.c_int], C.c_int),
    "bd_run_steps": ([C.c_void_p, C.c_int, _dp, _dp, _dp], C.c_int),
    "bd_set_timing": ([C.c_void_p, C.c_int], C.c_int),
    "bd_get_timing": ([C.c_void_p, C.POINTER(BdKernelTime), C.c_int, _ip], C.c_int),
    "bd_get_work_counts": ([C.c_void_p, C.POINTER(BdWorkCounts)], C.c_int),
    "bd_get_sizes": ([C.c_void_p, _ip, _ip, _ip, _ip], C.c_int),
    "bd_device_alloc": ([C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
    "bd_device_free": ([C.c_void_p], C.c_int),
    "bd_memcpy_htod": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "bd_memcpy_dtoh": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "bd_device_copy": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "bd_synchronize": ([C.c_void_p], C.c_int),
    "bd_rng_create": ([C.c_uint, C.POINTER(C.c_void_p)], C.c_int),
    "bd_rng_destroy": ([C.c_void_p], C.c_int),
    "bd_rng_u32": ([C.c_void_p], C.c_uint),
    "bd_rng_uniform": ([C.c_void_p], C.c_double),
    "bd_rng_normal": ([C.c_void_p, C.c_double, C.c_double], C.c_double),
    "bd_rng_step_noise": ([C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, _dp], C.c_int),
    "bd_rng_get_state": ([C.c_void_p, C.POINTER(C.c_uint), _ip, _ip, _dp], C.c_int),
    "bd_rng_set_state": ([C.c_void_p, C.POINTER(C.c_uint), C.c_int, C.c_int, C.c_double], C.c_int),
    "bd_last_error": ([], C.c_char_p),
    "bd_version": ([], C.c_int),
}

_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libbdhip.so; raises if it was not built (no fallback path)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not found -- build it with "
                               "`python -m brownian_dynamics_cuda_amd.build`")
        lib = C.CDLL(path)
        for name, (args, res) in API.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = lib
    return _LIB


class BDError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != 0:
        msg = load_library().bd_last_error().decode(errors="replace")
        raise BDError(f"{what} failed ({rc}): {msg}")


def _d(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    return a


def _pd(a):
    return a.ctypes.data_as(_dp) if a is not None else C.cast(None, _dp)


class _Keep:
    def __init__(self):
        self.bufs = []

    def d(self, x):
        a = _d(x)
        if a.size == 0:
            a = np.zeros(1)
        self.bufs.append(a)
        return _pd(a)

    def i(self, x):
        a = np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))
        if a.size == 0:
            a = np.zeros(1, np.int32)
        self.bufs.append(a)
        return a.ctypes.data_as(_ip)


def system_struct(fib: Fiber):
    """bd_system for a Fiber (buffers kept alive by the returned keeper)."""
    k = _Keep()
    ph = fib.phys
    s = BdSystem()
    s.n, s.n_c, s.n_tail, s.n_t = fib.n, fib.n_c, fib.n_tail, fib.n_t
    s.n_lh_n, s.n_lh_g, s.n_lh_c, s.n_LH = fib.n_lh_n, fib.n_lh_g, fib.n_lh_c, fib.n_LH
    s.Nq, s.groups_per_core, s.hi_all = fib.Nq, fib.groups_per_core, fib.hi_all
    s.type = k.i(fib.types)
    s.phi_o, s.rad = k.d(fib.phi_o), k.d(fib.rad)
    s.core_pos, s.core_q = k.d(fib.core_pos), k.d(fib.core_q)
    s.tail_pos, s.tail_fix_tpl = k.d(fib.tail_pos), k.i(fib.tail_fix_tpl)
    for name in ("h_t", "g_t", "lo_t", "beta_o_t", "t_q", "t_rad"):
        setattr(s, name, k.d(getattr(fib, name)))
    s.t_grp, s.t_fix = k.i(fib.t_grp), k.i(fib.t_fix)
    s.LH_g_pos, s.LH_conn = k.d(fib.LH_g_pos), k.i(fib.LH_conn)
    for name in ("LH_q", "LH_vdw_hh", "LH_vdw_hc", "LH_vdw_hl", "LH_vdw_ht", "LH_kstr",
                 "LH_kben", "LH_streq", "LH_betaeq"):
        setattr(s, name, k.d(getattr(fib, name)))
    s.nc_lh_flag, s.lh_rad = k.i(fib.nc_lh_flag), k.d(fib.lh_rad)
    for name in ("h", "g", "s", "debyell", "debye", "q_l", "k_e", "k_ex", "sig_dd", "sig_dc",
                 "sig_cc", "sig_tt", "sig_tl", "sig_tc", "a1", "a2", "s2dt", "time_step"):
        setattr(s, name, getattr(ph, name))
    s.del_ = ph.dell
    return s, k


DERIVED_KEYS = ("alpha", "beta", "gamma", "length", "a_dna", "b_dna", "c_dna", "alpha_p",
                "beta_p", "gamma_p", "beta_t", "beta_lh")


class BDEngine:
    """One fiber resident on one GPU."""

    def __init__(self, fib: Fiber, chol_mode: int = CHOL_FP64, use_graph: bool = True,
                 device: int = 0, lh_capacity: int = 0):
        self.lib = load_library()
        self.fib = fib
        sysc, self._keep = system_struct(fib)
        opt = BdOptions(device, chol_mode, 1 if use_graph else 0, lh_capacity)
        h = C.c_void_p()
        _check(self.lib.bd_create(C.byref(sysc), C.byref(opt), C.byref(h)), "bd_create")
        self.h = h
        self.n, self.n_tail, self.n_LH = fib.n, fib.n_tail, fib.n_LH
        self.n_hi = fib.n_hi
        self.n_D3 = 3 * (fib.n + fib.n_tail + fib.n_LH)

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.bd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state ---------------------------------------------------------------
    def set_state(self, st: State, rebuild: bool = True):
        bufs = [_d(x) for x in (st.r, st.a, st.b, st.c, st.r_t, st.r_lh)]
        ptrs = [_pd(b) if b.size else C.cast(None, _dp) for b in bufs]
        _check(self.lib.bd_set_state(self.h, *ptrs, 1 if rebuild else 0), "bd_set_state")

    def get_state(self) -> State:
        n, nt, nl = self.n, self.n_tail, self.n_LH
        out = [np.zeros(3 * n) for _ in range(4)] + [np.zeros(max(3 * nt, 1)),
                                                     np.zeros(max(3 * nl, 1))]
        _check(self.lib.bd_get_state(self.h, *[_pd(o) for o in out]), "bd_get_state")
        return State(out[0], out[1], out[2], out[3], out[4][:3 * nt], out[5][:3 * nl])

    def _derived_bufs(self):
        n, nc, nt, nl = self.n, self.fib.n_c, self.n_tail, self.n_LH
        sizes = dict(alpha=n, beta=n, gamma=n, length=n, a_dna=3 * nc, b_dna=3 * nc,
                     c_dna=3 * nc, alpha_p=nc, beta_p=nc, gamma_p=nc, beta_t=nt, beta_lh=nl)
        return {k: np.zeros(max(v, 1)) for k, v in sizes.items()}, sizes

    def get_derived(self) -> dict:
        bufs, sizes = self._derived_bufs()
        d = BdDerived(*[_pd(bufs[k]) for k in DERIVED_KEYS])
        _check(self.lib.bd_get_derived(self.h, C.byref(d)), "bd_get_derived")
        return {k: bufs[k][:sizes[k]] for k in DERIVED_KEYS}

    def set_derived(self, vals: dict):
        bufs, sizes = self._derived_bufs()
        for k in DERIVED_KEYS:
            v = _d(vals[k])[:sizes[k]]
            bufs[k][:len(v)] = v
        d = BdDerived(*[_pd(bufs[k]) for k in DERIVED_KEYS])
        _check(self.lib.bd_set_derived(self.h, C.byref(d)), "bd_set_derived")

    def set_lh(self, nc_lh_flag, r_lh, lh_rad):
        fl = np.ascontiguousarray(np.asarray(nc_lh_flag, np.int32))
        r = _d(r_lh)
        rad = _d(lh_rad)
        n_LH = len(rad)
        _check(self.lib.bd_set_lh(self.h, fl.ctypes.data_as(_ip), n_LH, _pd(r), _pd(rad)),
               "bd_set_lh")
        self.n_LH = n_LH
        self.n_D3 = 3 * (self.n + self.n_tail + n_LH)

    # -- the step --------------------------------------------------------------
    def step(self, rr, p) -> float:
        rr, p = _d(rr), _d(p)
        if rr.size != 3 * self.n or p.size != self.n_D3:
            raise ValueError(f"rr needs {3 * self.n} and p {self.n_D3} values")
        e = C.c_double(0.0)
        _check(self.lib.bd_step(self.h, _pd(rr), _pd(p), C.byref(e)), "bd_step")
        return e.value

    def step_device(self, d_rr: int, d_p: int, d_energy: int = 0):
        _check(self.lib.bd_step_device(self.h, C.c_void_p(d_rr), C.c_void_p(d_p),
                                       C.c_void_p(d_energy) if d_energy else None),
               "bd_step_device")

    def run_steps(self, rr, p):
        """k steps with host noise rr[k, 3n], p[k, n_D3]; returns the k energies."""
        rr = np.ascontiguousarray(rr, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        k = rr.shape[0]
        if rr.shape != (k, 3 * self.n) or p.shape != (k, self.n_D3):
            raise ValueError(f"rr needs shape ({k}, {3 * self.n}) and p ({k}, {self.n_D3})")
        e = np.zeros(max(k, 1))
        _check(self.lib.bd_run_steps(self.h, k, _pd(rr), _pd(p), _pd(e)), "bd_run_steps")
        return e[:k]

    def synchronize(self):
        _check(self.lib.bd_synchronize(self.h), "bd_synchronize")

    def forces(self, terms: int = TERM_ALL) -> dict:
        n, nt, nl = self.n, self.n_tail, self.n_LH
        F, T = np.zeros(3 * n), np.zeros(3 * n)
        tF, lF = np.zeros(max(3 * nt, 1)), np.zeros(max(3 * nl, 1))
        e = C.c_double(0.0)
        _check(self.lib.bd_forces(self.h, terms, _pd(F), _pd(T), _pd(tF), _pd(lF), C.byref(e)),
               "bd_forces")
        return dict(F=F, T=T, tF=tF[:3 * nt], lhF=lF[:3 * nl], E=e.value)

    def factor(self):
        _check(self.lib.bd_factor(self.h), "bd_factor")

    def get_factor(self):
        m = 3 * self.n_hi
        L = np.zeros((m, m))
        neg = C.c_int(0)
        _check(self.lib.bd_get_factor(self.h, _pd(L), C.byref(neg)), "bd_get_factor")
        return L, neg.value

    def get_tensor(self):
        m = 3 * self.n_hi
        D = np.zeros((m, m))
        _check(self.lib.bd_get_tensor(self.h, _pd(D)), "bd_get_tensor")
        return D

    def noise_displacement(self, p):
        p = _d(p)
        rd = np.zeros(self.n_D3)
        _check(self.lib.bd_noise_displacement(self.h, _pd(p), _pd(rd)), "bd_noise_displacement")
        return rd

    def set_chol_mode(self, mode: int):
        _check(self.lib.bd_set_chol_mode(self.h, mode), "bd_set_chol_mode")

    def set_chol_schedule(self, panel_tiles: int):
        """Super-panel width of the factor in 64-column block columns: 0
        chooses by size, -1 one block column per launch, k >= 2 always k."""
        _check(self.lib.bd_set_chol_schedule(self.h, int(panel_tiles)), "bd_set_chol_schedule")

    def set_cell_list(self, enable: bool):
        """Partner search of the 7 nm terms: cell grids (True; the default
        from 20000 tail and LH beads up) or bounding spheres of 16
        consecutive partners."""
        _check(self.lib.bd_set_cell_list(self.h, 1 if enable else 0), "bd_set_cell_list")

    def set_refine(self, enable: bool):
        """FP64 refinement of the noise displacement in the FP32 / BF16
        factor modes (on by default; FP64 mode never refines)."""
        _check(self.lib.bd_set_refine(self.h, 1 if enable else 0), "bd_set_refine")

    def chol_schedule(self) -> int:
        """Super-panel width in use (0: one block column per launch)."""
        v = C.c_int(0)
        _check(self.lib.bd_get_chol_schedule(self.h, C.byref(v)), "bd_get_chol_schedule")
        return v.value

    def set_frequency_rp(self, frequency_rp: int):
        """Rebuild D and L every `frequency_rp` steps (1: every step)."""
        _check(self.lib.bd_set_frequency_rp(self.h, int(frequency_rp)), "bd_set_frequency_rp")

    def set_timing(self, on: bool):
        _check(self.lib.bd_set_timing(self.h, 1 if on else 0), "bd_set_timing")

    def timing(self) -> dict:
        arr = (BdKernelTime * 64)()
        cnt = C.c_int(0)
        _check(self.lib.bd_get_timing(self.h, arr, 64, C.byref(cnt)), "bd_get_timing")
        return {arr[i].name.decode(): (arr[i].total_ms, arr[i].launches) for i in range(cnt.value)}

    def work_counts(self) -> dict:
        w = BdWorkCounts()
        _check(self.lib.bd_get_work_counts(self.h, C.byref(w)), "bd_get_work_counts")
        return dict(core_pairs=w.core_pairs, core_charge_pairs=w.core_charge_pairs)


class DeviceBuffer:
    """A device allocation through the library (no torch needed)."""

    def __init__(self, nbytes: int):
        self.lib = load_library()
        p = C.c_void_p()
        _check(self.lib.bd_device_alloc(nbytes, C.byref(p)), "bd_device_alloc")
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, arr: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        _check(self.lib.bd_memcpy_htod(C.c_void_p(self.ptr + offset), a.ctypes.data, a.nbytes),
               "bd_memcpy_htod")

    def download(self, shape, dtype=np.float64, offset: int = 0):
        out = np.zeros(shape, dtype)
        assert offset + out.nbytes <= self.nbytes
        _check(self.lib.bd_memcpy_dtoh(out.ctypes.data, C.c_void_p(self.ptr + offset),
                                       out.nbytes), "bd_memcpy_dtoh")
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int):
        """Device-to-device copy of nbytes (the library's streaming copy kernel)."""
        assert nbytes <= min(self.nbytes, other.nbytes)
        _check(self.lib.bd_device_copy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), nbytes),
               "bd_device_copy")

    def free(self):
        if self.ptr:
            self.lib.bd_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def cholesky_device(D: np.ndarray, mode: int = CHOL_FP64, entry: str = "bd_cholesky"):
    """L and the negative-pivot count of a symmetric matrix, factored on the
    GPU from a device copy of D through the C-ABI entry point `entry`."""
    lib = load_library()
    D = np.ascontiguousarray(D, np.float64)
    m = D.shape[0]
    dD, dL = DeviceBuffer(D.nbytes), DeviceBuffer(D.nbytes)
    try:
        dD.upload(D)
        neg = C.c_int(-1)
        _check(getattr(lib, entry)(m, dD.ptr, m, dL.ptr, m, mode, C.byref(neg), None), entry)
        L = dL.download((m, m))
    finally:
        dD.free()
        dL.free()
    return L, neg.value


class HostRng:
    """Host noise of the step (bd_rng_* in libbdhip.so)."""

    def __init__(self, seed: int = 5489):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.bd_rng_create(seed, C.byref(h)), "bd_rng_create")
        self.h = h

    def u32(self) -> int:
        return self.lib.bd_rng_u32(self.h)

    def uniform(self) -> float:
        return self.lib.bd_rng_uniform(self.h)

    def normal(self, mean=0.0, sd=1.0) -> float:
        return self.lib.bd_rng_normal(self.h, mean, sd)

    def step_noise(self, step: int, r_var, n_D3: int):
        rv = _d(r_var)
        rr = np.zeros(rv.size)
        p = np.zeros(n_D3)
        u = np.zeros(2)
        _check(self.lib.bd_rng_step_noise(self.h, step, rv.size, _pd(rv), n_D3, _pd(rr), _pd(p),
                                          _pd(u)), "bd_rng_step_noise")
        return rr, p

    def get_state(self):
        words = (C.c_uint * 624)()
        idx, cached, n2 = C.c_int(), C.c_int(), C.c_double()
        _check(self.lib.bd_rng_get_state(self.h, words, C.byref(idx), C.byref(cached),
                                         C.byref(n2)), "bd_rng_get_state")
        return np.array(words, np.uint32), idx.value, cached.value, n2.value

    def set_state(self, words, idx, cached, n2):
        w = (C.c_uint * 624)(*[int(x) for x in words])
        _check(self.lib.bd_rng_set_state(self.h, w, idx, cached, n2), "bd_rng_set_state")

    def close(self):
        if getattr(self, "h", None):
            self.lib.bd_rng_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
