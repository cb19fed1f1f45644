"""Fiber description: static per-bead data, physics constants and state.

Host-side mirror of what the reference driver assembles before it calls the
GPU step: derived constants (``main.cpp:68-99``), the rotational noise scale
(``main.cpp:288-298``), the tail and linker-histone builders
(``func.cpp:1672-1713, 2029-2059``) and the readers of the reference input
formats (``readfile.cpp:14-337``), with the same unit scalings.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

# constants.h:6-12
PI = 3.141592653589790
LO = 3.0
RO = 4.8
D1 = 1.8
THETA = 1.2 * PI / 2
ETA = 1.137076e-12
R_H = 1.2


@dataclasses.dataclass
class Physics:
    """Scalars of the model, derived as in main.cpp:68-99."""

    T: float
    Cs: float
    Mg: int
    h: float
    g: float
    s: float
    debyell: float
    debye: float
    q_l: float
    k_e: float
    k_ex: float
    k_h1: float
    sig_dd: float
    sig_dc: float
    sig_cc: float
    sig_tt: float
    sig_tl: float
    sig_tc: float
    a1: float
    a2: float
    s2dt: float
    dell: float
    time_step: float
    kbt: float

    @staticmethod
    def derive(T: float, Cs: float, Mg: int = 0, sigmas=(3.6, 2.4, 1.2, 1.8, 2.7, 1.8),
               q_l: Optional[float] = None, time_step: float = 1e-12) -> "Physics":
        kb = 1.380649e-5
        kbt = kb * 300.0  # main.cpp:69 -- T enters only the Debye length
        cs = Cs * 1000
        ql = LO * (-5.8824) * (1.901912136835033e-8 * cs ** 3.0
                               + -8.211102025728157e-6 * cs ** 2.0
                               + 7.554037628581672e-3 * cs + 3.524292543853884e-1)
        if q_l is not None:
            ql = q_l
        debye = 0.736 * math.sqrt((Cs / 0.05) * (298 / T))
        if Mg == 0:
            per, debyell = 50.0, debye
        else:
            per, debyell = 30.0, 2.5
        return Physics(
            T=T, Cs=Cs, Mg=Mg,
            h=100 * kbt / (LO * LO), g=per * kbt / LO, s=72.429 * kbt / LO,
            debyell=debyell, debye=debye, q_l=ql,
            k_e=0.4151 * kb * 300, k_ex=0.001 * kbt, k_h1=10.0 * kbt,
            sig_dd=sigmas[0], sig_dc=sigmas[1], sig_cc=sigmas[2],
            sig_tt=sigmas[3], sig_tl=sigmas[4], sig_tc=sigmas[5],
            a1=kbt / (6 * PI * ETA), a2=kbt / (8 * PI * ETA),
            s2dt=math.sqrt(2 * time_step), dell=time_step / kbt, time_step=time_step, kbt=kbt)


def twist_angle_phi(w: int) -> float:
    """func.cpp:13-99."""
    table = {2: 0.9007, 3: -0.6754, 4: -0.3519, 5: -0.1463, 6: 0.0215, 7: 0.1178,
             8: 0.2025, 9: 0.2726, 10: -0.2415, 11: -0.15}
    if w <= 11:
        if w < 2:
            raise ValueError("linker of fewer than 2 beads (the reference exits here)")
        return table[w]
    nbp = (w + 1) * 3 / 0.34
    per_bp = round(nbp / 10.3) * 360 / nbp
    whole = nbp * (per_bp - 34.95)
    return PI * (whole / (w + 1)) / 180.00


def phi_o_of(types: np.ndarray) -> np.ndarray:
    """update_phi_o, func.cpp:412-431.

    A fiber that ends with a core (n = n_c + 6(n_c - 1)) gives that core no
    linker; the reference would stop in twist_angle_phi(0).  Its phi_o entry
    is never read by the step (segments end at n-2), so it is set to 0."""
    n = len(types)
    phi = np.zeros(n)
    for i in range(n):
        if types[i] == 1:
            k = 1
            while i + k < n and types[i + k] == 0:
                k += 1
            if k == 1 and i == n - 1:
                phi[i] = 0.0
                continue
            phi[i:i + k] = twist_angle_phi(k - 1)
    return phi


def rot_var(types: np.ndarray, ph: Physics) -> np.ndarray:
    """Rotational noise scale, main.cpp:288-298 (a variance passed as sigma)."""
    n = len(types)
    rv = np.zeros(3 * n)
    core = 2.0 * ph.time_step * ph.kbt / (8 * PI * ETA * 125.0)
    lnk = 2.0 * ph.time_step * ph.kbt / (4 * PI * ETA * R_H * R_H * LO)
    for i in range(n):
        if types[i] != 0:
            rv[3 * i:3 * i + 3] = core
        else:
            rv[3 * i] = lnk
    return rv


@dataclasses.dataclass
class Fiber:
    """Static description of one fiber (everything the step reads but never writes)."""

    types: np.ndarray          # int32[n]
    rad: np.ndarray            # [n]
    phi_o: np.ndarray          # [n]
    core_pos: np.ndarray       # [Nq,3] nm body frame
    core_q: np.ndarray         # [Nq]
    # tail template (per core) and expanded per-bead arrays
    tail_pos: np.ndarray       # [n_t,3]
    tail_fix_tpl: np.ndarray   # int32[n_t]
    t_q: np.ndarray            # [n_tail]
    t_rad: np.ndarray
    h_t: np.ndarray
    g_t: np.ndarray
    lo_t: np.ndarray
    beta_o_t: np.ndarray
    t_grp: np.ndarray          # int32[n_tail]
    t_fix: np.ndarray          # int32[n_tail]
    # linker histone template: rows N-term, G-head, C-term
    n_lh_n: int
    n_lh_g: int
    n_lh_c: int
    LH_g_pos: np.ndarray       # [n_lh_g,3]
    LH_n_pos: np.ndarray       # [n_lh_n,3]
    LH_c_pos: np.ndarray       # [n_lh_c,3] (indexed by chain position, func.cpp:2047)
    LH_conn: np.ndarray        # int32[ntpl]
    LH_q: np.ndarray
    LH_vdw_hh: np.ndarray
    LH_vdw_hc: np.ndarray
    LH_vdw_hl: np.ndarray
    LH_vdw_ht: np.ndarray
    LH_kstr: np.ndarray
    LH_kben: np.ndarray
    LH_streq: np.ndarray
    LH_betaeq: np.ndarray
    LH_radius: np.ndarray
    nc_lh_flag: np.ndarray     # int32[n_c]
    lh_rad: np.ndarray         # [n_LH]
    phys: Physics
    groups_per_core: int = 10  # func.cpp:2368 literal
    hi_all: int = 0            # 1: D over every bead (func_cuda.cu:3524), 0: cores+linkers

    @property
    def n(self) -> int:
        return int(len(self.types))

    @property
    def n_c(self) -> int:
        return int(np.sum(self.types == 1))

    @property
    def Nq(self) -> int:
        return int(len(self.core_q))

    @property
    def n_t(self) -> int:
        return int(len(self.tail_fix_tpl))

    @property
    def n_tail(self) -> int:
        return int(len(self.t_q))

    @property
    def n_LH(self) -> int:
        return int(len(self.lh_rad))

    @property
    def n_hi(self) -> int:
        return self.n + self.n_tail + self.n_LH if self.hi_all else self.n

    def r_var(self) -> np.ndarray:
        return rot_var(self.types, self.phys)

    def rad_all(self) -> np.ndarray:
        return np.concatenate([self.rad, self.t_rad, self.lh_rad])

    def validate(self) -> None:
        t = self.types
        if t[0] != 1:
            raise ValueError("the first bead must be a core (func.cpp:1062 reads type[i-1])")
        if self.n < 3:
            raise ValueError("need at least three beads")
        if self.n_lh_n != 0 and self.n_LH:
            raise ValueError("N-terminal LH beads are not supported by the reference "
                             "bending rule (func.cpp:2067, i % n_lh_c)")


@dataclasses.dataclass
class State:
    """Dynamic state of a fiber: positions and body frames."""

    r: np.ndarray
    a: np.ndarray
    b: np.ndarray
    c: np.ndarray
    r_t: np.ndarray
    r_lh: np.ndarray

    def copy(self) -> "State":
        return State(*(np.array(x, copy=True) for x in
                       (self.r, self.a, self.b, self.c, self.r_t, self.r_lh)))


def body_to_world(r, a, b, c, p):
    """r + a*p0 + b*p1 + c*p2 for frames r,a,b,c [3] and points p [...,3]."""
    return r + np.multiply.outer(p[..., 0], a) + np.multiply.outer(p[..., 1], b) \
        + np.multiply.outer(p[..., 2], c)


def build_tail_positions(types, tail_pos, r, a, b, c):
    """Tail bead positions of build_tail (func.cpp:1689-1711), all cores tailed."""
    out = []
    for i in np.nonzero(types == 1)[0]:
        sl = slice(3 * i, 3 * i + 3)
        out.append(body_to_world(r[sl], a[sl], b[sl], c[sl], tail_pos))
    return np.concatenate(out).reshape(-1) if out else np.zeros(0)


def build_lh_positions(types, nc_lh_flag, n_lh_n, LH_n_pos, LH_c_pos, r, a, b, c):
    """Free LH bead positions of build_LH (func.cpp:2038-2057)."""
    n_lh_c = len(LH_c_pos)
    nlh = n_lh_n + n_lh_c
    out = []
    for k1, i in enumerate(np.nonzero(types == 1)[0]):
        if nc_lh_flag[k1] != 1:
            continue
        sl = slice(3 * i, 3 * i + 3)
        for j in range(nlh):
            src = LH_n_pos[j] if j < n_lh_n else LH_c_pos[j]
            out.append(body_to_world(r[sl], a[sl], b[sl], c[sl], src))
    return np.asarray(out).reshape(-1) if out else np.zeros(0)


def expand_tails(types, tail_q, tail_grp, tail_fix, tail_rad, bond_c, bond_v, angle_c,
                 angle_v):
    """Per-bead tail constants of build_tail (func.cpp:1678-1705)."""
    n_t = len(tail_q)
    n_grp = 1 + int(np.sum(np.asarray(tail_grp[1:]) != np.asarray(tail_grp[:-1])))
    n_c = int(np.sum(types == 1))
    rep = lambda x: np.tile(np.asarray(x, dtype=np.float64), n_c)
    t_grp = np.concatenate([np.asarray(tail_grp, np.int32) + k * n_grp for k in range(n_c)])
    return dict(t_q=rep(tail_q), t_rad=rep(tail_rad), t_fix=np.tile(np.asarray(tail_fix, np.int32), n_c),
                t_grp=t_grp.astype(np.int32), h_t=rep(bond_c), lo_t=rep(bond_v), g_t=rep(angle_c),
                beta_o_t=rep(angle_v), n_grp=n_grp, n_t=n_t)


# ---------------------------------------------------------------------------
# readers of the reference input formats (readfile.cpp)
# ---------------------------------------------------------------------------

def read_setup(path):
    """Read_const, readfile.cpp:14-38: token 2 of every line after the first."""
    vals = []
    with open(path) as f:
        for i, line in enumerate(f):
            if i == 0:
                continue
            tok = line.split()
            if len(tok) >= 2:
                vals.append(float(tok[1]))
    keys = ["T", "number_of_steps", "frequency_of_sampling", "frequency_RP", "sig_dd",
            "sig_dc", "sig_cc", "sig_tt", "sig_tl", "sig_tc", "Mg", "Cs", "restart",
            "unbind", "Pbind"]
    return dict(zip(keys, vals))


def read_data_mod(path):
    """Read_initial_struc, readfile.cpp:40-134: 4 lines per bead."""
    rows = [l.split() for l in open(path) if l.strip()]
    n = len(rows) // 4
    r = np.array([[float(x) for x in rows[4 * i][:3]] for i in range(n)]).reshape(-1)
    types = np.array([int(rows[4 * i][3]) for i in range(n)], np.int32)
    a = np.array([[float(x) for x in rows[4 * i + 1][:3]] for i in range(n)]).reshape(-1)
    b = np.array([[float(x) for x in rows[4 * i + 2][:3]] for i in range(n)]).reshape(-1)
    c = np.array([[float(x) for x in rows[4 * i + 3][:3]] for i in range(n)]).reshape(-1)
    rad = np.where(types == 1, 5.0, 1.5)
    return types, r, a, b, c, rad


def read_core(path):
    """Read_core, readfile.cpp:137-174: positions in Angstrom (/10), then charges."""
    pos, q = [], []
    for l in open(path):
        t = l.split()
        if len(t) == 3:
            pos.append([float(x) / 10 for x in t])
        elif t:
            q.append(float(t[0]))
    return np.array(pos), np.array(q)


def read_tail(path):
    """Read_tail, readfile.cpp:176-235 with its unit scalings."""
    cols = [l.split() for l in open(path) if l.strip()]
    f = lambda k: np.array([float(c[k]) for c in cols])
    return dict(
        tail_grp=np.array([int(c[1]) for c in cols], np.int32),
        tail_fix=np.array([int(c[2]) for c in cols], np.int32),
        tail_q=f(3) * 1.12, tail_rad=f(4) / 10,
        tail_pos=np.stack([f(6) / 10, f(7) / 10, f(8) / 10], 1),
        bond_v=f(9) / 10, bond_c=f(10) * 100 * 4.142e-3 / 0.5962,
        angle_v=f(11) * PI / 180, angle_c=f(12) * 4.142e-3 / 0.5962)


def read_lh(path):
    """Read_LH, readfile.cpp:238-337 with its unit scalings (positions in nm)."""
    sect = None
    rows = {"N": [], "G": [], "C": []}
    for l in open(path):
        t = l.split()
        if not t:
            continue
        if len(t) > 1 and t[1] in ("N-term", "G-head", "C-term"):
            sect = t[1][0]
            continue
        if t[0] == "#":
            continue
        rows[sect].append(t)
    allrows = rows["N"] + rows["G"] + rows["C"]
    f = lambda k: np.array([float(t[k]) for t in allrows])
    pos = lambda key: np.array([[float(x) for x in t[:3]] for t in rows[key]]).reshape(-1, 3)
    return dict(
        n_lh_n=len(rows["N"]), n_lh_g=len(rows["G"]), n_lh_c=len(rows["C"]),
        LH_n_pos=pos("N"), LH_g_pos=pos("G"), LH_c_pos=pos("C"),
        LH_q=f(3) * 0.66, LH_vdw_hh=f(4), LH_vdw_hc=f(5), LH_vdw_hl=f(6), LH_vdw_ht=f(7),
        LH_kstr=f(8) * 4.142e-3 / 0.5962, LH_kben=f(9) * 4.142e-3 / 0.5962,
        LH_conn=np.array([int(t[12]) for t in allrows], np.int32),
        LH_streq=f(13), LH_betaeq=f(14) * PI / 180, LH_radius=np.full(len(allrows), 0.5))


def read_lh_in(path, n_c):
    """Read_LH_core, readfile.cpp:563-584."""
    vals = [int(float(l.split()[0])) for l in open(path) if l.strip()]
    return np.array(vals[:n_c], np.int32)


# ---------------------------------------------------------------------------
# writers of the reference input formats (inverse of the readers above)
# ---------------------------------------------------------------------------

_E_UNIT = 4.142e-3 / 0.5962  # energy scaling of the tail and LH force constants


def _g(x):
    return "%.17g" % float(x)


def write_deck(directory, fib: "Fiber", st: "State", number_of_steps=1000,
               frequency_of_sampling=100, frequency_RP=20, restart=0, unbind=0.0, Pbind=0.0):
    """Writes a fiber and its state as an input deck the reference driver and
    load_reference_deck read (setup.txt, data_mod, core_data.reg.150mM,
    tail_data.mod.200mM, LH_N0G6C22.in, LH.in), undoing the readers' unit
    scalings.  Numbers carry 17 significant digits.  The reference derives
    q_l from Cs (main.cpp:78-84) and uses 10 tail groups per core
    (func.cpp:2368); a fiber built with other values needs them passed
    again when the deck is read."""
    os.makedirs(directory, exist_ok=True)
    j = lambda name: os.path.join(directory, name)
    ph = fib.phys
    rows = [("Temperature", ph.T), ("number_of_steps", number_of_steps),
            ("frequency_of_sampling", frequency_of_sampling), ("frequency_RP", frequency_RP),
            ("evd_DNA_DNA", ph.sig_dd), ("evd_DNA_Core", ph.sig_dc), ("evd_Core_Core", ph.sig_cc),
            ("evd_Tail_Tail", ph.sig_tt), ("evd_Tail_DNA", ph.sig_tl), ("evd_Tail_Core", ph.sig_tc),
            ("Mg", ph.Mg), ("Cs", ph.Cs), ("restart", restart), ("unbind", unbind),
            ("Pbind", Pbind)]
    with open(j("setup.txt"), "w") as f:
        f.write("$ configuration file for the Brownian dynamics simulation\n")
        for k, v in rows:
            f.write(f"{k} {v if isinstance(v, int) else _g(v)}\n")
    r, a, b, c = (np.asarray(x, np.float64).reshape(-1, 3) for x in (st.r, st.a, st.b, st.c))
    with open(j("data_mod"), "w") as f:
        for i in range(fib.n):
            f.write(f"{_g(r[i, 0])} {_g(r[i, 1])} {_g(r[i, 2])} {int(fib.types[i])}\n")
            for v in (a[i], b[i], c[i]):
                f.write(f"{_g(v[0])} {_g(v[1])} {_g(v[2])}\n")
    with open(j("core_data.reg.150mM"), "w") as f:
        for p in fib.core_pos.reshape(-1, 3):
            f.write(" ".join(_g(10 * x) for x in p) + "\n")
        for q in fib.core_q:
            f.write(_g(q) + "\n")
    nt = fib.n_t
    with open(j("tail_data.mod.200mM"), "w") as f:
        for k in range(nt):
            p = fib.tail_pos.reshape(-1, 3)[k]
            vals = [fib.t_q[k] / 1.12, fib.t_rad[k] * 10, 0.0, 10 * p[0], 10 * p[1], 10 * p[2],
                    fib.lo_t[k] * 10, fib.h_t[k] / (100 * _E_UNIT), fib.beta_o_t[k] * 180 / PI,
                    fib.g_t[k] / _E_UNIT]
            f.write(f"T{k + 1} {int(fib.t_grp[k])} {int(fib.tail_fix_tpl[k])} "
                    + " ".join(_g(v) for v in vals) + "\n")
    nn, ng, ncc = fib.n_lh_n, fib.n_lh_g, fib.n_lh_c
    with open(j("LH_N0G6C22.in"), "w") as f:
        f.write("# bead position (x,y,z) charge VdW_hh VdW_hc VdW_hl VdW_ht kstr kben ktor Seq "
                "Conn rstreq betaeq\n")
        pos = {"N": fib.LH_n_pos.reshape(-1, 3), "G": fib.LH_g_pos.reshape(-1, 3),
               "C": fib.LH_c_pos.reshape(-1, 3)}
        k = 0
        for sect, cnt in (("N-term", nn), ("G-head", ng), ("C-term", ncc)):
            f.write(f"# {sect}\n")
            for m in range(cnt):
                p = pos[sect[0]][m]
                vals = [p[0], p[1], p[2], fib.LH_q[k] / 0.66, fib.LH_vdw_hh[k], fib.LH_vdw_hc[k],
                        fib.LH_vdw_hl[k], fib.LH_vdw_ht[k], fib.LH_kstr[k] / _E_UNIT,
                        fib.LH_kben[k] / _E_UNIT, 0.0]
                f.write(" ".join(_g(v) for v in vals)
                        + f" {k + 1} {int(fib.LH_conn[k])} {_g(fib.LH_streq[k])} "
                        + _g(fib.LH_betaeq[k] * 180 / PI) + "\n")
                k += 1
    with open(j("LH.in"), "w") as f:
        f.write("".join(f"{int(x)}\n" for x in fib.nc_lh_flag))


def load_reference_deck(directory, hi_all=0, q_l=None):
    """Assemble a Fiber and State from a directory in the reference formats,
    as main.cpp:40-283 does (restart=0, every core tailed)."""
    st = read_setup(os.path.join(directory, "setup.txt"))
    ph = Physics.derive(st["T"], st["Cs"], int(st["Mg"]),
                        sigmas=(st["sig_dd"], st["sig_dc"], st["sig_cc"], st["sig_tt"],
                                st["sig_tl"], st["sig_tc"]), q_l=q_l)
    types, r, a, b, c, rad = read_data_mod(os.path.join(directory, "data_mod"))
    core_pos, core_q = read_core(os.path.join(directory, "core_data.reg.150mM"))
    tl = read_tail(os.path.join(directory, "tail_data.mod.200mM"))
    lh = read_lh(os.path.join(directory, "LH_N0G6C22.in"))
    n_c = int(np.sum(types == 1))
    flags = read_lh_in(os.path.join(directory, "LH.in"), n_c)
    return assemble(types, rad, r, a, b, c, core_pos, core_q, tl, lh, flags, ph, hi_all=hi_all)


def assemble(types, rad, r, a, b, c, core_pos, core_q, tl, lh, flags, ph, hi_all=0,
             groups_per_core=10):
    types = np.asarray(types, np.int32)
    et = expand_tails(types, tl["tail_q"], tl["tail_grp"], tl["tail_fix"], tl["tail_rad"],
                      tl["bond_c"], tl["bond_v"], tl["angle_c"], tl["angle_v"])
    r_t = build_tail_positions(types, tl["tail_pos"], r, a, b, c)
    n_lh_n, n_lh_c = lh["n_lh_n"], lh["n_lh_c"]
    r_lh = build_lh_positions(types, flags, n_lh_n, lh["LH_n_pos"], lh["LH_c_pos"], r, a, b, c)
    nlh = n_lh_n + n_lh_c
    n_LH = int(np.sum(flags == 1)) * nlh
    fib = Fiber(
        types=types, rad=np.asarray(rad, np.float64), phi_o=phi_o_of(types),
        core_pos=np.asarray(core_pos, np.float64).reshape(-1, 3),
        core_q=np.asarray(core_q, np.float64),
        tail_pos=np.asarray(tl["tail_pos"], np.float64).reshape(-1, 3),
        tail_fix_tpl=np.asarray(tl["tail_fix"], np.int32),
        t_q=et["t_q"], t_rad=et["t_rad"], h_t=et["h_t"], g_t=et["g_t"], lo_t=et["lo_t"],
        beta_o_t=et["beta_o_t"], t_grp=et["t_grp"], t_fix=et["t_fix"],
        n_lh_n=n_lh_n, n_lh_g=lh["n_lh_g"], n_lh_c=n_lh_c,
        LH_g_pos=lh["LH_g_pos"].reshape(-1, 3), LH_n_pos=lh["LH_n_pos"].reshape(-1, 3),
        LH_c_pos=lh["LH_c_pos"].reshape(-1, 3),
        LH_conn=np.asarray(lh["LH_conn"], np.int32),
        LH_q=lh["LH_q"], LH_vdw_hh=lh["LH_vdw_hh"], LH_vdw_hc=lh["LH_vdw_hc"],
        LH_vdw_hl=lh["LH_vdw_hl"], LH_vdw_ht=lh["LH_vdw_ht"], LH_kstr=lh["LH_kstr"],
        LH_kben=lh["LH_kben"], LH_streq=lh["LH_streq"], LH_betaeq=lh["LH_betaeq"],
        LH_radius=lh["LH_radius"], nc_lh_flag=np.asarray(flags, np.int32),
        lh_rad=np.tile(lh["LH_radius"][:nlh], int(np.sum(flags == 1))) if n_LH else np.zeros(0),
        phys=ph, groups_per_core=groups_per_core, hi_all=hi_all)
    state = State(r=np.asarray(r, np.float64).copy(), a=np.asarray(a, np.float64).copy(),
                  b=np.asarray(b, np.float64).copy(), c=np.asarray(c, np.float64).copy(),
                  r_t=np.asarray(r_t, np.float64), r_lh=np.asarray(r_lh, np.float64))
    return fib, state
