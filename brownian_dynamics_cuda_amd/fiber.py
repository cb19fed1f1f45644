"""Seeded synthetic fibers for the benchmark configurations.

A zigzag chromatin fiber of ``n_c`` nucleosome cores joined by straight
linker DNA of ``linker_beads`` beads.  Bead, tail and linker-histone
conventions follow the reference input deck (``readfile.cpp``): the first
bead is a core, core charges live in the core body frame (the DNA wraps the
body c axis at radius 4.8 nm, ``constants.h``), tails are anchored by their
last bead (``tail_fix``), and the linker histone has an N-terminal,
globular and C-terminal part (``LH_N0G6C22.in`` layout).

Recipe (all charges are final values, after the reader's scalings):
  * 300 core charges uniform on the side of a cylinder r = 5.5 nm,
    h = 5.5 nm about body c, charge U[-1, 1] e;
  * 5 tails x 10 beads per core, +1 e per bead, zigzag chains;
  * linker beads at about 3 nm spacing, -1 e per bead;
  * optionally 6 globular + 22 C-terminal LH beads per core, q U[0, 1] e.
"""
from __future__ import annotations

import math

import numpy as np

from . import system as S


def _orthonormal(b_dir, a_hint):
    b = b_dir / np.linalg.norm(b_dir)
    a = a_hint - np.dot(a_hint, b) * b
    a /= np.linalg.norm(a)
    c = np.cross(a, b)
    return a, b, c


def _exit_point(r, b, c):
    # second_coord, func.cpp:711-723
    return r - (S.RO * b + S.D1 * c)


def _entry_point(r, a, b, c):
    # first_coord, func.cpp:687-709
    bm = -math.sin(S.THETA) * a + math.cos(S.THETA) * b
    return r - S.RO * bm + S.D1 * c


def chain_geometry(types, r, a, b, c):
    """Conventions of a bead chain as the reference's frame update reads them
    (update_Euler_Angle, func.cpp:448-505; first_coord/second_coord,
    func.cpp:687-723): frame orthonormality and handedness, the bond vector
    from each linker bead to the next bead (a core's entry point), the
    distance from each core's exit point to the bead after it, and the
    alignment of each linker bead's a axis with its bond."""
    types = np.asarray(types)
    r, a, b, c = (np.asarray(x, np.float64).reshape(-1, 3) for x in (r, a, b, c))
    dots = np.concatenate([np.einsum("ij,ij->i", a, b), np.einsum("ij,ij->i", a, c),
                           np.einsum("ij,ij->i", b, c)])
    norms = np.concatenate([np.linalg.norm(x, axis=1) for x in (a, b, c)])
    hand = np.einsum("ij,ij->i", np.cross(a, b), c)
    bonds, exits, align = [], [], []
    for i in range(len(types) - 1):
        if types[i] == 0:
            if types[i + 1] == 0:
                f = r[i + 1] - r[i]
            else:
                f = _entry_point(r[i + 1], a[i + 1], b[i + 1], c[i + 1]) - r[i]
            bonds.append(np.linalg.norm(f))
            align.append(np.dot(f, a[i]) / np.linalg.norm(f))
        else:
            exits.append(np.linalg.norm(r[i + 1] - _exit_point(r[i], b[i], c[i])))
    cores = np.nonzero(types == 1)[0]
    return dict(max_dot=float(np.max(np.abs(dots))), max_norm_err=float(np.max(np.abs(norms - 1))),
                min_handedness=float(np.min(hand)), bonds=np.array(bonds), exits=np.array(exits),
                align=np.array(align), linkers_per_core=np.diff(cores))


def zigzag_geometry(n_c, linker_beads=6, radius=13.2, twist=0.85 * math.pi, rise=2.1,
                    rng=None):
    """Core centres on a two-start zigzag; straight linkers between the exit
    point of core i and the entry point of core i+1."""
    if rng is None:
        rng = np.random.default_rng(0)
    cen = np.array([[radius * math.cos(i * twist), radius * math.sin(i * twist), i * rise]
                    for i in range(n_c)])
    frames = []
    for i in range(n_c):
        nxt = cen[i + 1] - cen[i] if i + 1 < n_c else cen[i] - cen[i - 1]
        prv = cen[i - 1] - cen[i] if i > 0 else -(cen[i + 1] - cen[i])
        # exit (-b) faces the next core, entry faces the previous one
        a, b, c = _orthonormal(-nxt, prv + 0.3 * rng.standard_normal(3))
        frames.append((a, b, c))
    types, rs, as_, bs, cs = [], [], [], [], []
    for i in range(n_c):
        a, b, c = frames[i]
        types.append(1); rs.append(cen[i]); as_.append(a); bs.append(b); cs.append(c)
        if i + 1 == n_c:
            break
        p0 = _exit_point(cen[i], b, c)
        a1, b1, c1 = frames[i + 1]
        p1 = _entry_point(cen[i + 1], a1, b1, c1)
        seg = (p1 - p0) / (linker_beads + 1)
        la = seg / np.linalg.norm(seg)
        for j in range(1, linker_beads + 1):
            lb = np.cross(la, c) if abs(np.dot(la, c)) < 0.9 else np.cross(la, a)
            lb /= np.linalg.norm(lb)
            lc = np.cross(la, lb)
            types.append(0); rs.append(p0 + j * seg); as_.append(la); bs.append(lb); cs.append(lc)
    rad = np.where(np.array(types) == 1, 5.0, 1.5)
    flat = lambda v: np.asarray(v, np.float64).reshape(-1)
    return np.array(types, np.int32), flat(rs), flat(as_), flat(bs), flat(cs), rad


def core_template(rng, nq=300, radius=5.5, height=5.5):
    phi = rng.uniform(0.0, 2 * math.pi, nq)
    z = rng.uniform(-height / 2, height / 2, nq)
    pos = np.stack([radius * np.cos(phi), radius * np.sin(phi), z], 1)
    q = rng.uniform(-1.0, 1.0, nq)
    return pos, q


def _zigzag_chain(start, direction, side, nbeads, bond, angle_deg):
    """Planar zigzag of nbeads beads starting at `start` going along
    `direction` with bond angle angle_deg (no three beads collinear)."""
    d = direction / np.linalg.norm(direction)
    s = side - np.dot(side, d) * d
    s /= np.linalg.norm(s)
    half = math.radians(angle_deg) / 2
    step_along = bond * math.sin(half)
    step_side = bond * math.cos(half)
    pts = [np.asarray(start, np.float64)]
    for j in range(1, nbeads):
        pts.append(pts[-1] + step_along * d + (step_side if j % 2 else -step_side) * s)
    return np.array(pts)


def tail_template(n_tails=5, beads=10, bond=1.4, angle=115.0, charge=1.0):
    """Tail beads listed free end first, anchor (tail_fix=1) last, as in
    tail_data.mod.200mM."""
    pos, grp, fix = [], [], []
    for t in range(n_tails):
        phi = 2 * math.pi * (t + 0.5) / n_tails
        out = np.array([math.cos(phi), math.sin(phi), 0.0])
        z0 = 1.2 if t % 2 else -1.2
        anchor = 6.2 * out + np.array([0, 0, z0])
        chain = _zigzag_chain(anchor, out + np.array([0, 0, 0.35 * np.sign(z0)]),
                              np.array([0, 0, 1.0]), beads, bond, angle)
        for j in range(beads - 1, -1, -1):
            pos.append(chain[j]); grp.append(t + 1); fix.append(1 if j == 0 else 0)
    nt = len(pos)
    kb = 4.142e-3 / 0.5962
    bond_v = np.full(nt, bond)
    bond_c = np.full(nt, 0.08 * 100 * kb)
    angle_v = np.full(nt, math.radians(angle))
    angle_c = np.full(nt, 1.1 * kb)
    grp = np.array(grp, np.int32)
    for i in range(nt):      # last bead of a group: no bond / angle (file convention)
        if i == nt - 1 or grp[i + 1] != grp[i]:
            bond_v[i] = bond_c[i] = angle_v[i] = angle_c[i] = 0.0
    return dict(tail_pos=np.array(pos), tail_grp=grp, tail_fix=np.array(fix, np.int32),
                tail_q=np.full(nt, charge), tail_rad=np.full(nt, 0.6), bond_v=bond_v,
                bond_c=bond_c, angle_v=angle_v, angle_c=angle_c, n_grp=n_tails)


def lh_template(rng, n_g=6, n_c=22, bond=1.5, angle=110.0):
    """Globular head near the dyad plus a zigzag C-terminal chain; charges
    U[0, 1] e (final)."""
    kb = 4.142e-3 / 0.5962
    dyad = np.array([math.cos(-0.65), math.sin(-0.65), 0.0])
    gcen = 6.9 * dyad
    gpos = gcen + 0.6 * rng.standard_normal((n_g, 3))
    cpos = _zigzag_chain(gcen + 1.2 * dyad, dyad, np.array([0, 0, 1.0]), n_c, bond, angle)
    ntpl = n_g + n_c
    conn = np.zeros(ntpl, np.int32)
    conn[n_g - 1] = 1               # globular bead anchoring the C-term chain
    conn[n_g:n_g + n_c - 1] = 1     # C-term chain bonds
    kstr = np.zeros(ntpl)
    kstr[n_g - 1] = 10 * kb
    kstr[n_g:] = 10 * kb
    kben = np.zeros(ntpl)
    kben[n_g:] = 1 * kb
    streq = np.zeros(ntpl)
    streq[n_g:] = bond
    betaeq = np.zeros(ntpl)
    betaeq[n_g + 1:] = math.radians(angle)
    return dict(n_lh_n=0, n_lh_g=n_g, n_lh_c=n_c, LH_n_pos=np.zeros((0, 3)), LH_g_pos=gpos,
                LH_c_pos=cpos, LH_q=rng.uniform(0.0, 1.0, ntpl),
                LH_vdw_hh=np.full(ntpl, 1.6), LH_vdw_hc=np.full(ntpl, 1.75),
                LH_vdw_hl=np.full(ntpl, 2.3), LH_vdw_ht=np.full(ntpl, 1.7),
                LH_kstr=kstr, LH_kben=kben, LH_conn=conn, LH_streq=streq, LH_betaeq=betaeq,
                LH_radius=np.full(ntpl, 0.5))


def make_fiber(n_c, seed, with_lh=False, with_tails=True, linker_beads=6, hi_all=0,
               Cs=0.15, T=293.0, q_linker=-1.0, nq=300):
    """Seeded fiber of n_c cores.  Returns (Fiber, State)."""
    rng = np.random.default_rng(seed)
    types, r, a, b, c, rad = zigzag_geometry(n_c, linker_beads, rng=rng)
    core_pos, core_q = core_template(rng, nq)
    if with_tails:
        tl = tail_template()
        gpc = tl.pop("n_grp")
    else:
        tl = dict(tail_pos=np.zeros((0, 3)), tail_grp=np.zeros(0, np.int32),
                  tail_fix=np.zeros(0, np.int32), tail_q=np.zeros(0), tail_rad=np.zeros(0),
                  bond_v=np.zeros(0), bond_c=np.zeros(0), angle_v=np.zeros(0),
                  angle_c=np.zeros(0))
        gpc = 10
    lh = lh_template(rng)
    flags = np.full(n_c, 1 if with_lh else 0, np.int32)
    ph = S.Physics.derive(T, Cs, 0, q_l=q_linker)
    return S.assemble(types, rad, r, a, b, c, core_pos, core_q, tl, lh, flags, ph,
                      hi_all=hi_all, groups_per_core=gpc)


# The five benchmark configurations (n_c, LH, steps).
CONFIGS = {
    1: dict(n_c=12, with_lh=False, steps=1000),
    2: dict(n_c=100, with_lh=True, steps=10000),
    3: dict(n_c=500, with_lh=False, steps=None),
    4: dict(n_c=50, with_lh=False, steps=200),
    5: dict(n_c=1000, with_lh=True, steps=1000),
}


def make_config(idx, hi_all=0):
    cfg = CONFIGS[idx]
    return make_fiber(cfg["n_c"], seed=20261202 + idx, with_lh=cfg["with_lh"], hi_all=hi_all)
