"""Differentiable BD trajectory on the GPU (PyTorch-ROCm + HIP).

Gradients of a loss at the end of a trajectory -- the fiber energy, or the
squared end-to-end distance |r_{n-1} - r_0|^2 -- with respect to the model
parameters: the core charge template (300 charges shared by every core),
the tail-bead charges, and the stretch, bend and twist stiffnesses h, g, s
(h also sets the tail anchor springs, as in func.cpp:1971-2022).

The step is the Ermak-McCammon predictor-corrector of main_cuda
(func_cuda.cu:3434-3680) with the CPU semantics of func.cpp, written as
batched torch operations:
  * frames and Euler angles, update_Euler_Angle (func.cpp:433-685);
  * linker-DNA mechanics with the explicit torques (func.cpp:951-1289);
  * core/linker electrostatics (func.cpp:1291-1605), whose core-core charge
    sums -- the hot pair term -- run in the hand-written HIP kernels of
    bd_adjoint.hip, forward and adjoint, behind CoreCorePairs, a
    torch.autograd.Function;
  * histone tails (func.cpp:1763-2026);
  * the RPY tensor over cores and linker beads (func.cpp:728-804), its
    Cholesky factor in FP64, rd = s2dt L p, rotation (func.cpp:1607-1670)
    and translation (func_cuda.cu:650-700).
Forces are evaluated in `force_dtype` (float32 by default; float64 for
checks), the state, the frames, D and its factor in float64.  Cut-off
tests (7, 4, 8, 25 nm) and the +-0.9 force clamp are kept: the map is
piecewise smooth and its gradient is the gradient of the piece the
trajectory is on.  Fibers with linker histones are refused (the
differentiable form covers the 50-nucleosome fiber without them), and so
is the all-bead hydrodynamic set (tails move with their self-mobility).

Trajectories are checkpointed every `checkpoint_every` steps
(torch.utils.checkpoint): only the segment boundaries are kept and each
segment is recomputed during the backward pass.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import engine as E
from . import system as S

THETA = 1.2 * S.PI / 2
SI, CO = math.sin(THETA), math.cos(THETA)
ZETA_CORE = 8 * S.PI * S.ETA * 125.0
ZETA_LINKER = 4 * S.PI * S.ETA * S.R_H * S.R_H * S.LO

DERIVED = ("alpha", "beta", "gamma", "length", "a_dna", "b_dna", "c_dna", "alpha_p", "beta_p",
           "gamma_p", "beta_t")
STATE = ("r", "a", "b", "c", "r_t") + DERIVED


def _dot(u, v):
    return (u * v).sum(-1)


class _Acos1(torch.autograd.Function):
    """acos of the argument clamped to [-1, 1] (the reference's clamp before
    acos, func.cpp:522-545), with the derivative of the clamped map: zero
    where the clamp acts (also at exactly +-1, where acos' is unbounded),
    so branches that torch.where discards carry no NaN into the gradient."""

    @staticmethod
    def forward(ctx, x):
        ctx.save_for_backward(x)
        return torch.acos(torch.clamp(x, -1.0, 1.0))

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        inside = (x > -1.0) & (x < 1.0)
        one_m = torch.where(inside, 1.0 - x * x, torch.ones_like(x))
        return torch.where(inside, -g / torch.sqrt(one_m), torch.zeros_like(g))


def _acos1(x):
    return _Acos1.apply(x)


def _clamp1(x):
    return torch.clamp(x, -1.0, 1.0)


def _cross_ref(u, v):
    """cross product with the component order of utilities.cpp:139-148"""
    return torch.stack([u[..., 1] * v[..., 2] - u[..., 2] * v[..., 1],
                        u[..., 2] * v[..., 0] - u[..., 0] * v[..., 2],
                        u[..., 0] * v[..., 1] - u[..., 1] * v[..., 0]], -1)


def _body(f, a, b, c):
    """components of world vector f along the frame (a, b, c)"""
    return torch.stack([_dot(a, f), _dot(b, f), _dot(c, f)], -1)


class CoreCorePairs(torch.autograd.Function):
    """Core-core charge forces and energies (func.cpp:1489-1566) on the HIP
    kernels of bd_adjoint.hip: forward bd_cc_forward, backward the adjoint
    kernel bd_cc_backward (gradients with respect to the charge positions
    and the charge template)."""

    @staticmethod
    def forward(ctx, Q, q, off, nbr, consts):
        lib = E.load_library()
        Q = Q.contiguous()
        q = q.contiguous()
        n_c, nq = Q.shape[0], Q.shape[1]
        f = torch.empty_like(Q)
        e = torch.empty(n_c, dtype=Q.dtype, device=Q.device)
        fp64 = 1 if Q.dtype == torch.float64 else 0
        stream = C.c_void_p(torch.cuda.current_stream(Q.device).cuda_stream)
        rc = lib.bd_cc_forward(fp64, n_c, nq, Q.data_ptr(), q.data_ptr(), off.data_ptr(),
                               nbr.data_ptr(), *consts, f.data_ptr(), e.data_ptr(), stream)
        if rc:
            raise E.BDError(f"bd_cc_forward: {lib.bd_last_error().decode()}")
        ctx.save_for_backward(Q, q, off, nbr)
        ctx.consts = consts
        return f, e

    @staticmethod
    def backward(ctx, gf, ge):
        lib = E.load_library()
        Q, q, off, nbr = ctx.saved_tensors
        n_c, nq = Q.shape[0], Q.shape[1]
        gf = (torch.zeros_like(Q) if gf is None else gf).contiguous()
        ge = (torch.zeros(n_c, dtype=Q.dtype, device=Q.device) if ge is None else ge).contiguous()
        gQ = torch.empty_like(Q)
        gqs = torch.empty(n_c, nq, dtype=Q.dtype, device=Q.device)
        fp64 = 1 if Q.dtype == torch.float64 else 0
        stream = C.c_void_p(torch.cuda.current_stream(Q.device).cuda_stream)
        rc = lib.bd_cc_backward(fp64, n_c, nq, Q.data_ptr(), q.data_ptr(), off.data_ptr(),
                                nbr.data_ptr(), *ctx.consts, gf.data_ptr(), ge.data_ptr(),
                                gQ.data_ptr(), gqs.data_ptr(), stream)
        if rc:
            raise E.BDError(f"bd_cc_backward: {lib.bd_last_error().decode()}")
        return gQ, gqs.sum(0), None, None, None


class DiffBD:
    """The differentiable step of one fiber on one GPU."""

    def __init__(self, fib: S.Fiber, device="cuda", force_dtype=torch.float32):
        if not torch.cuda.is_available():
            raise RuntimeError("the differentiable form runs on the GPU (HIP kernels)")
        E.load_library()
        if fib.n_LH or int(np.sum(fib.nc_lh_flag)):
            raise ValueError("the differentiable form covers fibers without linker histones")
        if fib.hi_all:
            raise ValueError("the differentiable form uses the core + linker hydrodynamic set")
        self.fib = fib
        self.dev = torch.device(device)
        self.fdt = force_dtype
        ph = fib.phys
        self.ph = ph
        d64 = dict(dtype=torch.float64, device=self.dev)
        ti = dict(dtype=torch.long, device=self.dev)
        types = np.asarray(fib.types)
        n = fib.n
        self.n, self.n_c, self.nq = n, fib.n_c, fib.Nq
        self.n_tail, self.n_t = fib.n_tail, fib.n_t
        self.types = torch.tensor(types, **ti)
        self.is_core = self.types == 1
        self.core_idx = torch.nonzero(self.is_core).flatten()
        # core ordinal of the last core at or before each bead (func.cpp:587-591)
        self.ic = torch.tensor(np.cumsum(types == 1) - 1, **ti)
        self.rad = torch.tensor(fib.rad, **d64)
        self.phi_o = torch.tensor(fib.phi_o, **d64)
        self.core_pos = torch.tensor(np.asarray(fib.core_pos).reshape(-1, 3), **d64)
        self.tail_pos = torch.tensor(np.asarray(fib.tail_pos).reshape(-1, 3), **d64)
        self.t_rad = torch.tensor(fib.t_rad, **d64)
        self.h_t = torch.tensor(fib.h_t, **d64)
        self.g_t = torch.tensor(fib.g_t, **d64)
        self.lo_t = torch.tensor(fib.lo_t, **d64)
        self.beta_o_t = torch.tensor(fib.beta_o_t, **d64)
        tg = np.asarray(fib.t_grp)
        nt = self.n_tail
        self.t_str = torch.tensor(np.nonzero(tg[:-1] == tg[1:])[0] if nt > 1 else np.zeros(0), **ti)
        if nt > 2:
            ok = (tg[:-2] == tg[1:-1]) & (tg[:-2] == tg[2:])
            self.t_bend = torch.tensor(np.nonzero(ok)[0], **ti)
        else:
            self.t_bend = torch.zeros(0, **ti)
        # tail-tail pairs that the reference evaluates: i < j, other group
        # or more than two apart (func.cpp:1863-1891)
        ii, jj = np.triu_indices(nt, 1)
        keep = (tg[ii] != tg[jj]) | (jj > ii + 2)
        self.tt_i = torch.tensor(ii[keep], **ti)
        self.tt_j = torch.tensor(jj[keep], **ti)
        self.linkers = torch.nonzero(self.types == 0).flatten()
        # anchored tail beads: (core bead, tail bead, template row)
        fix = np.asarray(fib.tail_fix_tpl)
        cores = np.nonzero(types == 1)[0]
        anc = [(cores[k], k * fib.n_t + j, j) for k in range(len(cores)) for j in range(fib.n_t)
               if fix[j] == 1]
        anc = np.array(anc, dtype=np.int64).reshape(-1, 3)
        self.anc_core = torch.tensor(anc[:, 0], **ti)
        self.anc_tail = torch.tensor(anc[:, 1], **ti)
        self.anc_tpl = torch.tensor(anc[:, 2], **ti)
        # pair candidates of the core/linker term: i < j with the `ch`
        # filter's type rules (func.cpp:1316-1323); the 25 nm rule is applied
        # per evaluation
        ok_i = np.zeros(n, bool)
        ok_j = np.zeros(n, bool)
        for k in range(n):
            nxt_ok = k + 1 < n and types[k + 1] == 0
            prv_ok = k == 0 or types[k - 1] == 0
            ok_i[k] = nxt_ok and prv_ok
            ok_j[k] = (k == n - 1 or types[k + 1] == 0) and (k > 0 and types[k - 1] == 0)
        pi, pj = np.triu_indices(n, 1)
        sel = ok_i[pi] & ok_j[pj]
        pi, pj = pi[sel], pj[sel]
        tpi, tpj = types[pi], types[pj]
        dd = (tpi == 0) & (tpj == 0) & (np.abs(pi - pj) > 1)
        self.dd_i, self.dd_j = torch.tensor(pi[dd], **ti), torch.tensor(pj[dd], **ti)
        dc = (tpi + tpj) == 1
        lk = np.where(tpi[dc] == 0, pi[dc], pj[dc])
        co = np.where(tpi[dc] == 0, pj[dc], pi[dc])
        self.dc_l, self.dc_c = torch.tensor(lk, **ti), torch.tensor(co, **ti)
        self.dc_sgn = torch.tensor(np.where(tpi[dc] == 0, 1.0, -1.0), **d64)
        self.dc_far = torch.tensor(np.abs(pi[dc] - pj[dc]) > 1, dtype=torch.bool, device=self.dev)
        cc = (tpi == 1) & (tpj == 1)
        self.cc_i, self.cc_j = pi[cc], pj[cc]
        self.core_ord = np.cumsum(types == 1) - 1
        # integrator constants
        self.zeta = torch.where(self.is_core, torch.tensor(ZETA_CORE, **d64),
                                torch.tensor(ZETA_LINKER, **d64))
        self.rot_mask = torch.ones(n, 3, **d64)
        self.rot_mask[self.types == 0, 1:] = 0.0

    # ------------------------------------------------------------------
    def params(self, requires_grad=True):
        """The differentiable parameters as leaf tensors (float64)."""
        ph = self.ph
        d64 = dict(dtype=torch.float64, device=self.dev)
        p = dict(q_core=torch.tensor(self.fib.core_q, **d64), t_q=torch.tensor(self.fib.t_q, **d64),
                 h=torch.tensor(ph.h, **d64), g=torch.tensor(ph.g, **d64),
                 s=torch.tensor(ph.s, **d64))
        for v in p.values():
            v.requires_grad_(requires_grad)
        return p

    def initial_state(self, st: S.State):
        """Positions and frames with the derived arrays, as the reference
        prepares them (main.cpp:164, 223)."""
        d64 = dict(dtype=torch.float64, device=self.dev)
        x = dict(r=torch.tensor(st.r, **d64).reshape(-1, 3), a=torch.tensor(st.a, **d64).reshape(-1, 3),
                 b=torch.tensor(st.b, **d64).reshape(-1, 3), c=torch.tensor(st.c, **d64).reshape(-1, 3),
                 r_t=torch.tensor(st.r_t, **d64).reshape(-1, 3))
        return self.rebuild(x)

    # ------------------------------------------------------------------
    # frames, update_Euler_Angle (func.cpp:433-685) and update_tail_beta
    def rebuild(self, x):
        r, a, b, c = x["r"], x["a"], x["b"], x["c"]
        n = self.n
        lk = ~self.is_core[:-1]
        nxt_core = self.is_core[1:]
        ri, rf = r[:-1], r[1:]
        ai, bi, ci = a[:-1], b[:-1], c[:-1]
        af, bf, cf = a[1:], b[1:], c[1:]
        # pass 1 (448-505): linker frames and segment lengths
        bm_n = -SI * af + CO * bf
        fw_l = torch.where(nxt_core[:, None], rf - S.RO * bm_n + S.D1 * cf - ri, rf - ri)
        fw_c = rf - (ri - S.RO * bi - S.D1 * ci)
        fw = torch.where(lk[:, None], fw_l, fw_c)
        length = torch.sqrt(_dot(fw, fw))
        a_new = fw / length[:, None]
        da = a_new - ai
        b1 = bi - _dot(bi, da)[:, None] * ai
        b2 = b1 - _dot(b1, a_new)[:, None] * a_new
        b_new = b2 / torch.sqrt(_dot(b2, b2))[:, None]
        c_new = _cross_ref(a_new, b_new)
        A = torch.cat([torch.where(lk[:, None], a_new, ai), a[-1:]])
        B = torch.cat([torch.where(lk[:, None], b_new, bi), b[-1:]])
        Cc = torch.cat([torch.where(lk[:, None], c_new, ci), c[-1:]])
        # pass 2 (507-681)
        Ai, Bi, Ci = A[:-1], B[:-1], Cc[:-1]
        Af, Bf, Cf = A[1:], B[1:], Cc[1:]
        # linker beads
        am = torch.where(nxt_core[:, None], CO * Af + SI * Bf, Af)
        bm = torch.where(nxt_core[:, None], -SI * Af + CO * Bf, Bf)
        al_l, be_l, ga_l = self._angles(Ai, Bi, Ci, am, bm, Cf, strict=True, valid=lk)
        # cores: the exit frame (a_dna, b_dna, c_dna)
        ad = (rf - (ri - S.RO * Bi - S.D1 * Ci)) / length[:, None]
        cb = _clamp1(_dot(Ai, ad))
        beta_p = _acos1(_dot(Ai, ad))
        sb = torch.sin(beta_p)
        big = beta_p >= 1e-10
        sb_safe = torch.where(big, sb, torch.ones_like(sb))
        bmv = (ad - cb[:, None] * Ai) / sb_safe[:, None]
        ca = _clamp1(_dot(bmv, Bi))
        Ac = _acos1(_dot(bmv, Bi))
        alpha_p = torch.where(_dot(ad, Ci) >= 0, Ac, -Ac)
        alpha_p = torch.where(big, alpha_p, torch.zeros_like(alpha_p))
        gamma_p = -alpha_p
        sa, sg, cg = torch.sin(alpha_p), torch.sin(gamma_p), torch.cos(gamma_p)
        R21, R22, R23 = -cg * sb, cg * cb * ca - sg * sa, cg * cb * sa + sg * ca
        R31, R32, R33 = sg * sb, -sg * cb * ca - cg * sa, -sg * cb * sa + cg * ca
        bd = R21[:, None] * Ai + R22[:, None] * Bi + R23[:, None] * Ci
        cd = R31[:, None] * Ai + R32[:, None] * Bi + R33[:, None] * Ci
        bd = torch.where(big[:, None], bd, Bi)
        cd = torch.where(big[:, None], cd, Ci)
        al_c, be_c, ga_c = self._angles(ad, bd, cd, Af, Bf, Cf, strict=False, valid=~lk)
        isl = lk
        alpha = torch.cat([torch.where(isl, al_l, al_c), r.new_zeros(1)])
        beta = torch.cat([torch.where(isl, be_l, be_c), r.new_zeros(1)])
        gamma = torch.cat([torch.where(isl, ga_l, ga_c), r.new_zeros(1)])
        length_full = torch.cat([length, r.new_zeros(1)])
        # exit frames by core ordinal (a core as the last bead keeps zeros)
        cidx = torch.nonzero(~lk).flatten()
        zc = r.new_zeros(self.n_c, 3)
        zc1 = r.new_zeros(self.n_c)
        ordv = self.ic[cidx]
        a_dna = zc.index_copy(0, ordv, ad[cidx])
        b_dna = zc.index_copy(0, ordv, bd[cidx])
        c_dna = zc.index_copy(0, ordv, cd[cidx])
        al_p = zc1.index_copy(0, ordv, alpha_p[cidx])
        be_p = zc1.index_copy(0, ordv, beta_p[cidx])
        ga_p = zc1.index_copy(0, ordv, gamma_p[cidx])
        out = dict(r=r, a=A, b=B, c=Cc, r_t=x["r_t"], alpha=alpha, beta=beta, gamma=gamma,
                   length=length_full, a_dna=a_dna, b_dna=b_dna, c_dna=c_dna, alpha_p=al_p,
                   beta_p=be_p, gamma_p=ga_p, beta_t=self._tail_beta(x["r_t"]))
        del n
        return out

    @staticmethod
    def _angles(a1, b1, c1, a2, b2, c2, strict, valid):
        """alpha, beta, gamma of frame 2 in frame 1 (func.cpp:520-548).
        Lanes outside `valid` (the other bead kind's branch, which the
        caller discards) get two equal unit frames, so that no division by
        1 + a1.a2 = 0 or unbounded acos' reaches the gradient from them."""
        e = torch.eye(3, dtype=a1.dtype, device=a1.device)
        vm = valid[:, None]
        a1, b1, c1 = (torch.where(vm, u, e[k].expand_as(u)) for k, u in enumerate((a1, b1, c1)))
        a2, b2, c2 = (torch.where(vm, u, e[k].expand_as(u)) for k, u in enumerate((a2, b2, c2)))
        ada = _clamp1(_dot(a1, a2))
        beta = _acos1(_dot(a1, a2))
        sb = torch.sin(beta)
        big = (beta > 1e-10) if strict else (beta >= 1e-10)
        sb_safe = torch.where(big, sb, torch.ones_like(sb))
        f1 = torch.where(big, _dot(a2, b1) / sb_safe, _dot(b2, b1))
        Ac = _acos1(f1)
        alpha = torch.where(_dot(a2, c1) >= 0, Ac, -Ac)
        apg = _acos1((_dot(b1, b2) + _dot(c1, c2)) / (1.0 + ada))
        f2 = (_dot(c1, b2) - _dot(b1, c2)) / (1.0 + ada)
        gamma = torch.where(f2 >= 0.0, apg - alpha, -apg - alpha)
        return alpha, beta, gamma

    def _tail_beta(self, rt):
        out = rt.new_zeros(self.n_tail)
        if len(self.t_bend) == 0:
            return out
        i = self.t_bend
        p1, p2, p3 = rt[i], rt[i + 1], rt[i + 2]
        u, v = p1 - p2, p3 - p2
        cosv = _dot(u, v) / (torch.sqrt(_dot(u, u)) * torch.sqrt(_dot(v, v)))
        return out.index_copy(0, i, _acos1(cosv))

    # ------------------------------------------------------------------
    # force pass (func_cuda.cu:3556-3613 order: mechanics, electrostatics, tails)
    def forces(self, x, prm):
        """F, T (n x 3), tF (n_tail x 3) in float64 and the energy, computed
        in self.fdt."""
        dt = self.fdt
        y = {k: v.to(dt) for k, v in x.items()}
        P = {k: v.to(dt) for k, v in prm.items()}
        F, T, Em = self._mech(y, P)
        F2, T2, Ee = self._elec(y, P)
        tF, F3, T3, Et = self._tail(y, P)
        F = F + F2 + F3
        T = T + T2 + T3
        return F.double(), T.double(), tF.double(), (Em + Ee + Et).double()

    def _seg_torque(self, f, pf, pc, a, b, c):
        """torque_due_to_force (func.cpp:374-394) in the body frame"""
        comp = _body(pf - pc, a, b, c)
        return _cross_ref(comp, _body(f, a, b, c))

    def _mech(self, y, P):
        r, a, b, c = y["r"], y["a"], y["b"], y["c"]
        alpha, beta, gamma, length = y["alpha"], y["beta"], y["gamma"], y["length"]
        a_dna, b_dna, c_dna = y["a_dna"], y["b_dna"], y["c_dna"]
        alpha_p, beta_p, gamma_p = y["alpha_p"], y["beta_p"], y["gamma_p"]
        h, g, s = P["h"], P["g"], P["s"]
        phi = self.phi_o.to(r.dtype)
        n = self.n
        core = self.is_core
        ci_ = core[:-1]
        cn_ = core[1:]
        ic = self.ic[:-1]
        # segment ends: entry point of bead i+1, exit point of bead i (func.cpp:687-723)
        af, bf, cf, rf0 = a[1:], b[1:], c[1:], r[1:]
        bm_e = -SI * af + CO * bf
        rf = torch.where(cn_[:, None], rf0 - S.RO * bm_e + S.D1 * cf, rf0)
        rs = torch.where(ci_[:, None], r[:-1] - (S.RO * b[:-1] + S.D1 * c[:-1]), r[:-1])
        seg_f = r.new_zeros(n - 1, 3)
        Etot = r.new_zeros(())
        # stretching (978-1018)
        dv = rs - rf
        d = torch.sqrt(_dot(dv, dv))
        Etot = Etot + (h * (d - S.LO) ** 2 / 2).sum()
        seg_f = seg_f + (-h * (d - S.LO) / d)[:, None] * dv
        # bending (1022-1105)
        im1 = torch.clamp(torch.arange(n - 1, device=r.device) - 1, min=0)
        prev_core = core[im1]
        af_b = torch.where(ci_[:, None], af, torch.where(cn_[:, None], CO * af + SI * bf, af))
        ab_l = torch.where(prev_core[:, None], a_dna[ic], a[im1])
        ab = torch.where(ci_[:, None], a[:-1], ab_l)
        ao = torch.where(ci_[:, None], a_dna[ic], a[:-1])
        beta_b = torch.where(ci_, beta_p[ic], beta[im1])
        bi, ln = beta[:-1], length[:-1]
        Etot = Etot + (g * (bi - beta_b) ** 2 / 2).sum()
        big1 = bi >= 1e-10
        g1 = torch.where(big1, bi / (torch.where(big1, torch.sin(bi), torch.ones_like(bi)) * ln), 1.0 / ln)
        big2 = beta_b >= 1e-10
        g2 = torch.where(big2, beta_b / (torch.where(big2, torch.sin(beta_b), torch.ones_like(bi)) * ln),
                         1.0 / ln)
        Ai = g1[:, None] * (af_b - torch.cos(bi)[:, None] * ao)
        Bi = g2[:, None] * (ab - torch.cos(beta_b)[:, None] * ao)
        seg_f = seg_f - g * (Ai + Bi)
        # twisting (1110-1179)
        al_b = torch.where(ci_, alpha[:-1], alpha[im1])
        be_b = torch.where(ci_, beta_p[ic], beta[im1])
        ga_b = torch.where(ci_, gamma[:-1], gamma[im1])
        ph_b = torch.where(ci_, phi[:-1], phi[im1])
        ga_n = torch.where(ci_, gamma_p[ic], gamma[im1])
        bo = torch.where(ci_[:, None], b_dna[ic], b[:-1])
        cc = torch.where(ci_[:, None], c_dna[ic], c[:-1])
        tw = alpha[:-1] + gamma[:-1] - phi[:-1]
        Etot = Etot + (s * tw * tw / 2).sum()
        g1 = tw * torch.tan(0.5 * beta[:-1]) / ln
        ca, sa = torch.cos(alpha[:-1]), torch.sin(alpha[:-1])
        g2 = (al_b + ga_b - ph_b) * torch.tan(0.5 * be_b) / ln
        cg, sg = torch.cos(ga_n), torch.sin(ga_n)
        Chi = g1[:, None] * (ca[:, None] * cc - sa[:, None] * bo)
        Zhi = g2[:, None] * (cg[:, None] * cc + sg[:, None] * bo)
        seg_f = seg_f + s * (Chi + Zhi)
        # every segment force acts as +f on bead i (at rs) and -f on bead i+1 (at rf)
        F = r.new_zeros(n, 3).index_add(0, torch.arange(n - 1, device=r.device), seg_f)
        F = F.index_add(0, torch.arange(1, n, device=r.device), -seg_f)
        t1 = self._seg_torque(seg_f, rs, r[:-1], a[:-1], b[:-1], c[:-1])
        t2 = self._seg_torque(-seg_f, rf, r[1:], af, bf, cf)
        T = r.new_zeros(n, 3).index_add(0, torch.arange(n - 1, device=r.device), t1)
        T = T.index_add(0, torch.arange(1, n, device=r.device), t2)
        # explicit torques (1185-1277): linker beads assign, cores add
        ai_, bi_, ci3 = a[:-1], b[:-1], c[:-1]
        has_prev = torch.arange(n - 1, device=r.device) > 0
        twp = alpha[im1] + gamma[im1] - phi[im1]
        lin_T0 = s * (alpha[:-1] + gamma[:-1] + phi[:-1] - alpha[im1] - gamma[im1] - phi[im1])
        ad, bd, cd = a_dna[ic], b_dna[ic], c_dna[ic]
        Tc = T[:-1] + (s * tw)[:, None] * torch.stack([_dot(ad, ai_), _dot(ad, bi_), _dot(ad, ci3)], -1)
        prevv = torch.stack([torch.full_like(tw, CO), torch.full_like(tw, SI), torch.zeros_like(tw)], -1)
        Tc = Tc + torch.where(has_prev[:, None], (-s * twp)[:, None] * prevv, torch.zeros_like(Tc))
        bp = beta_p[ic]
        sbp = torch.sin(bp)
        adb, adc = _dot(ad, bi_), _dot(ad, ci3)
        Tc = Tc + torch.stack([torch.zeros_like(tw), -g * bp * adc / sbp, g * bp * adb / sbp], -1)
        abv = a[im1]
        xa, xb, xc = _dot(abv, ai_), _dot(abv, bi_), _dot(abv, ci3)
        bmv = beta[im1]
        sbm = torch.where(has_prev, torch.sin(bmv), torch.ones_like(bmv))
        eb = torch.stack([g * bmv * (SI * xc) / sbm, -g * bmv * (CO * xc) / sbm,
                          g * bmv * (CO * xb - SI * xa) / sbm], -1)
        Tc = Tc + torch.where(has_prev[:, None], eb, torch.zeros_like(eb))
        s1, c1 = torch.sin(alpha_p[ic]), torch.cos(alpha_p[ic])
        mag = s * tw * torch.tan(0.5 * bp)
        et = torch.stack([s1 * _dot(cd, ai_) + c1 * _dot(bd, ai_), s1 * _dot(cd, bi_) + c1 * _dot(bd, bi_),
                          s1 * _dot(cd, ci3) + c1 * _dot(bd, ci3)], -1)
        Tc = Tc - mag[:, None] * et
        s1p, c1p = torch.sin(gamma[im1]), torch.cos(gamma[im1])
        magp = s * twp * torch.tan(0.5 * beta[im1])
        ep = torch.stack([magp * c1p * SI, -magp * c1p * CO, magp * s1p], -1)
        Tc = Tc - torch.where(has_prev[:, None], ep, torch.zeros_like(ep))
        Tl = torch.stack([lin_T0, torch.zeros_like(tw), torch.zeros_like(tw)], -1)
        Tnew = torch.where(ci_[:, None], Tc, Tl)
        Tlast = torch.stack([-s * (alpha[n - 2] + gamma[n - 2] - phi[n - 2]), r.new_zeros(()), r.new_zeros(())])
        T = torch.cat([Tnew, Tlast[None]])
        return F, T, Etot

    def _charge_sites(self, r, a, b, c):
        """world positions of every core's charges, n_c x Nq x 3"""
        k = self.core_idx
        p = self.core_pos.to(r.dtype)
        return (r[k][:, None, :] + a[k][:, None, :] * p[None, :, 0:1] + b[k][:, None, :] * p[None, :, 1:2]
                + c[k][:, None, :] * p[None, :, 2:3])

    def _elec(self, y, P):
        r, a, b, c = y["r"], y["a"], y["b"], y["c"]
        ph = self.ph
        n = self.n
        ke, kex, ql = ph.k_e, ph.k_ex, ph.q_l
        F = r.new_zeros(n, 3)
        T = r.new_zeros(n, 3)
        Etot = r.new_zeros(())
        # DNA-DNA (func.cpp:1328-1354), 25 nm rule
        i, j = self.dd_i, self.dd_j
        z = r[i] - r[j]
        d = torch.sqrt(_dot(z, z))
        with torch.no_grad():
            m = d <= 25.0
        i, j, z, d = i[m], j[m], z[m], d[m]
        kll = ph.debyell
        g1 = ke * ql * ql * torch.exp(-kll * d) * (kll * d + 1.0) / (d * d)
        Etot = Etot + (ke * ql * ql / d * torch.exp(-kll * d)).sum()
        s1 = ph.sig_dd
        ev = d <= 8
        g1 = g1 + torch.where(ev, kex * ((12.0 / s1) * (s1 / d) ** 13 - (6.0 / s1) * (s1 / d) ** 7),
                              torch.zeros_like(d))
        Etot = Etot + torch.where(ev, kex * ((s1 / d) ** 12 - (s1 / d) ** 6), torch.zeros_like(d)).sum()
        fz = (g1 / d)[:, None] * z
        F = F.index_add(0, i, fz).index_add(0, j, -fz)
        # DNA-core per charge (func.cpp:1356-1487), 25 nm rule on bead centres
        Q = self._charge_sites(r, a, b, c)
        lk, co, sgn, far = self.dc_l, self.dc_c, self.dc_sgn.to(r.dtype), self.dc_far
        with torch.no_grad():
            dz = r[lk] - r[co]
            m = torch.sqrt(_dot(dz, dz)) <= 25.0
        lk, co, sgn, far = lk[m], co[m], sgn[m], far[m]
        oc = self.ic[co]
        qv = P["q_core"]
        zq = sgn[:, None, None] * (r[lk][:, None, :] - Q[oc])        # charge -> linker (DNA i) or back
        dq = torch.sqrt(_dot(zq, zq))
        zh = zq / dq[..., None]
        kap = ph.debye
        qq = ke * ql * qv[None, :]
        g_dh = torch.where(far[:, None], qq * torch.exp(-kap * dq) * (kap * dq + 1.0) / (dq * dq),
                           torch.zeros_like(dq))
        e_dh = torch.where(far[:, None], qq / dq * torch.exp(-kap * dq), torch.zeros_like(dq))
        s1 = ph.sig_dc
        evm = (dq <= 8.0) & (qv[None, :] > 0)
        g_ev = torch.where(evm, kex * ((12.0 / s1) * (s1 / dq) ** 13 - (6.0 / s1) * (s1 / dq) ** 7),
                           torch.zeros_like(dq))
        e_ev = torch.where(evm, kex * ((s1 / dq) ** 12 - (s1 / dq) ** 6), torch.zeros_like(dq))
        Etot = Etot + e_dh.sum() + e_ev.sum()
        fi = (g_dh + g_ev)[..., None] * zh                              # force on bead i
        # bead i gets +fi, bead j -fi; with the linker as i (sgn > 0) the
        # linker receives +fi and the core -fi, with the core as i the core
        # receives +fi and the linker -fi
        f_link = sgn[:, None, None] * fi
        F = F.index_add(0, lk, f_link.sum(1))
        F = F.index_add(0, co, -f_link.sum(1))
        fc = -f_link                                                     # force on each charge
        ac, bc, cc_ = a[co], b[co], c[co]
        fb = torch.stack([_dot(ac[:, None, :], fc), _dot(bc[:, None, :], fc), _dot(cc_[:, None, :], fc)], -1)
        p = self.core_pos.to(r.dtype)[None]
        tq = _cross_ref(p.expand_as(fb), fb).sum(1)
        T = T.index_add(0, co, tq)
        # core-core charge pairs (func.cpp:1489-1566): the HIP term
        with torch.no_grad():
            ci_np, cj_np = self.cc_i, self.cc_j
            if len(ci_np):
                ri_ = r[torch.as_tensor(ci_np, device=r.device)]
                rj_ = r[torch.as_tensor(cj_np, device=r.device)]
                dcc = torch.sqrt(_dot(ri_ - rj_, ri_ - rj_)).cpu().numpy()
                keep = dcc <= 25.0
                oi, oj = self.core_ord[ci_np[keep]], self.core_ord[cj_np[keep]]
            else:
                oi = oj = np.zeros(0, np.int64)
            nb = [[] for _ in range(self.n_c)]
            for u, v in zip(oi, oj):
                nb[u].append(v)
                nb[v].append(u)
            off = np.zeros(self.n_c + 1, np.int32)
            off[1:] = np.cumsum([len(x) for x in nb])
            flat = np.array([v for x in nb for v in sorted(x)], np.int32)
            off_t = torch.tensor(off, dtype=torch.int32, device=r.device)
            nbr_t = torch.tensor(flat if len(flat) else np.zeros(1, np.int32), dtype=torch.int32,
                                 device=r.device)
        consts = (ke, ph.debye, kex, ph.sig_cc)
        fs, es = CoreCorePairs.apply(Q, qv, off_t, nbr_t, consts)
        Etot = Etot + es.sum()
        k = self.core_idx
        F = F.index_add(0, k, fs.sum(1))
        fb = torch.stack([_dot(a[k][:, None, :], fs), _dot(b[k][:, None, :], fs), _dot(c[k][:, None, :], fs)], -1)
        p = self.core_pos.to(r.dtype)[None]
        T = T.index_add(0, k, _cross_ref(p.expand_as(fb), fb).sum(1))
        return F, T, Etot

    @staticmethod
    def _clamp09(v):
        return torch.clamp(v, -0.9, 0.9)

    def _tail(self, y, P):
        r, a, b, c, rt = y["r"], y["a"], y["b"], y["c"], y["r_t"]
        ph = self.ph
        dt = r.dtype
        n, nt = self.n, self.n_tail
        F = r.new_zeros(n, 3)
        T = r.new_zeros(n, 3)
        tF = r.new_zeros(nt, 3)
        Etot = r.new_zeros(())
        if nt == 0:
            return tF, F, T, Etot
        tq = P["t_q"]
        ke, kap, kex = ph.k_e, ph.debye, ph.k_ex
        # stretching (1811-1827)
        i = self.t_str
        dv = rt[i] - rt[i + 1]
        d = torch.sqrt(_dot(dv, dv))
        ht, lot = self.h_t.to(dt)[i], self.lo_t.to(dt)[i]
        Etot = Etot + (ht * (d - lot) ** 2 / 2).sum()
        f1 = (-ht * (d - lot) / d)[:, None] * dv
        tF = tF.index_add(0, i, f1).index_add(0, i + 1, -f1)
        # bending (1832-1858), Force_Bending (func.cpp:153-196)
        i = self.t_bend
        if len(i):
            p1, p2, p3 = rt[i], rt[i + 1], rt[i + 2]
            gt, b0 = self.g_t.to(dt)[i], self.beta_o_t.to(dt)[i]
            bt = y["beta_t"][i]
            Etot = Etot + (gt * (bt - b0) ** 2 / 2).sum()
            f = -gt * (bt - b0)
            d1 = torch.sqrt(_dot(p1 - p2, p1 - p2))
            d2 = torch.sqrt(_dot(p2 - p3, p2 - p3))
            ji, jk, kj = p2 - p1, p2 - p3, p3 - p2
            n0 = _cross_ref(ji, jk)
            rI = _cross_ref(ji, n0)
            rK = _cross_ref(kj, n0)
            u1 = (f / d1)[:, None] * rI / torch.sqrt(_dot(rI, rI))[:, None]
            u3 = (f / d2)[:, None] * rK / torch.sqrt(_dot(rK, rK))[:, None]
            # forces are subtracted (func.cpp:1852-1856): -f1, -f2, -f3
            tF = tF.index_add(0, i, -u1).index_add(0, i + 1, u1 + u3).index_add(0, i + 2, -u3)
        # tail-tail Debye-Hueckel within 7 nm (1863-1891), clamped per component
        ii, jj = self.tt_i, self.tt_j
        with torch.no_grad():
            zz = rt[ii] - rt[jj]
            m = _dot(zz, zz) < 49.0
        ii, jj = ii[m], jj[m]
        z = rt[ii] - rt[jj]
        d = torch.sqrt(_dot(z, z))
        qq = tq[ii] * tq[jj]
        ex = torch.exp(-kap * d)
        f = ke * qq * (kap * d + 1.0) / (d * d) * ex
        Etot = Etot + (ke * qq / d * ex).sum()
        fv = self._clamp09((f / d)[:, None] * z)
        tF = tF.index_add(0, ii, fv).index_add(0, jj, -fv)
        # tail-linker DH (7 nm) + EV (4 nm) (1898-1917)
        lk = self.linkers
        with torch.no_grad():
            zz = rt[:, None, :] - r[lk][None, :, :]
            m = _dot(zz, zz) < 49.0
            ti_, lj = torch.nonzero(m, as_tuple=True)
        lj = lk[lj]
        z = rt[ti_] - r[lj]
        d = torch.sqrt(_dot(z, z))
        qq = tq[ti_] * ph.q_l
        ex = torch.exp(-kap * d)
        f = ke * qq * (kap * d + 1.0) / (d * d) * ex
        Etot = Etot + (ke * qq / d * ex).sum()
        s1 = ph.sig_tl
        ev = d < 4.0
        fv = torch.where(ev, -kex * ((6 * s1 ** 6) / d ** 7 - (12 * s1 ** 12) / d ** 13), torch.zeros_like(d))
        Etot = Etot + torch.where(ev, kex * ((s1 / d) ** 12 - (s1 / d) ** 6), torch.zeros_like(d)).sum()
        fw = self._clamp09(((f + fv) / d)[:, None] * z)
        tF = tF.index_add(0, ti_, fw)
        F = F.index_add(0, lj, -fw)
        # anchor springs (1971-2022)
        ci, tb, tp = self.anc_core, self.anc_tail, self.anc_tpl
        pos = self.tail_pos.to(dt)[tp]
        ro = r[ci] + a[ci] * pos[:, 0:1] + b[ci] * pos[:, 1:2] + c[ci] * pos[:, 2:3]
        Sv = rt[tb] - ro
        h = P["h"]
        Etot = Etot + (h * _dot(Sv, Sv) / 2).sum()
        hs = h * Sv
        tF = tF.index_add(0, tb, -hs)
        F = F.index_add(0, ci, hs)
        T = T.index_add(0, ci, _cross_ref(pos, _body(hs, a[ci], b[ci], c[ci])))
        return tF, F, T, Etot

    # ------------------------------------------------------------------
    # tensor, factor, integrator
    def tensor(self, r):
        """RPY tensor over cores and linker beads (func.cpp:728-804), 3n x 3n."""
        ph = self.ph
        n = self.n
        rad = self.rad
        z = r[:, None, :] - r[None, :, :]
        eye = torch.eye(n, dtype=torch.bool, device=r.device)
        ssq = _dot(z, z) + eye.double()          # no zero distance on the diagonal
        s = torch.sqrt(ssq)
        ri, rj = rad[:, None], rad[None, :]
        zz = z[..., :, None] * z[..., None, :]
        I3 = torch.eye(3, dtype=r.dtype, device=r.device)
        f = (ri * ri + rj * rj) / ssq
        far = (ph.a2 / s)[..., None, None] * ((1.0 + f / 3)[..., None, None] * I3
                                              + (1.0 - f)[..., None, None] * zz / ssq[..., None, None])
        se = ((ri ** 3 + rj ** 3) / 2.0) ** (1.0 / 3)
        near = (ph.a1 / se)[..., None, None] * ((1.0 - 9.0 * s / (32 * se))[..., None, None] * I3
                                                + 3.0 * zz / (32.0 * se * s)[..., None, None])
        blk = torch.where((s >= ri + rj)[..., None, None], far, near)
        diag = (ph.a1 / rad)[:, None, None] * I3
        blk = torch.where(eye[..., None, None], diag[:, None].expand_as(blk), blk)
        return blk.permute(0, 2, 1, 3).reshape(3 * n, 3 * n)

    @staticmethod
    def _rotate(a, b, c, w):
        """rotate (func.cpp:1607-1670): Rodrigues update of the body frame"""
        wa, wb, wc = w[:, 0], w[:, 1], w[:, 2]
        z2 = wa * wa + wb * wb + wc * wc
        zero = z2 <= 0.0
        z2s = torch.where(zero, torch.ones_like(z2), z2)
        z = torch.sqrt(z2s)
        czt, szt = torch.cos(z), torch.sin(z)
        om = 1.0 - czt
        ga = torch.stack([((wb * wb + wc * wc) * czt + wa * wa) / z2s, wa * wb * om / z2s + wc * szt / z,
                          wa * wc * om / z2s - wb * szt / z], -1)
        gb = torch.stack([wa * wb * om / z2s - wc * szt / z, ((wa * wa + wc * wc) * czt + wb * wb) / z2s,
                          wb * wc * om / z2s + wa * szt / z], -1)
        gc = torch.stack([wa * wc * om / z2s + wb * szt / z, wb * wc * om / z2s - wa * szt / z,
                          ((wa * wa + wb * wb) * czt + wc * wc) / z2s], -1)
        an = ga[:, 0:1] * a + ga[:, 1:2] * b + ga[:, 2:3] * c
        bn = gb[:, 0:1] * a + gb[:, 1:2] * b + gb[:, 2:3] * c
        cn = gc[:, 0:1] * a + gc[:, 1:2] * b + gc[:, 2:3] * c
        zc = zero[:, None]
        return torch.where(zc, a, an), torch.where(zc, b, bn), torch.where(zc, c, cn)

    def _move(self, x0, D, rd, rr, F, T, tF):
        """rotation_cal + translation_cal from x0 with forces (F, T, tF)"""
        ph = self.ph
        n3 = 3 * self.n
        w = (ph.time_step * T / self.zeta[:, None] + rr) * self.rot_mask
        a, b, c = self._rotate(x0["a"], x0["b"], x0["c"], w)
        r = x0["r"] + rd[:n3].reshape(-1, 3) + ph.dell * (D @ F.reshape(-1)).reshape(-1, 3)
        rt = x0["r_t"] + rd[n3:].reshape(-1, 3) + ph.dell * (ph.a1 / self.t_rad)[:, None] * tF
        return dict(r=r, a=a, b=b, c=c, r_t=rt)

    def step(self, x0, rr, p, prm):
        """One predictor-corrector step from the full state x0 with host
        noise rr (n x 3) and p (3(n + n_tail)); returns the full state after
        it and the energy of force pass 1."""
        ph = self.ph
        n3 = 3 * self.n
        D = self.tensor(x0["r"])
        L = torch.linalg.cholesky(D)
        rd = torch.cat([ph.s2dt * (L @ p[:n3]),
                        ph.s2dt * torch.sqrt(ph.a1 / self.t_rad).repeat_interleave(3) * p[n3:]])
        F0, T0, tF0, E0 = self.forces(x0, prm)
        pred = self.rebuild(self._move(x0, D, rd, rr, F0, T0, tF0))
        F1, T1, tF1, _ = self.forces(pred, prm)
        x1 = self._move(x0, D, rd, rr, 0.5 * (F0 + F1), 0.5 * (T0 + T1), 0.5 * (tF0 + tF1))
        return self.rebuild(x1), E0

    def trajectory(self, x0, rr, p, prm, checkpoint_every=20):
        """len(rr) steps from x0 (rr: K x n x 3, p: K x 3(n + n_tail)),
        recomputed segment by segment in the backward pass."""
        keys = STATE
        state = tuple(x0[k] for k in keys)
        pv = tuple(prm[k] for k in ("q_core", "t_q", "h", "g", "s"))

        def seg(s0, s1):
            def run(*args):
                st = dict(zip(keys, args[:len(keys)]))
                pr = dict(zip(("q_core", "t_q", "h", "g", "s"), args[len(keys):]))
                for k in range(s0, s1):
                    st, _ = self.step(st, rr[k], p[k], pr)
                return tuple(st[k] for k in keys)
            return run

        K = len(rr)
        for s0 in range(0, K, checkpoint_every):
            s1 = min(K, s0 + checkpoint_every)
            state = checkpoint(seg(s0, s1), *state, *pv, use_reentrant=False)
        return dict(zip(keys, state))

    # ------------------------------------------------------------------
    def energy(self, x, prm):
        """fiber energy of the force pass at x (func_cuda.cu:3613-3615)"""
        return self.forces(x, prm)[3]

    @staticmethod
    def end_to_end2(x):
        """squared distance between the first and the last bead"""
        dv = x["r"][-1] - x["r"][0]
        return (dv * dv).sum()

    def noise(self, rr, p):
        """host noise arrays (K x 3n, K x 3(n + n_tail)) as device tensors"""
        d64 = dict(dtype=torch.float64, device=self.dev)
        rr = torch.tensor(np.asarray(rr), **d64).reshape(len(rr), -1, 3)
        p = torch.tensor(np.asarray(p), **d64).reshape(len(p), -1)
        return rr, p

    def gradients(self, st: S.State, rr, p, loss="end_to_end", checkpoint_every=20):
        """(loss value, {parameter: gradient}) for a trajectory of len(rr)
        steps from the positions and frames st."""
        prm = self.params(True)
        x0 = self.initial_state(st)
        rrt, pt = self.noise(rr, p)
        xT = self.trajectory(x0, rrt, pt, prm, checkpoint_every)
        L = self.end_to_end2(xT) if loss == "end_to_end" else self.energy(xT, prm)
        keys = ("q_core", "t_q", "h", "g", "s")
        grads = torch.autograd.grad(L, [prm[k] for k in keys])
        return float(L.detach()), {k: gr.detach().cpu().numpy() for k, gr in zip(keys, grads)}
