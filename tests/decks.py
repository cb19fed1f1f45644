"""Small seeded decks shared by the tests (sizes the oracle finishes in
seconds).  Everything is generated; nothing reads the reference tree."""
import numpy as np

from brownian_dynamics_cuda_amd import fiber as Fb


def deck(kind, hi_all=0):
    """kind: 'tiny' (3 cores, tails, LH), 'small' (6 cores, tails, LH),
    'cfg1' (12 cores, tails, no LH, the 12-nucleosome recipe)."""
    if kind == "tiny":
        return Fb.make_fiber(3, seed=11, with_lh=True, hi_all=hi_all)
    if kind == "small":
        return Fb.make_fiber(6, seed=12, with_lh=True, hi_all=hi_all)
    if kind == "cfg1":
        return Fb.make_config(1, hi_all=hi_all)
    raise KeyError(kind)


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    if b.size == 0:
        return 0.0
    den = np.max(np.abs(b))
    if den == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)


GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def load_golden(name):
    """(Fiber, State, outputs) from tests/golden/<name>.npz."""
    import os
    from brownian_dynamics_cuda_amd import system as S
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    ph = S.Physics(**{k[5:]: (int(z[k]) if k == "phys_Mg" else float(z[k]))
                      for k in z.files if k.startswith("phys_")})
    fa = {k[4:]: z[k] for k in z.files if k.startswith("fib_")}
    fib = S.Fiber(
        types=fa["types"].astype(np.int32), rad=fa["rad"], phi_o=fa["phi_o"],
        core_pos=fa["core_pos"], core_q=fa["core_q"], tail_pos=fa["tail_pos"],
        tail_fix_tpl=fa["tail_fix_tpl"].astype(np.int32), t_q=fa["t_q"], t_rad=fa["t_rad"],
        h_t=fa["h_t"], g_t=fa["g_t"], lo_t=fa["lo_t"], beta_o_t=fa["beta_o_t"],
        t_grp=fa["t_grp"].astype(np.int32), t_fix=fa["t_fix"].astype(np.int32),
        n_lh_n=int(fa["n_lh_n"]), n_lh_g=int(fa["n_lh_g"]), n_lh_c=int(fa["n_lh_c"]),
        LH_g_pos=fa["LH_g_pos"], LH_n_pos=fa["LH_n_pos"], LH_c_pos=fa["LH_c_pos"],
        LH_conn=fa["LH_conn"].astype(np.int32), LH_q=fa["LH_q"], LH_vdw_hh=fa["LH_vdw_hh"],
        LH_vdw_hc=fa["LH_vdw_hc"], LH_vdw_hl=fa["LH_vdw_hl"], LH_vdw_ht=fa["LH_vdw_ht"],
        LH_kstr=fa["LH_kstr"], LH_kben=fa["LH_kben"], LH_streq=fa["LH_streq"],
        LH_betaeq=fa["LH_betaeq"], LH_radius=fa["LH_radius"],
        nc_lh_flag=fa["nc_lh_flag"].astype(np.int32), lh_rad=fa["lh_rad"], phys=ph,
        groups_per_core=int(fa["groups_per_core"]), hi_all=0)
    st = S.State(r=z["st_r"], a=z["st_a"], b=z["st_b"], c=z["st_c"], r_t=z["st_r_t"],
                 r_lh=z["st_r_lh"])
    out = {k: z[k] for k in z.files if not (k.startswith("fib_") or k.startswith("phys_")
                                            or k.startswith("st_"))}
    return fib, st, out


def frame_err(a, b):
    """Max deviation of two sets of unit frame vectors (conditioning-aware
    comparison of derived angles: compare the frames they come from)."""
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) if np.size(a) else 0.0
