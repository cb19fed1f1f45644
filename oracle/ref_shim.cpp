// ref_shim.cpp -- extern "C" entry points over the reference CPU library.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's own
// BD_cuda/func.cpp, utilities.cpp and mt.cpp (read where they lie, never
// copied) into oracle/_ref/libbdref.so by oracle/Makefile.  The reference
// functions take std::vector arguments and keep RNG state in globals; this
// file only adapts plain pointers to those signatures, and composes one
// predictor-corrector step from them in the order of main_cuda
// (func_cuda.cu:3434-3680), which has no CPU driver in the reference.
// Functions that write side files (Diffusion_Tensor -> wrongD.txt,
// Linker_Histone_force -> wf.txt) write them into the current directory, and
// so do the reference's writers of readfile.cpp (out.xyz, restart_*.txt),
// which are exported here to pin the host driver's output formats.
#include "func.h"
#include "readfile.h"

#include <cstring>
#include <vector>

namespace {

std::vector<double> vd(const double *p, int n) { return std::vector<double>(p, p + (n > 0 ? n : 0)); }
std::vector<int> vi(const int *p, int n) { return std::vector<int>(p, p + (n > 0 ? n : 0)); }

}  // namespace

// Must match or_system in bd_oracle.h field by field.
extern "C" {

struct ref_system {
    int n, n_c, n_tail, n_t, n_lh_n, n_lh_g, n_lh_c, n_LH, Nq, groups_per_core, hi_all;
    const int *type;
    const double *phi_o, *rad, *core_pos, *core_q, *tail_pos;
    const int *tail_fix_tpl;
    const double *h_t, *g_t, *lo_t, *beta_o_t, *t_q, *t_rad;
    const int *t_grp, *t_fix;
    const double *LH_g_pos;
    const int *LH_conn;
    const double *LH_q, *LH_vdw_hh, *LH_vdw_hc, *LH_vdw_hl, *LH_vdw_ht, *LH_kstr, *LH_kben,
        *LH_streq, *LH_betaeq;
    const int *nc_lh_flag;
    const double *lh_rad;
    double h, g, s, debyell, debye, q_l, k_e, k_ex;
    double sig_dd, sig_dc, sig_cc, sig_tt, sig_tl, sig_tc;
    double a1, a2, s2dt, del, time_step;
};

struct ref_state {
    double *r, *a, *b, *c, *alpha, *beta, *gamma, *length, *a_dna, *b_dna, *c_dna, *alpha_p,
        *beta_p, *gamma_p, *r_t, *beta_t, *r_lh, *beta_lh;
};

void ref_init_genrand(unsigned long s) { init_genrand(s); }
double ref_genrand_real1(void) { return genrand_real1(); }
double ref_rand_normal(double mean, double sd) { return rand_normal(mean, sd); }
double ref_twist_angle_phi(int w) { return twist_angle_phi(w); }

void ref_update_phi_o(int n, int *type, double *phi_o) { update_phi_o(n, type, phi_o); }

void ref_update_euler(const ref_system *S, ref_state *X) {
    update_Euler_Angle(S->n_c, 3 * S->n_c, S->n, 3 * S->n, const_cast<int *>(S->type), X->r, X->a,
                       X->b, X->c, X->alpha, X->beta, X->gamma, X->length, X->a_dna, X->b_dna,
                       X->c_dna, X->alpha_p, X->beta_p, X->gamma_p);
}

void ref_update_tail_beta(int n_tail, double *r_t, double *beta_t, int *t_grp) {
    // the reference reads t_grp[n_tail] and t_grp[n_tail+1]; pad a copy
    std::vector<int> g(t_grp, t_grp + n_tail);
    g.push_back(-7);
    g.push_back(-9);
    update_tail_beta(n_tail, r_t, beta_t, g.data());
}

void ref_update_lh_beta(int n_LH, int n_lh_c, double *r_lh, double *beta_lh) {
    update_LH_beta(n_LH, n_lh_c, r_lh, beta_lh);
}

// Diffusion_Tensor over nb beads; D and L row-major 3nb x 3nb.
void ref_diffusion_tensor(int nb, double *r, double *rad, double a1, double a2, double *D,
                          double *L) {
    const int m = 3 * nb;
    std::vector<std::vector<double>> Dv(m, std::vector<double>(m, 0.0)),
        Lv(m, std::vector<double>(m, 0.0));
    Diffusion_Tensor(nb, m, r, a1, a2, rad, Dv, Lv);
    for (int i = 0; i < m; i++) {
        if (D) std::memcpy(D + (size_t)i * m, Dv[i].data(), sizeof(double) * m);
        if (L) std::memcpy(L + (size_t)i * m, Lv[i].data(), sizeof(double) * m);
    }
}

void ref_mech(const ref_system *S, ref_state *X, double *F, double *T, double *E) {
    double e = *E;
    mechanical_force_and_torque(S->n_c, 3 * S->n_c, S->n, 3 * S->n, const_cast<int *>(S->type),
                                X->r, X->a, X->b, X->c, X->alpha, X->beta, X->gamma, X->length,
                                X->a_dna, X->b_dna, X->c_dna, X->alpha_p, X->beta_p, X->gamma_p,
                                S->h, S->g, S->s, const_cast<double *>(S->phi_o), F, T, e, 0, 1);
    *E = e;
}

void ref_elec(const ref_system *S, ref_state *X, double *F, double *T, double *E) {
    double e = *E;
    Electrostatic_and_Excluded_volume_force(
        S->n, 3 * S->n, S->n_c, 3 * S->n_c, const_cast<int *>(S->type), X->r, X->a, X->b, X->c,
        S->debyell, S->debye, S->q_l, S->k_e, S->k_ex, 0.0, S->sig_dd, S->sig_dc, S->sig_cc, S->Nq,
        3 * S->Nq, vd(S->core_pos, 3 * S->Nq), vd(S->core_q, S->Nq), F, T, e, 0, 1);
    *E = e;
}

void ref_tail(const ref_system *S, ref_state *X, double *F, double *T, double *tF, double *E) {
    double e = *E;
    std::vector<int> nct(S->n_c, 1);
    tail_force(S->n, S->n_t, const_cast<int *>(S->type), S->n_tail, 3 * S->n_tail,
               vd(S->tail_pos, 3 * S->n_t), vi(S->tail_fix_tpl, S->n_t), nct.data(), X->r_t,
               X->beta_t, const_cast<double *>(S->h_t), const_cast<double *>(S->g_t),
               const_cast<double *>(S->lo_t), const_cast<double *>(S->beta_o_t), S->h, S->k_e,
               S->debye, S->k_ex, const_cast<double *>(S->t_q), const_cast<double *>(S->t_rad),
               const_cast<int *>(S->t_grp), const_cast<int *>(S->t_fix), X->r, X->a, X->b, X->c,
               S->q_l, S->Nq, 3 * S->Nq, vd(S->core_pos, 3 * S->Nq), vd(S->core_q, S->Nq), tF, F,
               T, S->sig_tt, S->sig_tl, S->sig_tc, e, 0, 1);
    *E = e;
}

void ref_lh(const ref_system *S, ref_state *X, double *F, double *T, double *tF, double *lhF,
            double *E) {
    double e = *E;
    int ntpl = S->n_lh_n + S->n_lh_g + S->n_lh_c;
    Linker_Histone_force(
        S->n, S->n_tail, S->n_lh_n, S->n_lh_g, S->n_lh_c, const_cast<int *>(S->t_grp),
        const_cast<int *>(S->t_fix), const_cast<int *>(S->type), S->n_LH, 3 * S->n_LH,
        vd(S->LH_g_pos, 3 * S->n_lh_g), vi(S->LH_conn, ntpl), const_cast<int *>(S->nc_lh_flag),
        X->beta_lh, X->r_lh, vd(S->LH_q, ntpl), S->k_e, S->debye, S->k_ex, vd(S->LH_vdw_hh, ntpl),
        vd(S->LH_vdw_hc, ntpl), vd(S->LH_vdw_hl, ntpl), vd(S->LH_vdw_ht, ntpl),
        vd(S->LH_kstr, ntpl), vd(S->LH_kben, ntpl), vd(S->LH_streq, ntpl), vd(S->LH_betaeq, ntpl),
        X->r, X->a, X->b, X->c, S->q_l, S->Nq, 3 * S->Nq, vd(S->core_pos, 3 * S->Nq),
        vd(S->core_q, S->Nq), X->r_t, const_cast<double *>(S->t_q), tF, lhF, F, T, e, 0, 1);
    *E = e;
}

void ref_forces(const ref_system *S, ref_state *X, double *F, double *T, double *tF, double *lhF,
                double *E) {
    std::memset(F, 0, sizeof(double) * 3 * S->n);
    std::memset(T, 0, sizeof(double) * 3 * S->n);
    if (S->n_tail) std::memset(tF, 0, sizeof(double) * 3 * S->n_tail);
    if (S->n_LH) std::memset(lhF, 0, sizeof(double) * 3 * S->n_LH);
    *E = 0.0;
    ref_mech(S, X, F, T, E);
    ref_elec(S, X, F, T, E);
    if (S->n_tail) ref_tail(S, X, F, T, tF, E);
    if (S->n_LH) ref_lh(S, X, F, T, tF, lhF, E);
}

void ref_rotate(int n, double *a, double *b, double *c, double *a_n, double *b_n, double *c_n,
                double *d_theta, double dt) {
    rotate(n, 3 * n, a, b, c, a_n, b_n, c_n, d_theta, dt);
}

// One predictor-corrector step built from the reference functions, in the
// order of main_cuda.  Only the pieces with no CPU function in the reference
// (rd = s2dt*L*p, the species-block drift, the rotation increment) are
// written here; D is formed over all beads (hi_all) or the main beads.
int ref_step(const ref_system *S, ref_state *X, const double *rr, const double *p,
             double *energy) {
    const int n = S->n, n3 = 3 * n, nt = S->n_tail, nt3 = 3 * nt, nl = S->n_LH, nl3 = 3 * nl;
    const int n_D = n + nt + nl, n_D3 = 3 * n_D, n_hi = S->hi_all ? n_D : n, m = 3 * n_hi;
    std::vector<double> r_all(n_D3), rad_all(n_D);
    std::memcpy(r_all.data(), X->r, sizeof(double) * n3);
    if (nt) std::memcpy(r_all.data() + n3, X->r_t, sizeof(double) * nt3);
    if (nl) std::memcpy(r_all.data() + n3 + nt3, X->r_lh, sizeof(double) * nl3);
    for (int i = 0; i < n; i++) rad_all[i] = S->rad[i];
    for (int i = 0; i < nt; i++) rad_all[n + i] = S->t_rad[i];
    for (int i = 0; i < nl; i++) rad_all[n + nt + i] = S->lh_rad[i];
    if (nl) update_LH_beta(nl, S->n_lh_c, X->r_lh, X->beta_lh);

    std::vector<std::vector<double>> D(m, std::vector<double>(m, 0.0)),
        L(m, std::vector<double>(m, 0.0));
    Diffusion_Tensor(n_hi, m, r_all.data(), S->a1, S->a2, rad_all.data(), D, L);
    std::vector<double> rd(n_D3);
    for (int row = 0; row < n_D3; row++) {
        double v = 0.0;
        if (row < m)
            for (int k = 0; k <= row; k++) v += S->s2dt * L[row][k] * p[k];
        else
            v = S->s2dt * std::sqrt(S->a1 / rad_all[row / 3]) * p[row];
        rd[row] = v;
    }

    std::vector<double> F0(n3), T0(n3), F1(n3), T1(n3), tF0(nt3 + 3), tF1(nt3 + 3), lF0(nl3 + 3),
        lF1(nl3 + 3);
    double E0 = 0, E1 = 0;
    ref_forces(S, X, F0.data(), T0.data(), tF0.data(), lF0.data(), &E0);

    std::vector<double> pr_r(n3), pr_a(n3), pr_b(n3), pr_c(n3), al(n), be(n), ga(n), le(n),
        ad(3 * S->n_c), bd(3 * S->n_c), cd(3 * S->n_c), alp(S->n_c), bep(S->n_c), gap(S->n_c),
        prt(nt3 + 3), bt(nt + 1), prl(nl3 + 3), bl(nl + 1), dth(n3);
    const double zc = 8 * PI * eta * 125.0, zl = 4 * PI * eta * r_h * r_h * lo;
    auto rot_inc = [&](const std::vector<double> &T) {
        for (int j = 0; j < n; j++) {
            if (S->type[j] != 0) {
                for (int k = 0; k < 3; k++)
                    dth[3 * j + k] = S->time_step * T[3 * j + k] / zc + rr[3 * j + k];
            } else {
                dth[3 * j] = S->time_step * T[3 * j] / zl + rr[3 * j];
                dth[3 * j + 1] = 0.0;
                dth[3 * j + 2] = 0.0;
            }
        }
    };
    auto drift = [&](const double *x0, const double *F, int off, int cnt3, const double *rad,
                     double *out) {
        for (int q = 0; q < cnt3; q++) {
            int row = off + q;
            double v = x0[q] + rd[row];
            if (row < m)
                for (int k = 0; k < cnt3; k++) v = v + S->del * D[row][off + k] * F[k];
            else
                v = v + S->del * (S->a1 / rad[q / 3]) * F[q];
            out[q] = v;
        }
    };

    rot_inc(T0);
    rotate(n, n3, X->a, X->b, X->c, pr_a.data(), pr_b.data(), pr_c.data(), dth.data(), 1.0);
    drift(X->r, F0.data(), 0, n3, S->rad, pr_r.data());
    if (nt3) drift(X->r_t, tF0.data(), n3, nt3, S->t_rad, prt.data());
    if (nl3) drift(X->r_lh, lF0.data(), n3 + nt3, nl3, S->lh_rad, prl.data());

    ref_state P{pr_r.data(), pr_a.data(), pr_b.data(), pr_c.data(), al.data(), be.data(),
                ga.data(),   le.data(),   ad.data(),   bd.data(),   cd.data(),  alp.data(),
                bep.data(),  gap.data(),  prt.data(),  bt.data(),   prl.data(), bl.data()};
    if (nt) ref_update_tail_beta(nt, prt.data(), bt.data(), const_cast<int *>(S->t_grp));
    if (nl) update_LH_beta(nl, S->n_lh_c, prl.data(), bl.data());
    ref_update_euler(S, &P);
    ref_forces(S, &P, F1.data(), T1.data(), tF1.data(), lF1.data(), &E1);

    std::vector<double> Fa(n3), Ta(n3), tFa(nt3 + 3), lFa(nl3 + 3);
    for (int q = 0; q < n3; q++) {
        Fa[q] = 0.5 * (F0[q] + F1[q]);
        Ta[q] = 0.5 * (T0[q] + T1[q]);
    }
    for (int q = 0; q < nt3; q++) tFa[q] = 0.5 * (tF0[q] + tF1[q]);
    for (int q = 0; q < nl3; q++) lFa[q] = 0.5 * (lF0[q] + lF1[q]);
    rot_inc(Ta);
    rotate(n, n3, X->a, X->b, X->c, pr_a.data(), pr_b.data(), pr_c.data(), dth.data(), 1.0);
    drift(X->r, Fa.data(), 0, n3, S->rad, pr_r.data());
    if (nt3) drift(X->r_t, tFa.data(), n3, nt3, S->t_rad, prt.data());
    if (nl3) drift(X->r_lh, lFa.data(), n3 + nt3, nl3, S->lh_rad, prl.data());

    std::memcpy(X->r, pr_r.data(), sizeof(double) * n3);
    std::memcpy(X->a, pr_a.data(), sizeof(double) * n3);
    std::memcpy(X->b, pr_b.data(), sizeof(double) * n3);
    std::memcpy(X->c, pr_c.data(), sizeof(double) * n3);
    if (nt3) std::memcpy(X->r_t, prt.data(), sizeof(double) * nt3);
    if (nl3) std::memcpy(X->r_lh, prl.data(), sizeof(double) * nl3);
    ref_update_euler(S, X);
    if (nt) ref_update_tail_beta(nt, X->r_t, X->beta_t, const_cast<int *>(S->t_grp));
    if (nl) update_LH_beta(nl, S->n_lh_c, X->r_lh, X->beta_lh);
    if (energy) *energy = E0;
    return 0;
}

// build_LH, func.cpp:2029-2059
void ref_build_lh(int n, int *type, int n_LH, const double *LH_n_pos, int n_lh_n, const double *LH_c_pos,
                  int n_lh_c, const double *LH_radius, double *r, double *a, double *b, double *c,
                  int *nc_lh_flag, double *r_lh, double *lh_rad) {
    const int nlh = n_lh_n + n_lh_c;
    build_LH(n, type, n_LH, vd(LH_n_pos, 3 * n_lh_n), vd(LH_c_pos, 3 * n_lh_c), vd(LH_radius, nlh), r, a, b, c,
             n_lh_n, n_lh_c, nc_lh_flag, r_lh, lh_rad);
}

// write_xyz_append, readfile.cpp:346-405 (appends to `filename`)
void ref_write_xyz_append(int n, int n_c, int n_tail, int n_lh_g, int nc_lh, int *nc_lh_flag, int n_LH, int Nq,
                          int *type, double *r, double *a, double *b, double *c, const double *core_pos,
                          double *r_t, const double *LH_g_pos, double *r_lh, const char *filename) {
    write_xyz_append(n, 3 * n, n_c, n_tail, n_lh_g, nc_lh, nc_lh_flag, n_LH, Nq, 3 * Nq, type, r, a, b, c,
                     vd(core_pos, 3 * Nq), r_t, vd(LH_g_pos, 3 * n_lh_g), r_lh, std::string(filename));
}

// write_restart, readfile.cpp:407-441 (restart_*.txt in the current directory)
void ref_write_restart(int n3, int n_tail3, int n_LH3, double *r, double *a, double *b, double *c, double *Er,
                       double *r_t, double *r_lh) {
    write_restart(n3, n_tail3, n_LH3, r, a, b, c, Er, r_t, r_lh);
}

// Read_restart_ini/_tail/_LH, readfile.cpp:443-561 (current directory)
void ref_read_restart(int n3, int n_tail3, int n_LH3, double *r, double *a, double *b, double *c, double *Er,
                      double *r_t, double *r_lh) {
    Read_restart_ini(n3, r, a, b, c, Er);
    if (n_tail3) Read_restart_tail(n_tail3, r_t);
    if (n_LH3) Read_restart_LH(n_LH3, r_lh);
}

// Read_initial_struc, readfile.cpp:40-134: the per-coordinate drag
// coefficients Er of `filename` (3 per bead)
int ref_initial_er(const char *filename, double *Er, int max3) {
    int n, n_c, n3, nc3;
    std::vector<double> r, a, b, c, rad, er;
    std::vector<int> type;
    Read_initial_struc(n, n_c, n3, nc3, r, a, b, c, rad, er, type, std::string(filename));
    if ((int)er.size() > max3) return -1;
    std::memcpy(Er, er.data(), sizeof(double) * er.size());
    return (int)er.size();
}

}  // extern "C"
