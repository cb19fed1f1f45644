// bd_legacy.hip -- the reference's four step entry points (func_cuda.h:37,
// 48, 50, 52) implemented over the context API of bd_hip.h.
//
// The reference keeps all device state in file-scope globals
// (func_cuda.cu:3-104); here one process-wide context plays that role.
// Errors, which the reference never reports, are printed to stderr once
// per call and leave the host arrays untouched.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/bd_hip.h"
#include "../../include/func_cuda.h"

namespace {

struct Legacy {
    bd_context *ctx = nullptr;
    int n_D3 = 0;
    int n = 0, n_c = 0, n_tail = 0, nlh = 0;
    std::vector<int> flags;
    int n_LH = -1;
    double a1 = 0, a2 = 0, s2dt = 0, del = 0, dt = 0;
    int rp = 1;
};
Legacy g_leg;

void report(const char *where, int rc) {
    if (rc) std::fprintf(stderr, "%s: %s (code %d)\n", where, bd_last_error(), rc);
}

}  // namespace

void cuda_application_init_D_Chol(int n3) { g_leg.n_D3 = n3; }

void cuda_application_init_data(int n_c, int nc3, int n, int n3, int *type, double *r, double *a, double *b,
                                double *c, double *alpha, double *beta, double *gamma, double *length,
                                double *a_dna, double *b_dna, double *c_dna, double *alpha_p, double *beta_p,
                                double *gamma_p, double h, double g, double s, double *phi_o, double debyell,
                                double debye, double q_l, double k_e, double k_ex, double k_h1,
                                double sigma_DNA_DNA, double sigma_DNA_Core, double sigma_Core_Core,
                                double sigma_Tail_Tail, double sigma_Tail_Linker, double sigma_Tail_Core, int Nq,
                                int Nq3, double *core_pos, double *core_q, int n_t, int n_tail, int n_tail3,
                                double *tail_pos, int *tail_fix, int *nc_t_flag, double *r_t, double *beta_t,
                                double *h_t, double *g_t, double *lo_t, double *beta_o_t, double *t_q,
                                double *t_rad, int *t_grp, int *t_fix, int n_lh_n, int n_lh_g, int n_lh_c,
                                int n_LH, int n_LH3, double *LH_g_pos, int *LH_conn, int *nc_lh_flag,
                                double *beta_lh, double *r_lh, double *LH_q, double *LH_vdw_hh, double *LH_vdw_hc,
                                double *LH_vdw_hl, double *LH_vdw_ht, double *LH_kstr, double *LH_kben,
                                double *LH_streq, double *LH_betaeq, double *LH_force, double *t_force,
                                double *force, double *torque, double *Energy, double *r_all, double *rad_all) {
    (void)nc3; (void)n3; (void)k_h1; (void)Nq3; (void)n_tail3; (void)nc_t_flag; (void)n_LH3;
    (void)LH_force; (void)t_force; (void)force; (void)torque; (void)Energy; (void)r_all;
    if (g_leg.ctx) { bd_destroy(g_leg.ctx); g_leg.ctx = nullptr; }
    const char *hs = std::getenv("BD_HI_SET");
    bd_system sys{};
    sys.n = n; sys.n_c = n_c; sys.n_tail = n_tail; sys.n_t = n_t;
    sys.n_lh_n = n_lh_n; sys.n_lh_g = n_lh_g; sys.n_lh_c = n_lh_c; sys.n_LH = n_LH; sys.Nq = Nq;
    sys.groups_per_core = 10;  // func.cpp:2368, func_cuda.cu:2066
    sys.hi_all = (hs && std::strcmp(hs, "main") == 0) ? 0 : 1;
    sys.type = type; sys.phi_o = phi_o;
    std::vector<double> rad(rad_all, rad_all + n);
    sys.rad = rad.data();
    sys.core_pos = core_pos; sys.core_q = core_q;
    sys.tail_pos = tail_pos; sys.tail_fix_tpl = tail_fix;
    sys.h_t = h_t; sys.g_t = g_t; sys.lo_t = lo_t; sys.beta_o_t = beta_o_t; sys.t_q = t_q; sys.t_rad = t_rad;
    sys.t_grp = t_grp; sys.t_fix = t_fix;
    sys.LH_g_pos = LH_g_pos; sys.LH_conn = LH_conn; sys.LH_q = LH_q; sys.LH_vdw_hh = LH_vdw_hh;
    sys.LH_vdw_hc = LH_vdw_hc; sys.LH_vdw_hl = LH_vdw_hl; sys.LH_vdw_ht = LH_vdw_ht; sys.LH_kstr = LH_kstr;
    sys.LH_kben = LH_kben; sys.LH_streq = LH_streq; sys.LH_betaeq = LH_betaeq;
    sys.nc_lh_flag = nc_lh_flag;
    sys.lh_rad = rad_all + n + n_tail;
    sys.h = h; sys.g = g; sys.s = s; sys.debyell = debyell; sys.debye = debye; sys.q_l = q_l;
    sys.k_e = k_e; sys.k_ex = k_ex;
    sys.sig_dd = sigma_DNA_DNA; sys.sig_dc = sigma_DNA_Core; sys.sig_cc = sigma_Core_Core;
    sys.sig_tt = sigma_Tail_Tail; sys.sig_tl = sigma_Tail_Linker; sys.sig_tc = sigma_Tail_Core;
    // a1, a2, s2dt, del, time_step arrive with every main_cuda call
    bd_options opt{0, BD_CHOL_FP64, 1, n_c * (n_lh_n + n_lh_c)};
    int rc = bd_create(&sys, &opt, &g_leg.ctx);
    if (rc) { report("cuda_application_init_data", rc); g_leg.ctx = nullptr; return; }
    rc = bd_set_state(g_leg.ctx, r, a, b, c, r_t, r_lh, 0);
    if (!rc) {
        bd_derived d{alpha, beta, gamma, length, a_dna, b_dna, c_dna, alpha_p, beta_p, gamma_p, beta_t, beta_lh};
        rc = bd_set_derived(g_leg.ctx, &d);
    }
    report("cuda_application_init_data", rc);
    g_leg.n = n; g_leg.n_c = n_c; g_leg.n_tail = n_tail; g_leg.nlh = n_lh_n + n_lh_c;
    g_leg.flags.assign(nc_lh_flag, nc_lh_flag + n_c);
    g_leg.n_LH = n_LH;
    g_leg.rp = 1;
}

void main_cuda(int n_c, int nc3, int step, int number_of_steps, double time_step, double del, int frequency_RP,
               int frequency_of_sampling, double h, double g, double s, double debyell, double debye, double q_l,
               double k_e, double k_ex, double k_h1, double sigma_DNA_DNA, double sigma_DNA_Core,
               double sigma_Core_Core, double sigma_Tail_Tail, double sigma_Tail_Linker, double sigma_Tail_Core,
               int Nq, int Nq3, int n_t, int n_lh_n, int n_lh_g, int n_lh_c, int n, int n3, int n_tail,
               int n_tail3, int n_LH, int n_LH3, double a1, double a2, double s2dt, double *rr, double *p,
               double *Energy, double *h_r, double *h_a, double *h_b, double *h_c, double *h_r_t, double *h_r_lh,
               double *h_rad_all, int *nc_lh_flag) {
    (void)nc3; (void)h; (void)g; (void)s; (void)debyell; (void)debye; (void)q_l;
    (void)k_e; (void)k_ex; (void)k_h1; (void)sigma_DNA_DNA; (void)sigma_DNA_Core; (void)sigma_Core_Core;
    (void)sigma_Tail_Tail; (void)sigma_Tail_Linker; (void)sigma_Tail_Core; (void)Nq; (void)Nq3; (void)n_t;
    (void)n_lh_n; (void)n_lh_g; (void)n_lh_c; (void)n3; (void)n_tail3; (void)n_LH3;
    bd_context *ctx = g_leg.ctx;
    if (!ctx) { std::fprintf(stderr, "main_cuda: no context (cuda_application_init_data failed)\n"); return; }
    // linker-histone arrays: the host copy is authoritative each step (func_cuda.cu:3436-3492)
    int rc = bd_set_lh(ctx, nc_lh_flag, n_LH, h_r_lh, h_rad_all + n + n_tail);
    if (rc) { report("main_cuda", rc); return; }
    if (a1 != g_leg.a1 || a2 != g_leg.a2 || s2dt != g_leg.s2dt || del != g_leg.del || time_step != g_leg.dt) {
        rc = bd_set_integrator(ctx, a1, a2, s2dt, del, time_step);
        if (rc) { report("main_cuda", rc); return; }
        g_leg.a1 = a1; g_leg.a2 = a2; g_leg.s2dt = s2dt; g_leg.del = del; g_leg.dt = time_step;
    }
    // frequency_RP is ignored by the reference (func_cuda.cu:3504);
    // BD_FREQUENCY_RP=1 in the environment honours it
    const char *fr = std::getenv("BD_FREQUENCY_RP");
    const int want_rp = (fr && std::strcmp(fr, "1") == 0 && frequency_RP > 1) ? frequency_RP : 1;
    if (want_rp != g_leg.rp) {
        rc = bd_set_frequency_rp(ctx, want_rp);
        if (rc) { report("main_cuda", rc); return; }
        g_leg.rp = want_rp;
    }
    rc = bd_step(ctx, rr, p, Energy);
    if (rc) { report("main_cuda", rc); return; }
    // D2H r_lh every step; r, a, b, c, r_t on sampling steps (3662-3674)
    const bool sample = (step % frequency_of_sampling == 0) || step == number_of_steps - 1;
    rc = bd_get_state(ctx, sample ? h_r : nullptr, sample ? h_a : nullptr, sample ? h_b : nullptr,
                      sample ? h_c : nullptr, sample ? h_r_t : nullptr, h_r_lh);
    report("main_cuda", rc);
    (void)n_c;
}

void free_all() {
    if (g_leg.ctx) bd_destroy(g_leg.ctx);
    g_leg.ctx = nullptr;
}
