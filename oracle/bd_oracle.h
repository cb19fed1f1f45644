/*
 * bd_oracle.h -- CPU restatement of the reference Brownian-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function restates one reference function (file:line cited at the
 * definition in bd_oracle.c) in plain C99, FP64, single thread, with the
 * reference's loop order and quirks kept.  Arrays use the reference layout:
 * AoS xyz (x[3i+k]) and row-major matrices.
 */
#ifndef BD_ORACLE_H
#define BD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Static description of a fiber: sizes, per-bead constants, physics. */
typedef struct or_system {
    int n, n_c;            /* main beads (cores + linker beads), cores      */
    int n_tail, n_t;       /* tail beads in total, per-core template beads  */
    int n_lh_n, n_lh_g, n_lh_c; /* LH template: N-term, globular, C-term    */
    int n_LH;              /* free LH beads currently bound (N + C per core) */
    int Nq;                /* charges per core                              */
    int groups_per_core;   /* reference literal 10 (func.cpp:2368)          */
    int hi_all;            /* 1: D over every bead, 0: cores + linkers only */
    const int *type;       /* n: 1 core, 0 linker                           */
    const double *phi_o;   /* n                                             */
    const double *rad;     /* n                                             */
    const double *core_pos, *core_q;         /* 3Nq (nm, body), Nq          */
    const double *tail_pos; const int *tail_fix_tpl; /* 3n_t, n_t           */
    const double *h_t, *g_t, *lo_t, *beta_o_t, *t_q, *t_rad; /* n_tail     */
    const int *t_grp, *t_fix;                /* n_tail                      */
    const double *LH_g_pos;                  /* 3 n_lh_g                    */
    const int *LH_conn;                      /* n_lh_n+n_lh_g+n_lh_c        */
    const double *LH_q, *LH_vdw_hh, *LH_vdw_hc, *LH_vdw_hl, *LH_vdw_ht;
    const double *LH_kstr, *LH_kben, *LH_streq, *LH_betaeq;
    const int *nc_lh_flag;                   /* n_c                         */
    const double *lh_rad;                    /* n_LH                        */
    double h, g, s, debyell, debye, q_l, k_e, k_ex;
    double sig_dd, sig_dc, sig_cc, sig_tt, sig_tl, sig_tc;
    double a1, a2, s2dt, del, time_step;
} or_system;

/* Dynamic state (positions, frames) and the quantities derived from it. */
typedef struct or_state {
    double *r, *a, *b, *c;                   /* 3n                          */
    double *alpha, *beta, *gamma, *length;   /* n                           */
    double *a_dna, *b_dna, *c_dna;           /* 3 n_c                       */
    double *alpha_p, *beta_p, *gamma_p;      /* n_c                         */
    double *r_t, *beta_t;                    /* 3 n_tail, n_tail            */
    double *r_lh, *beta_lh;                  /* 3 n_LH, n_LH                */
} or_state;

/* Optional per-step intermediates (any pointer may be NULL). */
typedef struct or_step_out {
    double *D;        /* n_hi3 x n_hi3 row-major diffusion tensor          */
    double *L;        /* n_hi3 x n_hi3 row-major lower Cholesky factor     */
    double *rd;       /* n_D3 noise displacement                            */
    double *F0, *T0, *tF0, *lhF0;   /* predictor-pass forces                */
    double *F1, *T1, *tF1, *lhF1;   /* corrector-pass forces                */
    int neg_pivots;   /* Cholesky pivots that were negative (|f| taken)     */
} or_step_out;

/* constants.h:6-12 */
#define OR_PI    3.141592653589790
#define OR_LO    3.0
#define OR_RO    4.8
#define OR_D1    1.8
#define OR_THETA (1.2 * OR_PI / 2)
#define OR_ETA   1.137076e-12
#define OR_RH    1.2

double or_twist_angle_phi(int w);
void or_update_phi_o(int n, const int *type, double *phi_o);
void or_update_euler(const or_system *sys, double *r, double *a, double *b, double *c,
                     double *alpha, double *beta, double *gamma, double *length,
                     double *a_dna, double *b_dna, double *c_dna,
                     double *alpha_p, double *beta_p, double *gamma_p);
void or_update_tail_beta(int n_tail, const double *r_t, double *beta_t, const int *t_grp);
void or_update_lh_beta(int n_LH, int n_lh_c, const double *r_lh, double *beta_lh);
void or_build_tail(const or_system *sys, const double *tail_q, const int *tail_grp,
                   const double *tail_rad, const double *bond_c, const double *bond_v,
                   const double *angle_c, const double *angle_v,
                   const double *r, const double *a, const double *b, const double *c,
                   double *r_t, double *h_t, double *g_t, double *lo_t, double *beta_o_t,
                   double *t_q, double *t_rad, int *t_grp, int *t_fix);
void or_build_lh(int n, const int *type, int n_lh_n, int n_lh_c, const double *LH_n_pos,
                 const double *LH_c_pos, const double *LH_radius, const int *nc_lh_flag,
                 const double *r, const double *a, const double *b, const double *c,
                 double *r_lh, double *lh_rad);

/* RPY tensor over nb beads (row-major 3nb x 3nb, full). */
void or_diffusion_tensor(int nb, const double *r, const double *rad, double a1, double a2,
                         double *D);
/* Row-oriented Cholesky of an m x m row-major matrix; returns #negative pivots. */
int or_cholesky(int m, const double *D, double *L);

/* Force terms; each accumulates into F/T/tF/lhF and *E. */
void or_mech(const or_system *sys, const or_state *st, double *F, double *T, double *E);
void or_elec(const or_system *sys, const or_state *st, double *F, double *T, double *E);
void or_tail(const or_system *sys, const or_state *st, double *F, double *T, double *tF,
             double *E);
void or_lh(const or_system *sys, const or_state *st, double *F, double *T, double *tF,
           double *lhF, double *E);
/* Zero the outputs and run the four terms in the reference order. */
void or_forces(const or_system *sys, const or_state *st, double *F, double *T, double *tF,
               double *lhF, double *E);

void or_rotate(int n, const double *a, const double *b, const double *c, double *a_n,
               double *b_n, double *c_n, const double *d_theta, double dt);

/* One Ermak-McCammon predictor-corrector step (in place on st). */
int or_step(const or_system *sys, or_state *st, const double *rr, const double *p,
            double *energy, or_step_out *out);
/* The same step with the factor of an earlier step (frequency_RP,
 * main.cpp:45; the rebuild gate that func_cuda.cu:3504 has commented out):
 * rd = s2dt L_reuse p, while D, which the drift uses, is assembled at the
 * current state as every step does (L_reuse: n_hi3 x n_hi3 row-major). */
int or_step_reuse(const or_system *sys, or_state *st, const double *rr, const double *p,
                  const double *L_reuse, double *energy, or_step_out *out);

/* Host RNG: MT19937 (mt.cpp:48-141) and the polar normal sampler
 * (utilities.cpp:304-336) with its cached second value. */
typedef struct or_rng {
    uint32_t mt[624];
    int mti;
    int cached;
    double n2;
} or_rng;
void or_rng_seed(or_rng *g, uint32_t s);
uint32_t or_rng_u32(or_rng *g);
double or_rng_real1(or_rng *g);
double or_rng_normal(or_rng *g, double mean, double stddev);
/* Draw one step's noise in the reference order (main.cpp:386-388,479-486). */
void or_rng_step_noise(or_rng *g, int step, int n3, const double *r_var, int n_D3,
                       double *rr, double *p);

#ifdef __cplusplus
}
#endif
#endif
