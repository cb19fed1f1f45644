/*
 * bd_hip.h -- C-ABI of the MI355X Brownian-dynamics step (libbdhip.so).
 *
 * The step is the Ermak-McCammon predictor-corrector of the reference
 * (BD_cuda/func_cuda.cu:3434-3680, main_cuda) with the CPU semantics of
 * BD_cuda/func.cpp.  Plain pointers and sizes only; every call returns 0 on
 * success or a negative BD_E_* code, with a message in bd_last_error().
 *
 * The reference's own entry points (func_cuda.h) are provided on top of this
 * API by include/func_cuda.h; this header is the extended interface used by
 * the parity harness, the benchmark and the Python host mirror.
 */
#ifndef BD_HIP_H
#define BD_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BD_OK 0
#define BD_E_ARG (-1)      /* invalid argument or size                      */
#define BD_E_HIP (-2)      /* HIP runtime error                              */
#define BD_E_STATE (-3)    /* call out of order (no state set, ...)          */
#define BD_E_CAPACITY (-4) /* a size exceeds what the context was made for  */

/* Static description of a fiber.  Sizes, per-bead constants and physics
 * scalars; the same quantities the reference passes to
 * cuda_application_init_data (func_cuda.h:48, func_cuda.cu:2676-2887).
 * Arrays are host pointers in the reference layout (AoS xyz, x[3i+k]) and
 * are copied at bd_create; the caller keeps ownership. */
typedef struct bd_system {
    int n, n_c;              /* main beads (cores + linker beads), cores        */
    int n_tail, n_t;         /* tail beads in total, tail beads per core        */
    int n_lh_n, n_lh_g, n_lh_c; /* LH template: N-term, globular head, C-term   */
    int n_LH;                /* free LH beads bound now ((N+C) per bound core)  */
    int Nq;                  /* point charges per core                          */
    int groups_per_core;     /* tail groups per core; reference literal 10       */
    int hi_all;              /* 1: D over every bead; 0: cores + linker beads    */
    const int *type;         /* n: 1 core, 0 linker bead                         */
    const double *phi_o;     /* n: intrinsic twist (func.cpp:412-431)           */
    const double *rad;       /* n: hydrodynamic radius                          */
    const double *core_pos;  /* 3 Nq: charge positions, nm, core body frame     */
    const double *core_q;    /* Nq                                              */
    const double *tail_pos;  /* 3 n_t: tail template, body frame                */
    const int *tail_fix_tpl; /* n_t: 1 = bead anchored to its core              */
    const double *h_t, *g_t, *lo_t, *beta_o_t, *t_q, *t_rad; /* n_tail           */
    const int *t_grp, *t_fix;                                 /* n_tail           */
    const double *LH_g_pos;  /* 3 n_lh_g: globular head, body frame             */
    const int *LH_conn;      /* n_lh_n + n_lh_g + n_lh_c                         */
    const double *LH_q, *LH_vdw_hh, *LH_vdw_hc, *LH_vdw_hl, *LH_vdw_ht;
    const double *LH_kstr, *LH_kben, *LH_streq, *LH_betaeq;   /* same length     */
    const int *nc_lh_flag;   /* n_c: 1 = linker histone bound                   */
    const double *lh_rad;    /* n_LH                                            */
    double h, g, s, debyell, debye, q_l, k_e, k_ex;
    double sig_dd, sig_dc, sig_cc, sig_tt, sig_tl, sig_tc;
    double a1, a2, s2dt, del, time_step;
} bd_system;

/* Cholesky arithmetic for the factor of D (the noise B = L). */
#define BD_CHOL_FP64 0      /* everything FP64 (FP64 MFMA updates)              */
#define BD_CHOL_FP32 1      /* trailing updates FP32 MFMA, FP32 accumulation    */
#define BD_CHOL_BF16 2      /* trailing updates as three-slice BF16 MFMA products
                               (each operand split exactly into three BF16
                               numbers, six products), FP32 accumulation   */

typedef struct bd_options {
    int device;             /* HIP device ordinal                               */
    int chol_mode;          /* BD_CHOL_*                                        */
    int use_graph;          /* 1: replay each step as one captured hipGraph      */
    int lh_capacity;        /* free LH beads to reserve (>= n_LH; 0 = n_LH)     */
} bd_options;

typedef struct bd_context bd_context;

/* Create a context on one GPU: copies the static data, allocates every
 * device buffer once (the reference reallocates per step,
 * func_cuda.cu:3436-3534).  Replaces cuda_application_init_D_Chol +
 * cuda_application_init_data (func_cuda.h:37,48). */
int bd_create(const bd_system *sys, const bd_options *opt, bd_context **out);
int bd_destroy(bd_context *ctx); /* free_all (func_cuda.h:52) */

/* Upload positions and frames (host).  With rebuild != 0 the derived frame
 * quantities are recomputed on the GPU, as main.cpp:164,223,283 do on the
 * host before the first step (linker frames are re-orthonormalised).
 * Any pointer may be NULL when its size is 0. */
int bd_set_state(bd_context *ctx, const double *r, const double *a, const double *b,
                 const double *c, const double *r_t, const double *r_lh, int rebuild);
int bd_get_state(bd_context *ctx, double *r, double *a, double *b, double *c, double *r_t,
                 double *r_lh);

/* Rebind the linker histones: which cores carry one (nc_lh_flag[n_c]), the
 * free-bead count n_LH = bound x (n_lh_n + n_lh_c) <= lh_capacity, their
 * positions and radii.  This is the bookkeeping main_cuda redoes every
 * step (func_cuda.cu:3436-3492) after the host's bind/unbind move
 * (main.cpp:386-465); here it is a copy into reserved buffers. */
int bd_set_lh(bd_context *ctx, const int *nc_lh_flag, int n_LH, const double *r_lh,
              const double *lh_rad);

/* Integrator scalars: RPY prefactors a1 = kT/(6 pi eta), a2 = kT/(8 pi eta),
 * s2dt = sqrt(2 dt), del = dt/kT and dt (main.cpp:88-99); bd_create takes
 * them from bd_system, the legacy main_cuda passes them every step. */
int bd_set_integrator(bd_context *ctx, double a1, double a2, double s2dt, double del,
                      double time_step);

/* Derived per-bead quantities (Euler angles, segment lengths, core exit
 * frames, tail and LH bending angles), in the layout of func.h:66. */
typedef struct bd_derived {
    double *alpha, *beta, *gamma, *length;     /* n    */
    double *a_dna, *b_dna, *c_dna;             /* 3n_c */
    double *alpha_p, *beta_p, *gamma_p;        /* n_c  */
    double *beta_t;                            /* n_tail */
    double *beta_lh;                           /* n_LH */
} bd_derived;
int bd_set_derived(bd_context *ctx, const bd_derived *d);
int bd_get_derived(bd_context *ctx, bd_derived *d);

/* One BD step with host noise: rr[3n] rotational, p[3(n+n_tail+n_LH)]
 * translational (main.cpp:479-486 draw order), energy of the predictor
 * pass (func_cuda.cu:3613-3615).  This is main_cuda without the legacy
 * H2D/D2H traffic. */
int bd_step(bd_context *ctx, const double *rr, const double *p, double *energy);

/* Same with noise already in device memory; energy written to a device
 * double (may be NULL).  No host synchronisation. */
int bd_step_device(bd_context *ctx, const double *d_rr, const double *d_p, double *d_energy);

/* k consecutive steps with host noise rr[k][3n] and p[k][3(n+n_tail+n_LH)]
 * (step-major, the draw order of main.cpp:479-486 for each step) and the k
 * pass-1 energies back in energies[k] (may be NULL): one upload, k steps,
 * one download.  Returns after the last step has finished.  The host loop of
 * main.cpp:384-512 between two linker-histone moves or two samples. */
int bd_run_steps(bd_context *ctx, int k, const double *rr, const double *p, double *energies);

/* Forces at the current state (one force pass, func_cuda.cu:3556-3613).
 * terms: bit mask BD_TERM_*.  Outputs are host arrays (NULL to skip). */
#define BD_TERM_MECH 1
#define BD_TERM_ELEC 2
#define BD_TERM_TAIL 4
#define BD_TERM_LH 8
#define BD_TERM_ALL 15
int bd_forces(bd_context *ctx, int terms, double *F, double *T, double *tF, double *lhF,
              double *energy);

/* Diffusion tensor and its factor at the current state.  bd_factor builds D
 * and factors it in place; bd_get_factor copies L (m x m row-major, lower,
 * m = 3 * n_hi) and the count of negative pivots replaced by |f|
 * (func.cpp:827-833).  bd_get_tensor assembles D only (no factor). */
int bd_factor(bd_context *ctx);
int bd_get_factor(bd_context *ctx, double *L, int *neg_pivots);
int bd_get_tensor(bd_context *ctx, double *D);
/* rd = s2dt * L * p for the current factor (rd_cal, func_cuda.cu:634-647);
 * in the FP32 / BF16 modes with refinement on (bd_set_refine), the refined
 * s2dt * L~ p' of the step.  The refinement reads the positions D was
 * assembled from, which an assembly keeps only while a refinement is on:
 * BD_E_STATE when the mode or the refinement was switched on after the
 * factor was formed (factor again). */
int bd_noise_displacement(bd_context *ctx, const double *p, double *rd);
int bd_set_chol_mode(bd_context *ctx, int chol_mode);

/* Partner search of the short-range terms (tail, linker-histone, head-site
 * and linker partners within the 7 nm cut-off of func.cpp:278-371, which
 * the reference evaluates over all pairs, func.cpp:1863-2022, 2109-2777):
 * enable = 1 through cell grids of edge >= 7 nm rebuilt every force pass,
 * 0 through bounding spheres of 16 consecutive partners.  Both evaluate
 * exactly the reference's nonzero pairs; only the summation order differs.
 * bd_create chooses cells for fibers of 20000 tail and LH beads or more;
 * the environment variable BD_CELLS overrides that choice. */
int bd_set_cell_list(bd_context *ctx, int enable);

/* FP64 refinement of the noise displacement in the BD_CHOL_FP32 / BF16
 * modes (the reference factors in FP64 throughout, func_cuda.cu:3513-3552,
 * so this has no counterpart there).  With L~ the reduced-precision factor,
 * rd = s2dt (L~ p + D u) / 2 with u = inv(L~)^T p: the first-order term of
 * B p for B = L~ (inv(L~) D inv(L~)^T)^(1/2), a square root of D itself, so
 * B B^T - D shrinks from O(E) to O(E^2) for a factor error E.  One blocked
 * FP64 triangular solve with the factor and one matrix-free D u product
 * from the positions D was assembled from.  enable = 1 (the default)
 * refines in the reduced-precision modes; FP64 mode never refines.  The
 * environment variable BD_REFINE sets the initial value.  A step after the
 * refinement is switched on forms D and L again even under frequency_RP. */
int bd_set_refine(bd_context *ctx, int enable);

/* Schedule of the blocked factor (no counterpart in the reference, which
 * calls cusolverDnDpotrf, func_cuda.cu:3526-3529).  panel_tiles = 0 (the
 * default) chooses: super-panels of 8 block columns of 64 from 64 block
 * columns (D of 4096 rows) up, one block column per launch below; -1 always
 * one block column per launch; k >= 2 super-panels of k block columns
 * whenever the matrix has more than k.  In a super-panel the panel columns
 * are factored with their updates restricted to the panel, then the
 * trailing matrix takes the panel's K = 64k update in one pass.  The
 * environment variable BD_PANEL_TILES sets the initial value.
 * bd_get_chol_schedule returns the super-panel width in use (0: none). */
int bd_set_chol_schedule(bd_context *ctx, int panel_tiles);
int bd_get_chol_schedule(bd_context *ctx, int *panel_tiles_used);

/* Cholesky factor of a caller's symmetric matrix in device memory with the
 * library's blocked factor: replaces cusolverDnDpotrf +
 * Cholesky_Decomposition_mod (func_cuda.cu:3513-3534, 614-630), with the CPU
 * rule of func.cpp:806-836 on a negative pivot (|f| taken and counted;
 * cuSOLVER instead stops and the reference maps the NaNs to 0).
 * D: m x m row-major, only its lower triangle read (ldd >= m); L: m x m
 * row-major (ldl >= m), the lower factor with zeros above the diagonal; D
 * and L may be the same buffer when ldd == ldl.  mode: BD_CHOL_*.  neg: host
 * int for the count of negative pivots (may be NULL).  stream: a hipStream_t
 * (NULL: the default stream); the call returns once the factor is done.
 * The workspace, (64 ceil(m/64))^2 + 4096 ceil(m/64) doubles, is allocated
 * and freed by the call. */
int bd_cholesky(int m, const double *D, int ldd, double *L, int ldl, int mode, int *neg,
                void *stream);

/* bd_cholesky without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The workspace is a
 * static device array of the library, BD_CHOLESKY_NOALLOC_MAX_M rows at
 * most (36.8 MB, held by every GPU the library runs on until the process
 * ends); larger m returns BD_E_CAPACITY.  All calls on one GPU share that
 * array: calls that overlap on one GPU (two host threads, or two streams)
 * race and must be serialised by the caller.  Same arguments and results as
 * bd_cholesky. */
#define BD_CHOLESKY_NOALLOC_MAX_M 2112
int bd_cholesky_noalloc(int m, const double *D, int ldd, double *L, int ldl, int mode, int *neg,
                        void *stream);

/* Core-core charge term of the force pass (func.cpp:1489-1566) and its
 * adjoint, for the differentiable trajectory (a torch.autograd.Function on
 * the Python side; no counterpart in the reference).  All arrays are device
 * pointers of one arithmetic: double when fp64 != 0, float otherwise.
 *   Q[n_c][Nq][3]  world positions of every core's charges
 *   q[Nq]          charges (one template shared by all cores)
 *   nbr_off[n_c+1], nbr[]: for each core, the cores it interacts with
 *                  (both directions: j in i's list iff i in j's)
 *   ke, kappa, kex, sigma: k_e, Debye kappa, k_ex, sigma_core-core
 *   f[n_c][Nq][3]  force on each charge (written)
 *   e[n_c]         half of the energy of each core's pairs (written)
 * bd_cc_backward takes gf = dLoss/df and ge = dLoss/de (same shapes) and
 * writes gQ = dLoss/dQ and gq_sites[n_c][Nq] = dLoss/dq per charge site
 * (the template's gradient is the sum over cores).  Asynchronous on
 * stream (a hipStream_t; NULL: the default stream). */
int bd_cc_forward(int fp64, int n_c, int Nq, const void *Q, const void *q, const int *nbr_off,
                  const int *nbr, double ke, double kappa, double kex, double sigma, void *f,
                  void *e, void *stream);
int bd_cc_backward(int fp64, int n_c, int Nq, const void *Q, const void *q, const int *nbr_off,
                   const int *nbr, double ke, double kappa, double kex, double sigma,
                   const void *gf, const void *ge, void *gQ, void *gq_sites, void *stream);

/* frequency_RP (main.cpp:45, read from setup.txt): rebuild D and its factor
 * only every frequency_rp-th step and reuse L for rd = s2dt L p in between,
 * the gate that func_cuda.cu:3504 has commented out.  1 (the default, and the
 * reference's behaviour) rebuilds every step.  Counting restarts, with a
 * rebuild, after bd_set_state, bd_set_chol_mode, bd_get_tensor and any
 * bd_set_lh that changes the hydrodynamic size; the drift is unaffected
 * (it uses the self-mobilities). */
int bd_set_frequency_rp(bd_context *ctx, int frequency_rp);

/* Per-kernel timing with HIP events on the context's stream (0 = off). */
int bd_set_timing(bd_context *ctx, int enable);
typedef struct bd_kernel_time {
    char name[48];
    double total_ms;
    long long launches;
} bd_kernel_time;
int bd_get_timing(bd_context *ctx, bd_kernel_time *out, int max_entries, int *n_entries);

/* Counters of the last force pass: charge pairs evaluated (core-core),
 * core pairs within 25 nm, and so on; used for algorithmic work figures. */
typedef struct bd_work_counts {
    long long core_pairs;          /* core-core pairs inside 25 nm                */
    long long core_charge_pairs;   /* charge-charge evaluations for those pairs   */
    long long core_linker_pairs;   /* core-linker pairs inside 25 nm              */
} bd_work_counts;
int bd_get_work_counts(bd_context *ctx, bd_work_counts *out);

/* Sizes of the context. */
int bd_get_sizes(bd_context *ctx, int *n, int *n_tail, int *n_LH, int *n_hi);

/* Device memory helpers for callers without their own allocator. */
int bd_device_alloc(size_t bytes, void **ptr);
int bd_device_free(void *ptr);
int bd_memcpy_htod(void *dst, const void *src, size_t bytes);
int bd_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte words; both
 * pointers 16-byte aligned, bytes a multiple of 16); synchronous.  Also the
 * known-size copy that HBM traffic counters are calibrated against. */
int bd_device_copy(void *dst, const void *src, size_t bytes);
int bd_synchronize(bd_context *ctx);

/* Host noise: MT19937 + Marsaglia polar normals with the cached second
 * value (replaces mt.cpp:48-141 and rand_normal, utilities.cpp:304-336).
 * bd_rng_step_noise draws one step in the order of main.cpp:386-388 and
 * 479-486 (two uniforms every 100th step, rr[n3] with sigma r_var[i],
 * then p[n_D3]); lh_uniforms (may be NULL) receives the two uniforms.
 * The state (624 words, index, cached normal) can be saved and restored,
 * which the reference cannot do (its seed is time(0), main.cpp:60-62). */
typedef struct bd_rng bd_rng;
int bd_rng_create(unsigned int seed, bd_rng **out);
int bd_rng_destroy(bd_rng *g);
unsigned int bd_rng_u32(bd_rng *g);
double bd_rng_uniform(bd_rng *g);
double bd_rng_normal(bd_rng *g, double mean, double stddev);
int bd_rng_step_noise(bd_rng *g, int step, int n3, const double *r_var, int n_D3, double *rr,
                      double *p, double *lh_uniforms);
int bd_rng_get_state(bd_rng *g, unsigned int *words, int *idx, int *cached, double *n2);
int bd_rng_set_state(bd_rng *g, const unsigned int *words, int idx, int cached, double n2);

const char *bd_last_error(void);
int bd_version(void);

#ifdef __cplusplus
}
#endif
#endif
