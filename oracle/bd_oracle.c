/*
 * bd_oracle.c -- CPU restatement of the reference Brownian-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY (see bd_oracle.h).  Each function names the
 * reference function and lines it restates.  Loop orders and the
 * reference's quirks (force clamps, torque overwrites, index maps, the
 * literal 10 groups per core, |pivot| in the Cholesky) are kept on purpose;
 * they are what the HIP product is checked against.
 */
#include "bd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* small vector helpers                                                 */
/* ------------------------------------------------------------------ */
static inline double dot3(const double *u, const double *v) {
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}
/* u.v + w.x summed left to right, as the reference writes its 6-term sums */
static inline double dot6(const double *u, const double *v, const double *w, const double *x) {
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2] + w[0] * x[0] + w[1] * x[1] + w[2] * x[2];
}
static inline void cross3(const double *u, const double *v, double *w) {
    /* utilities.cpp:139-148 (component order kept) */
    double t0 = u[1] * v[2] - u[2] * v[1];
    double t1 = u[2] * v[0] - u[0] * v[2];
    double t2 = u[0] * v[1] - u[1] * v[0];
    w[0] = t0; w[1] = t1; w[2] = t2;
}
static inline double dist3(const double *p, const double *q) {
    double x = p[0] - q[0], y = p[1] - q[1], z = p[2] - q[2];
    return sqrt(x * x + y * y + z * z);
}
static inline void add3(double *dst, const double *v) {
    dst[0] += v[0]; dst[1] += v[1]; dst[2] += v[2];
}
static inline void sub3(double *dst, const double *v) {
    dst[0] -= v[0]; dst[1] -= v[1]; dst[2] -= v[2];
}
/* body-frame point p (components along a,b,c of frame at r) to world */
static inline void body_to_world(const double *r, const double *a, const double *b,
                                 const double *c, const double *p, double *w) {
    for (int k = 0; k < 3; k++) w[k] = r[k] + a[k] * p[0] + b[k] * p[1] + c[k] * p[2];
}
/* torque_due_to_force_relative, func.cpp:396-410 */
static void torque_rel(const double *f, const double *comp, const double *a, const double *b,
                       const double *c, double *t) {
    double fa = a[0] * f[0] + a[1] * f[1] + a[2] * f[2];
    double fb = b[0] * f[0] + b[1] * f[1] + b[2] * f[2];
    double fc = c[0] * f[0] + c[1] * f[1] + c[2] * f[2];
    t[0] = fc * comp[1] - fb * comp[2];
    t[1] = fa * comp[2] - fc * comp[0];
    t[2] = fb * comp[0] - fa * comp[1];
}
/* torque_due_to_force, func.cpp:374-394: lever arm = coord_f - coord_c in the frame */
static void torque_abs(const double *f, const double *pf, const double *pc, const double *a,
                       const double *b, const double *c, double *t) {
    double d[3] = {pf[0] - pc[0], pf[1] - pc[1], pf[2] - pc[2]};
    double comp[3] = {dot3(d, a), dot3(d, b), dot3(d, c)};
    torque_rel(f, comp, a, b, c, t);
}

/* ------------------------------------------------------------------ */
/* pair primitives, func.cpp:134-371                                    */
/* ------------------------------------------------------------------ */
static inline double clamp09(double v) {
    if (v > 0.9) v = 0.9;
    if (v < -0.9) v = -0.9;
    return v;
}

/* Force_Stretching, func.cpp:134-151 */
static void pr_stretch(double k, const double *p1, const double *p2, double l0, double *f1,
                       double *f2, double *E) {
    double d = dist3(p1, p2);
    *E += k * (d - l0) * (d - l0) / 2;
    double f = -k * (d - l0);
    for (int i = 0; i < 3; i++) {
        f1[i] = f * (p1[i] - p2[i]) / d;
        f2[i] = -f * (p1[i] - p2[i]) / d;
    }
}

/* Force_Bending, func.cpp:153-196 (with norm/cross_product of utilities.cpp) */
static void pr_bend(double kb, double beta, double beta0, const double *p1, const double *p2,
                    const double *p3, double *f1, double *f2, double *f3, double *E) {
    *E += kb * (beta - beta0) * (beta - beta0) / 2;
    double f = -kb * (beta - beta0);
    double d1 = dist3(p1, p2), d2 = dist3(p2, p3);
    double ji[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double jk[3] = {p2[0] - p3[0], p2[1] - p3[1], p2[2] - p3[2]};
    double kj[3] = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
    double n0[3], ri[3], rk[3];
    cross3(ji, jk, n0);
    cross3(ji, n0, ri);
    double nri = sqrt(dot3(ri, ri));
    for (int i = 0; i < 3; i++) {
        f1[i] = (f / d1) * ri[i] / nri;
        f2[i] = -(f / d1) * ri[i] / nri;
    }
    cross3(kj, n0, rk);
    double nrk = sqrt(dot3(rk, rk));
    for (int i = 0; i < 3; i++) {
        f2[i] = f2[i] - (f / d2) * rk[i] / nrk;
        f3[i] = (f / d2) * rk[i] / nrk;
    }
}

/* Force_Electrostatics, func.cpp:278-300 (Rcut 7 nm, +-0.9 clamp) */
static void pr_dh(double q1, double q2, double eps, double kappa, const double *p1,
                  const double *p2, double *f1, double *f2, double *E) {
    double d = dist3(p1, p2), f = 0.0;
    if (d < 7.0) {
        f = ((q1 * q2 * (kappa * d + 1)) / (4 * OR_PI * eps * d * d)) * exp(-kappa * d);
        *E += (q1 * q2 / (4 * OR_PI * eps * d)) * exp(-kappa * d);
    }
    for (int i = 0; i < 3; i++) {
        f1[i] = clamp09(f * (p1[i] - p2[i]) / d);
        f2[i] = clamp09(-f * (p1[i] - p2[i]) / d);
    }
}

/* Force_Exclude_Volume, func.cpp:318-340 (cut 4 nm, +-0.9 clamp) */
static void pr_ev(double kev, double sigma, const double *p1, const double *p2, double *f1,
                  double *f2, double *E) {
    double d = dist3(p1, p2), f = 0.0;
    if (d < 4.0) {
        f = -kev * ((6 * pow(sigma, 6)) / pow(d, 7) - (12 * pow(sigma, 12)) / pow(d, 13));
        *E += kev * (pow(sigma / d, 12) - pow(sigma / d, 6));
    }
    for (int i = 0; i < 3; i++) {
        f1[i] = clamp09(f * (p1[i] - p2[i]) / d);
        f2[i] = clamp09(-f * (p1[i] - p2[i]) / d);
    }
}

/* Force_Ele_Vdw, func.cpp:342-371 */
static void pr_dhev(double q1, double q2, double eps, double kappa, double kev, double sigma,
                    const double *p1, const double *p2, double *f1, double *f2, double *E) {
    double d = dist3(p1, p2), f = 0.0, fv = 0.0;
    if (d < 7.0) {
        f = ((q1 * q2 * (kappa * d + 1)) / (4 * OR_PI * eps * d * d)) * exp(-kappa * d);
        *E += (q1 * q2 / (4 * OR_PI * eps * d)) * exp(-kappa * d);
    }
    if (d < 4.0) {
        fv = -kev * ((6 * pow(sigma, 6)) / pow(d, 7) - (12 * pow(sigma, 12)) / pow(d, 13));
        *E += kev * (pow(sigma / d, 12) - pow(sigma / d, 6));
    }
    for (int i = 0; i < 3; i++) {
        f1[i] = clamp09((f + fv) * (p1[i] - p2[i]) / d);
        f2[i] = clamp09(-(f + fv) * (p1[i] - p2[i]) / d);
    }
}

/* ------------------------------------------------------------------ */
/* setup-type functions                                                 */
/* ------------------------------------------------------------------ */

/* twist_angle_phi, func.cpp:13-99 (table for w<=11, formula above) */
double or_twist_angle_phi(int w) {
    static const double table[12] = {NAN,     NAN,     0.9007, -0.6754, -0.3519, -0.1463,
                                     0.0215,  0.1178,  0.2025, 0.2726,  -0.2415, -0.15};
    if (w <= 11) {
        if (w < 2) return NAN; /* the reference exits the program here */
        return table[w];
    }
    double nbp = (w + 1) * 3 / 0.34;
    double turns = nbp / 10.3;
    double rot = round(turns);
    double per_bp = rot * 360 / nbp;
    double tw = per_bp - 34.95;
    double whole = nbp * tw;
    double seg = whole / (w + 1);
    return OR_PI * seg / 180.00;
}

/* update_phi_o, func.cpp:412-431 */
void or_update_phi_o(int n, const int *type, double *phi_o) {
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        int k = 1;
        while (i + k < n && type[i + k] == 0) k++;
        double ph = or_twist_angle_phi(k - 1);
        for (int l = 0; l < k; l++) phi_o[i + l] = ph;
    }
}

/* index of the last core at or before bead i (func.cpp:587-591) */
static int core_ordinal(const int *type, int i) {
    int ic = 0;
    for (int q = 0; q <= i; q++)
        if (type[q] == 1) ic++;
    return ic - 1;
}

static inline double clamp1(double v) {
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    return v;
}

/* update_Euler_Angle, func.cpp:433-685: pass 1 rebuilds linker frames and
 * segment lengths (448-505); pass 2 computes Euler angles and the core exit
 * frames (507-681). */
void or_update_euler(const or_system *sys, double *r, double *a, double *b, double *c,
                     double *alpha, double *beta, double *gamma, double *length,
                     double *a_dna, double *b_dna, double *c_dna, double *alpha_p,
                     double *beta_p, double *gamma_p) {
    const int n = sys->n;
    const int *type = sys->type;
    const double si = sin(OR_THETA), co = cos(OR_THETA);

    for (int i = 0; i < n - 1; i++) {
        double *ri = r + 3 * i, *rf = r + 3 * i + 3;
        double *ai = a + 3 * i, *bi = b + 3 * i, *ci = c + 3 * i;
        double *af = a + 3 * i + 3, *bf = b + 3 * i + 3, *cf = c + 3 * i + 3;
        double fw[3];
        if (type[i] == 0) {
            double a_old[3] = {ai[0], ai[1], ai[2]};
            if (type[i + 1] == 0) {
                for (int k = 0; k < 3; k++) fw[k] = rf[k] - ri[k];
            } else {
                for (int k = 0; k < 3; k++) {
                    double bm = -si * af[k] + co * bf[k];
                    fw[k] = rf[k] - OR_RO * bm + OR_D1 * cf[k] - ri[k];
                }
            }
            length[i] = sqrt(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
            double mi = 1.0 / length[i];
            for (int k = 0; k < 3; k++) ai[k] = mi * fw[k];
            double da[3] = {ai[0] - a_old[0], ai[1] - a_old[1], ai[2] - a_old[2]};
            double bda = bi[0] * da[0] + bi[1] * da[1] + bi[2] * da[2];
            for (int k = 0; k < 3; k++) bi[k] = bi[k] - bda * a_old[k];
            bda = bi[0] * ai[0] + bi[1] * ai[1] + bi[2] * ai[2];
            for (int k = 0; k < 3; k++) bi[k] = bi[k] - bda * ai[k];
            mi = 1.0 / sqrt(bi[0] * bi[0] + bi[1] * bi[1] + bi[2] * bi[2]);
            for (int k = 0; k < 3; k++) bi[k] = mi * bi[k];
            ci[0] = ai[1] * bi[2] - ai[2] * bi[1];
            ci[1] = ai[2] * bi[0] - ai[0] * bi[2];
            ci[2] = ai[0] * bi[1] - ai[1] * bi[0];
        } else {
            for (int k = 0; k < 3; k++) fw[k] = rf[k] - (ri[k] - OR_RO * bi[k] - OR_D1 * ci[k]);
            length[i] = sqrt(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
        }
    }

    for (int i = 0; i < n - 1; i++) {
        double *ri = r + 3 * i, *rf = r + 3 * i + 3;
        double *ai = a + 3 * i, *bi = b + 3 * i, *ci = c + 3 * i;
        double *af = a + 3 * i + 3, *bf = b + 3 * i + 3, *cf = c + 3 * i + 3;
        double ada, f1, f2, Ac, apg, sb;
        if (type[i] == 0) {
            /* forward frame: next linker's (a,b) or the next core's entry frame */
            double am[3], bm[3];
            if (type[i + 1] == 0) {
                for (int k = 0; k < 3; k++) { am[k] = af[k]; bm[k] = bf[k]; }
            } else {
                for (int k = 0; k < 3; k++) {
                    am[k] = co * af[k] + si * bf[k];
                    bm[k] = -si * af[k] + co * bf[k];
                }
            }
            ada = clamp1(dot3(ai, am));
            beta[i] = acos(ada);
            sb = sin(beta[i]);
            if (beta[i] > 1e-10) f1 = dot3(am, bi) / sb;
            else f1 = dot3(bm, bi);
            f1 = clamp1(f1);
            Ac = acos(f1);
            f2 = dot3(am, ci);
            alpha[i] = (f2 >= 0) ? Ac : -Ac;
            f1 = dot6(bi, bm, ci, cf) / (1.0 + ada);
            f1 = clamp1(f1);
            apg = acos(f1);
            f2 = (dot3(ci, bm) - dot3(bi, cf)) / (1.0 + ada);
            gamma[i] = (f2 >= 0.0) ? apg - alpha[i] : -apg - alpha[i];
        } else {
            int ic = core_ordinal(type, i);
            double *ad = a_dna + 3 * ic, *bd = b_dna + 3 * ic, *cd = c_dna + 3 * ic;
            for (int k = 0; k < 3; k++)
                ad[k] = (rf[k] - (ri[k] - OR_RO * bi[k] - OR_D1 * ci[k])) / length[i];
            double cb = clamp1(dot3(ai, ad));
            beta_p[ic] = acos(cb);
            sb = sin(beta_p[ic]);
            if (beta_p[ic] >= 1e-10) {
                double bm[3];
                for (int k = 0; k < 3; k++) bm[k] = (ad[k] - cb * ai[k]) / sb;
                double ca = clamp1(dot3(bm, bi));
                Ac = acos(ca);
                f1 = dot3(ad, ci);
                alpha_p[ic] = (f1 >= 0) ? Ac : -Ac;
                gamma_p[ic] = -alpha_p[ic];
                double sa = sin(alpha_p[ic]);
                double sg = sin(gamma_p[ic]);
                double cg = cos(gamma_p[ic]);
                double R21 = -cg * sb, R22 = cg * cb * ca - sg * sa, R23 = cg * cb * sa + sg * ca;
                for (int k = 0; k < 3; k++) bd[k] = R21 * ai[k] + R22 * bi[k] + R23 * ci[k];
                double R31 = sg * sb, R32 = -sg * cb * ca - cg * sa, R33 = -sg * cb * sa + cg * ca;
                for (int k = 0; k < 3; k++) cd[k] = R31 * ai[k] + R32 * bi[k] + R33 * ci[k];
            } else {
                for (int k = 0; k < 3; k++) { bd[k] = bi[k]; cd[k] = ci[k]; }
            }
            ada = clamp1(dot3(ad, af));
            beta[i] = acos(ada);
            sb = sin(beta[i]);
            if (beta[i] >= 1e-10) f1 = dot3(af, bd) / sb;
            else f1 = dot3(bf, bd);
            f1 = clamp1(f1);
            Ac = acos(f1);
            f2 = dot3(af, cd);
            alpha[i] = (f2 >= 0) ? Ac : -Ac;
            f1 = clamp1(dot6(bd, bf, cd, cf) / (1.0 + ada));
            apg = acos(f1);
            f2 = (dot3(cd, bf) - dot3(bd, cf)) / (1.0 + ada);
            gamma[i] = (f2 >= 0.0) ? apg - alpha[i] : -apg - alpha[i];
        }
        (void)ri;
    }
}

/* update_tail_beta, func.cpp:1763-1790 (bounds test moved first; the
 * reference reads t_grp[i+1], t_grp[i+2] past the end before testing i) */
void or_update_tail_beta(int n_tail, const double *r_t, double *beta_t, const int *t_grp) {
    for (int i = 0; i < n_tail; i++) {
        if (i < n_tail - 2 && t_grp[i] == t_grp[i + 1] && t_grp[i] == t_grp[i + 2]) {
            const double *p1 = r_t + 3 * i, *p2 = p1 + 3, *p3 = p1 + 6;
            double x12 = p1[0] - p2[0], y12 = p1[1] - p2[1], z12 = p1[2] - p2[2];
            double r12 = sqrt(x12 * x12 + y12 * y12 + z12 * z12);
            double x32 = p3[0] - p2[0], y32 = p3[1] - p2[1], z32 = p3[2] - p2[2];
            double r32 = sqrt(x32 * x32 + y32 * y32 + z32 * z32);
            beta_t[i] = acos((x12 * x32 + y12 * y32 + z12 * z32) / (r12 * r32));
        } else {
            beta_t[i] = 0.0;
        }
    }
}

/* update_LH_beta, func.cpp:2062-2089 (chain position from i % n_lh_c) */
void or_update_lh_beta(int n_LH, int n_lh_c, const double *r_lh, double *beta_lh) {
    for (int i = 0; i < n_LH; i++) {
        if ((i % n_lh_c != 0) && (i % n_lh_c != n_lh_c - 1)) {
            const double *p1 = r_lh + 3 * i - 3, *p2 = r_lh + 3 * i, *p3 = r_lh + 3 * i + 3;
            double x12 = p1[0] - p2[0], y12 = p1[1] - p2[1], z12 = p1[2] - p2[2];
            double r12 = sqrt(x12 * x12 + y12 * y12 + z12 * z12);
            double x32 = p3[0] - p2[0], y32 = p3[1] - p2[1], z32 = p3[2] - p2[2];
            double r32 = sqrt(x32 * x32 + y32 * y32 + z32 * z32);
            beta_lh[i] = acos((x12 * x32 + y12 * y32 + z12 * z32) / (r12 * r32));
        } else {
            beta_lh[i] = 0.0;
        }
    }
}

/* build_tail, func.cpp:1672-1713 (every core carries tails: nc_t_flag = 1,
 * main.cpp:200-204) */
void or_build_tail(const or_system *sys, const double *tail_q, const int *tail_grp,
                   const double *tail_rad, const double *bond_c, const double *bond_v,
                   const double *angle_c, const double *angle_v, const double *r,
                   const double *a, const double *b, const double *c, double *r_t,
                   double *h_t, double *g_t, double *lo_t, double *beta_o_t, double *t_q,
                   double *t_rad, int *t_grp, int *t_fix) {
    const int n_t = sys->n_t;
    int n_grp = 1;
    for (int i = 1; i < n_t; i++)
        if (tail_grp[i] != tail_grp[i - 1]) n_grp++;
    int k = 0;
    for (int i = 0; i < sys->n; i++) {
        if (sys->type[i] != 1) continue;
        for (int j = 0; j < n_t; j++) {
            int q = k * n_t + j;
            body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, sys->tail_pos + 3 * j,
                          r_t + 3 * q);
            t_q[q] = tail_q[j];
            t_rad[q] = tail_rad[j];
            t_fix[q] = sys->tail_fix_tpl[j];
            t_grp[q] = tail_grp[j] + k * n_grp;
            h_t[q] = bond_c[j];
            lo_t[q] = bond_v[j];
            g_t[q] = angle_c[j];
            beta_o_t[q] = angle_v[j];
        }
        k++;
    }
}

/* build_LH, func.cpp:2029-2059 (C-term positions indexed by the chain
 * position j as in the reference) */
void or_build_lh(int n, const int *type, int n_lh_n, int n_lh_c, const double *LH_n_pos,
                 const double *LH_c_pos, const double *LH_radius, const int *nc_lh_flag,
                 const double *r, const double *a, const double *b, const double *c,
                 double *r_lh, double *lh_rad) {
    int nlh = n_lh_n + n_lh_c, k = 0, k1 = 0;
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        if (nc_lh_flag[k1] == 1) {
            for (int j = 0; j < nlh; j++) {
                const double *src = (j < n_lh_n) ? LH_n_pos + 3 * j : LH_c_pos + 3 * j;
                body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, src,
                              r_lh + 3 * (k * nlh + j));
                lh_rad[k * nlh + j] = LH_radius[j];
            }
            k++;
        }
        k1++;
    }
}

/* ------------------------------------------------------------------ */
/* RPY diffusion tensor and Cholesky, func.cpp:728-840                  */
/* ------------------------------------------------------------------ */
void or_diffusion_tensor(int nb, const double *r, const double *rad, double a1, double a2,
                         double *D) {
    const int m = 3 * nb;
    for (int i = 0; i < nb; i++) {
        for (int j = i; j < nb; j++) {
            if (i == j) {
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++)
                        D[(3 * i + k) * m + 3 * j + l] = (k == l) ? a1 / rad[i] : 0.0;
                continue;
            }
            double sij[3] = {r[3 * i] - r[3 * j], r[3 * i + 1] - r[3 * j + 1],
                             r[3 * i + 2] - r[3 * j + 2]};
            double ssq = sij[0] * sij[0] + sij[1] * sij[1] + sij[2] * sij[2];
            double s = sqrt(ssq);
            for (int k = 0; k < 3; k++) {
                for (int l = 0; l < 3; l++) {
                    double v;
                    if (s >= rad[i] + rad[j]) {
                        /* far field (func.cpp:759-779) */
                        double f = (rad[i] * rad[i] + rad[j] * rad[j]) / ssq;
                        double f1 = 1.0 + f / 3, f2 = 1.0 - f;
                        v = (k == l) ? (a2 / s) * (f1 + f2 * sij[k] * sij[k] / ssq)
                                     : (a2 / s) * (f2 * sij[k] * sij[l] / ssq);
                    } else {
                        /* overlap: effective radius (func.cpp:780-798) */
                        double se = pow((rad[i] * rad[i] * rad[i] + rad[j] * rad[j] * rad[j]) / 2.0,
                                        1.0 / 3);
                        v = (k == l) ? (a1 / se) * (1.0 - 9.0 * sqrt(ssq) / (32 * se) +
                                                    3.0 * sij[k] * sij[k] / (32.0 * se * sqrt(ssq)))
                                     : (a1 / se) * (3.0 * sij[k] * sij[l] / (32.0 * se * sqrt(ssq)));
                    }
                    D[(3 * i + k) * m + 3 * j + l] = v;
                    D[(3 * j + l) * m + 3 * i + k] = v;
                }
            }
        }
    }
}

/* Row-oriented Cholesky-Crout, func.cpp:806-836: a negative pivot f is
 * replaced by -f (the reference also appends to wrongD.txt). */
int or_cholesky(int m, const double *D, double *L) {
    int neg = 0;
    memset(L, 0, sizeof(double) * (size_t)m * m);
    L[0] = sqrt(D[0]);
    for (int i = 1; i < m; i++) {
        double *Li = L + (size_t)i * m;
        const double *Di = D + (size_t)i * m;
        Li[0] = Di[0] / L[0];
        for (int j = 1; j < i; j++) {
            const double *Lj = L + (size_t)j * m;
            double ssq = 0.0;
            for (int k = 0; k < j; k++) ssq += Li[k] * Lj[k];
            Li[j] = (Di[j] - ssq) / Lj[j];
        }
        double ssq = 0.0;
        for (int k = 0; k < i; k++) ssq += Li[k] * Li[k];
        double f = Di[i] - ssq;
        if (f < 0) { f = -f; neg++; }
        Li[i] = sqrt(f);
    }
    return neg;
}

/* ------------------------------------------------------------------ */
/* linker-DNA mechanics, func.cpp:951-1289                              */
/* ------------------------------------------------------------------ */

/* first_coord / second_coord, func.cpp:687-723 */
static void entry_point(int t, const double *r, const double *a, const double *b,
                        const double *c, double *out) {
    if (t == 0) { out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; return; }
    const double si = sin(OR_THETA), co = cos(OR_THETA);
    for (int k = 0; k < 3; k++) {
        double bm = -si * a[k] + co * b[k];
        out[k] = r[k] - OR_RO * bm + OR_D1 * c[k];
    }
}
static void exit_point(int t, const double *r, const double *b, const double *c,
                       double *out) {
    if (t == 0) { out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; return; }
    for (int k = 0; k < 3; k++) out[k] = r[k] - (OR_RO * b[k] + OR_D1 * c[k]);
}

/* add a segment force pair (f1 on bead i, f2 on bead i+1) and the lever-arm
 * torques about each bead centre (func.cpp:998-1016) */
static void seg_apply(const or_state *st, int i, const double *rs, const double *rf,
                      const double *f1, const double *f2, double *F, double *T) {
    add3(F + 3 * i, f1);
    add3(F + 3 * i + 3, f2);
    double t1[3], t2[3];
    torque_abs(f1, rs, st->r + 3 * i, st->a + 3 * i, st->b + 3 * i, st->c + 3 * i, t1);
    torque_abs(f2, rf, st->r + 3 * i + 3, st->a + 3 * i + 3, st->b + 3 * i + 3,
               st->c + 3 * i + 3, t2);
    add3(T + 3 * i, t1);
    add3(T + 3 * i + 3, t2);
}

void or_mech(const or_system *sys, const or_state *st, double *F, double *T, double *E) {
    const int n = sys->n, n3 = 3 * n;
    const int *type = sys->type;
    const double h = sys->h, g = sys->g, s = sys->s;
    const double *r = st->r, *a = st->a, *b = st->b, *c = st->c;
    const double *alpha = st->alpha, *beta = st->beta, *gamma = st->gamma;
    const double *phi_o = sys->phi_o;
    const double si = sin(OR_THETA), co = cos(OR_THETA);
    double rf[3], rs[3], f1[3], f2[3];

    /* stretching, func.cpp:978-1018 */
    for (int i = 0; i < n - 1; i++) {
        entry_point(type[i + 1], r + 3 * i + 3, a + 3 * i + 3, b + 3 * i + 3, c + 3 * i + 3, rf);
        exit_point(type[i], r + 3 * i, b + 3 * i, c + 3 * i, rs);
        pr_stretch(h, rs, rf, OR_LO, f1, f2, E);
        seg_apply(st, i, rs, rf, f1, f2, F, T);
    }

    /* bending, func.cpp:1022-1105 */
    for (int i = 0; i < n - 1; i++) {
        int ic = core_ordinal(type, i);
        entry_point(type[i + 1], r + 3 * i + 3, a + 3 * i + 3, b + 3 * i + 3, c + 3 * i + 3, rf);
        exit_point(type[i], r + 3 * i, b + 3 * i, c + 3 * i, rs);
        double af[3], ab[3], ao[3], beta_b;
        if (type[i] == 0) {
            for (int k = 0; k < 3; k++)
                af[k] = (type[i + 1] == 0) ? a[3 * i + 3 + k]
                                           : co * a[3 * i + 3 + k] + si * b[3 * i + 3 + k];
            for (int k = 0; k < 3; k++)
                ab[k] = (type[i - 1] == 0) ? a[3 * i - 3 + k] : st->a_dna[3 * ic + k];
            for (int k = 0; k < 3; k++) ao[k] = a[3 * i + k];
            beta_b = beta[i - 1];
        } else {
            for (int k = 0; k < 3; k++) {
                af[k] = a[3 * i + 3 + k];
                ab[k] = a[3 * i + k];
                ao[k] = st->a_dna[3 * ic + k];
            }
            beta_b = st->beta_p[ic];
        }
        /* Bending_force_projection, func.cpp:198-232 */
        double bi = beta[i], len = st->length[i];
        *E += g * (bi - beta_b) * (bi - beta_b) / 2;
        double g1 = (bi >= 1e-10) ? bi / (sin(bi) * len) : 1.0 / len;
        double c1 = cos(bi);
        double g2 = (beta_b >= 1e-10) ? beta_b / (sin(beta_b) * len) : 1.0 / len;
        double c2 = cos(beta_b);
        for (int k = 0; k < 3; k++) {
            double Ai = g1 * (af[k] - c1 * ao[k]);
            double Bi = g2 * (ab[k] - c2 * ao[k]);
            f1[k] = -g * (Ai + Bi);
            f2[k] = g * (Ai + Bi);
        }
        seg_apply(st, i, rs, rf, f1, f2, F, T);
    }

    /* twisting, func.cpp:1110-1179 */
    for (int i = 0; i < n - 1; i++) {
        int ic = core_ordinal(type, i);
        entry_point(type[i + 1], r + 3 * i + 3, a + 3 * i + 3, b + 3 * i + 3, c + 3 * i + 3, rf);
        exit_point(type[i], r + 3 * i, b + 3 * i, c + 3 * i, rs);
        double al_b, be_b, ga_b, ph_b, ga_n;
        const double *bo, *cc;
        if (type[i] == 0) {
            al_b = alpha[i - 1]; be_b = beta[i - 1]; ga_b = gamma[i - 1];
            ph_b = phi_o[i - 1]; ga_n = gamma[i - 1];
            bo = b + 3 * i; cc = c + 3 * i;
        } else {
            al_b = alpha[i]; be_b = st->beta_p[ic]; ga_b = gamma[i];
            ph_b = phi_o[i]; ga_n = st->gamma_p[ic];
            bo = st->b_dna + 3 * ic; cc = st->c_dna + 3 * ic;
        }
        /* Twisting_force_projection, func.cpp:234-260 */
        double tw = alpha[i] + gamma[i] - phi_o[i];
        *E += s * (tw * tw) / 2;
        double g1 = tw * tan(0.5 * beta[i]) / st->length[i];
        double ca = cos(alpha[i]), sa = sin(alpha[i]);
        double g2 = (al_b + ga_b - ph_b) * tan(0.5 * be_b) / st->length[i];
        double cg = cos(ga_n), sg = sin(ga_n);
        for (int k = 0; k < 3; k++) {
            double Chi = g1 * (ca * cc[k] - sa * bo[k]);
            double Zhi = g2 * (cg * cc[k] + sg * bo[k]);
            f1[k] = s * (Chi + Zhi);
            f2[k] = -s * (Chi + Zhi);
        }
        seg_apply(st, i, rs, rf, f1, f2, F, T);
    }

    /* explicit mechanical torques, func.cpp:1185-1277 (linker beads assign) */
    for (int i = 0; i < n - 1; i++) {
        int im1 = i - 1, ic = core_ordinal(type, i);
        double *Ti = T + 3 * i;
        const double *ai = a + 3 * i, *bi = b + 3 * i, *ci = c + 3 * i;
        if (type[i] == 0) {
            Ti[0] = s * (alpha[i] + gamma[i] + phi_o[i] - alpha[im1] - gamma[im1] - phi_o[im1]);
            Ti[1] = 0.0;
            Ti[2] = 0.0;
            continue;
        }
        const double *ad = st->a_dna + 3 * ic, *bd = st->b_dna + 3 * ic, *cd = st->c_dna + 3 * ic;
        double ada = dot3(ad, ai), adb = dot3(ad, bi), adc = dot3(ad, ci);
        double mag = s * (alpha[i] + gamma[i] - phi_o[i]);
        Ti[0] += mag * ada;
        Ti[1] += mag * adb;
        Ti[2] += mag * adc;
        if (i > 0) {
            mag = -s * (alpha[im1] + gamma[im1] - phi_o[im1]);
            Ti[0] += mag * co;
            Ti[1] += mag * si;
            Ti[2] += 0.0;
        }
        /* extra bending torque */
        double bp = st->beta_p[ic];
        Ti[1] = Ti[1] - g * bp * adc / sin(bp);
        Ti[2] = Ti[2] + g * bp * adb / sin(bp);
        if (i > 0) {
            const double *ab = a + 3 * im1;
            double xa = dot3(ab, ai), xb = dot3(ab, bi), xc = dot3(ab, ci);
            double bm = beta[im1];
            Ti[0] = Ti[0] + g * bm * (si * xc) / sin(bm);
            Ti[1] = Ti[1] - g * bm * (co * xc) / sin(bm);
            Ti[2] = Ti[2] + g * bm * (co * xb - si * xa) / sin(bm);
        }
        /* extra twisting torque */
        double s1 = sin(st->alpha_p[ic]), c1 = cos(st->alpha_p[ic]);
        mag = s * (alpha[i] + gamma[i] - phi_o[i]) * tan(0.5 * bp);
        double cda = dot3(cd, ai), bda = dot3(bd, ai), cdb = dot3(cd, bi), bdb = dot3(bd, bi);
        double cdc = dot3(cd, ci), bdc = dot3(bd, ci);
        Ti[0] = Ti[0] - mag * (s1 * cda + c1 * bda);
        Ti[1] = Ti[1] - mag * (s1 * cdb + c1 * bdb);
        Ti[2] = Ti[2] - mag * (s1 * cdc + c1 * bdc);
        if (i > 0) {
            s1 = sin(gamma[im1]);
            c1 = cos(gamma[im1]);
            mag = s * (alpha[im1] + gamma[im1] - phi_o[im1]) * tan(0.5 * beta[im1]);
            /* the previous segment's frame expressed in this core's entry frame */
            Ti[0] = Ti[0] - mag * (s1 * 0.0 - c1 * (-si));
            Ti[1] = Ti[1] - mag * (s1 * 0.0 - c1 * co);
            Ti[2] = Ti[2] - mag * (s1 * 1.0 - c1 * 0.0);
        }
    }

    /* last bead: assigned, func.cpp:1281-1287 */
    T[n3 - 3] = -s * (alpha[n - 2] + gamma[n - 2] - phi_o[n - 2]);
    T[n3 - 2] = 0.0;
    T[n3 - 1] = 0.0;
}

/* ------------------------------------------------------------------ */
/* core / linker nonbonded, func.cpp:1291-1605                          */
/* ------------------------------------------------------------------ */

/* body-frame torque of force f applied at body point p on a core frame */
static void core_torque_add(double *T, const double *a, const double *b, const double *c,
                            const double *f, const double *p) {
    double fa = dot3(a, f), fb = dot3(b, f), fc = dot3(c, f);
    T[0] += fc * p[1] - fb * p[2];
    T[1] += fa * p[2] - fc * p[0];
    T[2] += fb * p[0] - fa * p[1];
}

void or_elec(const or_system *sys, const or_state *st, double *F, double *T, double *E) {
    const int n = sys->n, Nq = sys->Nq;
    const int *type = sys->type;
    const double *r = st->r, *a = st->a, *b = st->b, *c = st->c;
    const double *cp = sys->core_pos, *cq = sys->core_q;
    const double ke = sys->k_e, kex = sys->k_ex, ql = sys->q_l;
    const double kll = sys->debyell, kap = sys->debye;
    const double qlql = ql * ql;

    for (int i = 0; i < n - 1; i++) {
        for (int j = i + 1; j < n; j++) {
            const double *ri = r + 3 * i, *rj = r + 3 * j;
            double dist = dist3(rj, ri);
            /* the `ch` filter, func.cpp:1316-1323 */
            int ch = 1;
            if (dist > 25.0) ch = 0;
            if (type[i + 1] != 0) ch = 0;
            if (i > 0 && type[i - 1] != 0) ch = 0;
            if (j < n - 1 && type[j + 1] != 0) ch = 0;
            if (type[j - 1] != 0) ch = 0;
            if (!ch) continue;

            if (type[i] == 0 && type[j] == 0) {
                /* DNA-DNA, func.cpp:1328-1354 */
                if (abs(i - j) <= 1) continue;
                double mi = 1.0 / dist;
                double z[3] = {mi * (ri[0] - rj[0]), mi * (ri[1] - rj[1]), mi * (ri[2] - rj[2])};
                double g1 = ke * qlql * exp(-kll * dist) * (kll * dist + 1.0) / (dist * dist);
                *E += (ke * qlql / dist) * exp(-kll * dist);
                if (dist <= 8) {
                    double s1 = sys->sig_dd;
                    g1 += kex * ((12.0 / s1) * pow(s1 / dist, 13) - (6.0 / s1) * pow(s1 / dist, 7));
                    *E += kex * (pow(s1 / dist, 12) - pow(s1 / dist, 6));
                }
                for (int k = 0; k < 3; k++) {
                    F[3 * i + k] += g1 * z[k];
                    F[3 * j + k] -= g1 * z[k];
                }
            } else if (type[i] == 0 || type[j] == 0) {
                /* DNA-core (func.cpp:1356-1425) and core-DNA (1431-1487):
                 * z points from the charge towards the linker bead (DNA i) or
                 * from the linker bead towards the charge (core i). */
                int lk = (type[i] == 0) ? i : j, co = (type[i] == 0) ? j : i;
                double sgn = (type[i] == 0) ? 1.0 : -1.0;
                const double *rl = r + 3 * lk, *rc = r + 3 * co;
                const double *ac = a + 3 * co, *bc = b + 3 * co, *cc = c + 3 * co;
                for (int k = 0; k < Nq; k++) {
                    double q[3];
                    body_to_world(rc, ac, bc, cc, cp + 3 * k, q);
                    double z[3];
                    if (sgn > 0) { z[0] = rl[0] - q[0]; z[1] = rl[1] - q[1]; z[2] = rl[2] - q[2]; }
                    else         { z[0] = q[0] - rl[0]; z[1] = q[1] - rl[1]; z[2] = q[2] - rl[2]; }
                    double d = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
                    double mi = 1.0 / d;
                    z[0] *= mi; z[1] *= mi; z[2] *= mi;
                    if (abs(i - j) > 1) {
                        double g1 = ke * ql * cq[k] * exp(-kap * d) * (kap * d + 1.0) / (d * d);
                        *E += (ke * ql * cq[k] / d) * exp(-kap * d);
                        double fi[3] = {g1 * z[0], g1 * z[1], g1 * z[2]};
                        add3(F + 3 * i, fi);
                        sub3(F + 3 * j, fi);
                        /* force on the core is +fi (core i) or -fi (core j) */
                        double fc[3] = {sgn * -fi[0], sgn * -fi[1], sgn * -fi[2]};
                        core_torque_add(T + 3 * co, ac, bc, cc, fc, cp + 3 * k);
                    }
                    if (d <= 8.0 && cq[k] > 0) {
                        double s1 = sys->sig_dc;
                        double g1 = kex * ((12.0 / s1) * pow(s1 / d, 13) - (6.0 / s1) * pow(s1 / d, 7));
                        *E += kex * (pow(s1 / d, 12) - pow(s1 / d, 6));
                        double fi[3] = {g1 * z[0], g1 * z[1], g1 * z[2]};
                        add3(F + 3 * i, fi);
                        sub3(F + 3 * j, fi);
                        double fc[3] = {sgn * -fi[0], sgn * -fi[1], sgn * -fi[2]};
                        core_torque_add(T + 3 * co, ac, bc, cc, fc, cp + 3 * k);
                    }
                }
            } else {
                /* core-core over all charge pairs, func.cpp:1489-1566 */
                const double *ai = a + 3 * i, *bi = b + 3 * i, *ci = c + 3 * i;
                const double *aj = a + 3 * j, *bj = b + 3 * j, *cj = c + 3 * j;
                for (int k = 0; k < Nq; k++) {
                    double qk[3];
                    body_to_world(ri, ai, bi, ci, cp + 3 * k, qk);
                    for (int l = 0; l < Nq; l++) {
                        double ql3[3];
                        body_to_world(rj, aj, bj, cj, cp + 3 * l, ql3);
                        double z[3] = {qk[0] - ql3[0], qk[1] - ql3[1], qk[2] - ql3[2]};
                        double d = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
                        double mi = 1.0 / d;
                        z[0] *= mi; z[1] *= mi; z[2] *= mi;
                        double g1 = ke * cq[k] * cq[l] * exp(-kap * d) * (kap * d + 1.0) / (d * d);
                        *E += (ke * cq[l] * cq[k] / d) * exp(-kap * d);
                        double fi[3] = {g1 * z[0], g1 * z[1], g1 * z[2]};
                        double fj[3] = {-fi[0], -fi[1], -fi[2]};
                        add3(F + 3 * i, fi);
                        sub3(F + 3 * j, fi);
                        core_torque_add(T + 3 * i, ai, bi, ci, fi, cp + 3 * k);
                        core_torque_add(T + 3 * j, aj, bj, cj, fj, cp + 3 * l);
                        if (d <= 8.0) {
                            double s1 = sys->sig_cc;
                            g1 = kex * ((12.0 / s1) * pow(s1 / d, 13) - (6.0 / s1) * pow(s1 / d, 7));
                            *E += kex * (pow(s1 / d, 12) - pow(s1 / d, 6));
                            double gi[3] = {g1 * z[0], g1 * z[1], g1 * z[2]};
                            double gj[3] = {-gi[0], -gi[1], -gi[2]};
                            add3(F + 3 * i, gi);
                            sub3(F + 3 * j, gi);
                            core_torque_add(T + 3 * i, ai, bi, ci, gi, cp + 3 * k);
                            core_torque_add(T + 3 * j, aj, bj, cj, gj, cp + 3 * l);
                        }
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* histone tails, func.cpp:1793-2026                                    */
/* ------------------------------------------------------------------ */
void or_tail(const or_system *sys, const or_state *st, double *F, double *T, double *tF,
             double *E) {
    const int n = sys->n, n_tail = sys->n_tail, n_t = sys->n_t;
    const int *tg = sys->t_grp;
    const double *rt = st->r_t;
    const double eps = 1 / (4 * OR_PI * sys->k_e);
    double f1[3], f2[3], f3[3];

    /* stretching, 1811-1827 */
    for (int i = 0; i < n_tail - 1; i++) {
        if (tg[i] != tg[i + 1]) continue;
        pr_stretch(sys->h_t[i], rt + 3 * i, rt + 3 * i + 3, sys->lo_t[i], f1, f2, E);
        add3(tF + 3 * i, f1);
        add3(tF + 3 * i + 3, f2);
    }
    /* bending, 1832-1858 */
    for (int i = 0; i < n_tail - 2; i++) {
        if (!(tg[i] == tg[i + 1] && tg[i] == tg[i + 2])) continue;
        pr_bend(sys->g_t[i], st->beta_t[i], sys->beta_o_t[i], rt + 3 * i, rt + 3 * i + 3,
                rt + 3 * i + 6, f1, f2, f3, E);
        sub3(tF + 3 * i, f1);
        sub3(tF + 3 * i + 3, f2);
        sub3(tF + 3 * i + 6, f3);
    }
    /* tail-tail Debye-Hueckel only, 1863-1891 */
    for (int i = 0; i < n_tail - 1; i++) {
        for (int j = i + 1; j < n_tail; j++) {
            if (!(tg[i] != tg[j] || j > i + 2)) continue;
            pr_dh(sys->t_q[i], sys->t_q[j], eps, sys->debye, rt + 3 * i, rt + 3 * j, f1, f2, E);
            add3(tF + 3 * i, f1);
            add3(tF + 3 * j, f2);
        }
    }
    /* tail-linker DH + EV, 1898-1917 */
    for (int i = 0; i < n_tail; i++) {
        for (int j = 0; j < n; j++) {
            if (sys->type[j] != 0) continue;
            pr_dhev(sys->t_q[i], sys->q_l, eps, sys->debye, sys->k_ex, sys->sig_tl, rt + 3 * i,
                    st->r + 3 * j, f1, f2, E);
            add3(tF + 3 * i, f1);
            add3(F + 3 * j, f2);
        }
    }
    /* anchor springs between fixed tail beads and their core, 1971-2022 */
    int k = 0;
    for (int i = 0; i < n; i++) {
        if (sys->type[i] != 1) continue;
        const double *ri = st->r + 3 * i, *ai = st->a + 3 * i, *bi = st->b + 3 * i,
                     *ci = st->c + 3 * i;
        for (int j = 0; j < n_t; j++) {
            if (sys->tail_fix_tpl[j] != 1) continue;
            double ro[3];
            body_to_world(ri, ai, bi, ci, sys->tail_pos + 3 * j, ro);
            const double *p = rt + 3 * (k * n_t + j);
            double S[3] = {p[0] - ro[0], p[1] - ro[1], p[2] - ro[2]};
            *E += sys->h * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2]) / 2;
            double hs[3] = {sys->h * S[0], sys->h * S[1], sys->h * S[2]};
            sub3(tF + 3 * (k * n_t + j), hs);
            add3(F + 3 * i, hs);
            double tq[3];
            torque_rel(hs, sys->tail_pos + 3 * j, ai, bi, ci, tq);
            add3(T + 3 * i, tq);
        }
        k++;
    }
}

/* ------------------------------------------------------------------ */
/* linker histone, func.cpp:2093-2781                                   */
/* ------------------------------------------------------------------ */

/* template index of free LH bead i (N-term first, then C-term after G) */
static inline int lh_tpl(const or_system *sys, int i) {
    int idx = i % (sys->n_lh_n + sys->n_lh_c);
    return (idx < sys->n_lh_n) ? idx : idx + sys->n_lh_g;
}

/* core bead carrying tail bead j: the ((t_grp-1)/groups_per_core)-th core
 * (func.cpp:2366-2375) */
static int tail_core_bead(const or_system *sys, int j) {
    int target = (sys->t_grp[j] - 1) / sys->groups_per_core, k = 0;
    for (int q = 0; q < sys->n; q++) {
        if (sys->type[q] != 1) continue;
        if (k == target) return q;
        k++;
    }
    return -1;
}

/* force f2 acting on tail bead j: on the bead itself, or moved to its core
 * (with the lever-arm torque) for anchored beads (func.cpp:2361-2391) */
static void tail_receive(const or_system *sys, const or_state *st, int j, const double *f2,
                         double *F, double *T, double *tF) {
    if (sys->t_fix[j] == 0) {
        add3(tF + 3 * j, f2);
        return;
    }
    int ic = tail_core_bead(sys, j);
    double t2[3];
    torque_abs(f2, st->r_t + 3 * j, st->r + 3 * ic, st->a + 3 * ic, st->b + 3 * ic,
               st->c + 3 * ic, t2);
    add3(F + 3 * ic, f2);
    add3(T + 3 * ic, t2);
}

void or_lh(const or_system *sys, const or_state *st, double *F, double *T, double *tF,
           double *lhF, double *E) {
    const int n = sys->n, n_LH = sys->n_LH, n_tail = sys->n_tail;
    const int nlh = sys->n_lh_n + sys->n_lh_c, n_lh_g = sys->n_lh_g, n_lh_c = sys->n_lh_c;
    const int *type = sys->type;
    const double *rl = st->r_lh, *r = st->r, *a = st->a, *b = st->b, *c = st->c;
    const double eps = 1 / (4 * OR_PI * sys->k_e);
    const double kap = sys->debye, kex = sys->k_ex;
    double f1[3], f2[3], f3[3];

    /* stretching, 2111-2143 */
    for (int i = 0; i < n_LH; i++) {
        int idx = lh_tpl(sys, i);
        if (sys->LH_conn[idx] != 1) continue;
        pr_stretch(sys->LH_kstr[idx], rl + 3 * i, rl + 3 * i + 3, sys->LH_streq[idx], f1, f2, E);
        add3(lhF + 3 * i, f1);
        add3(lhF + 3 * i + 3, f2);
    }
    /* bending, 2149-2181 */
    for (int i = 0; i < n_LH; i++) {
        int idx = lh_tpl(sys, i);
        if (!((i % n_lh_c != 0) && (i % n_lh_c != n_lh_c - 1))) continue;
        pr_bend(sys->LH_kben[idx], st->beta_lh[i], sys->LH_betaeq[idx], rl + 3 * i - 3,
                rl + 3 * i, rl + 3 * i + 3, f1, f2, f3, E);
        sub3(lhF + 3 * i - 3, f1);
        sub3(lhF + 3 * i, f2);
        sub3(lhF + 3 * i + 3, f3);
    }
    /* LH-LH: same LH excluded volume only (j > i+1), else DH + EV, 2190-2244 */
    for (int i = 0; i < n_LH; i++) {
        int i1 = lh_tpl(sys, i);
        for (int j = i; j < n_LH; j++) {
            int i2 = lh_tpl(sys, j);
            double sg = (sys->LH_vdw_hh[i1] + sys->LH_vdw_hh[i2]) / 2;
            if (i / nlh == j / nlh) {
                if (!(j > i + 1)) continue;
                pr_ev(kex, sg, rl + 3 * i, rl + 3 * j, f1, f2, E);
            } else {
                pr_dhev(sys->LH_q[i1], sys->LH_q[i2], eps, kap, kex, sg, rl + 3 * i, rl + 3 * j,
                        f1, f2, E);
            }
            add3(lhF + 3 * i, f1);
            add3(lhF + 3 * j, f2);
        }
    }
    /* LH-linker DNA, 2249-2279 */
    for (int i = 0; i < n_LH; i++) {
        int idx = lh_tpl(sys, i);
        for (int j = 0; j < n; j++) {
            if (type[j] != 0) continue;
            pr_dhev(sys->LH_q[idx], sys->q_l, eps, kap, kex, sys->LH_vdw_hl[idx], rl + 3 * i,
                    r + 3 * j, f1, f2, E);
            add3(lhF + 3 * i, f1);
            add3(F + 3 * j, f2);
        }
    }
    /* LH-tail, 2343-2398 */
    for (int i = 0; i < n_LH; i++) {
        int idx = lh_tpl(sys, i);
        for (int j = 0; j < n_tail; j++) {
            pr_dhev(sys->LH_q[idx], sys->t_q[j], eps, kap, kex, sys->LH_vdw_ht[idx], rl + 3 * i,
                    st->r_t + 3 * j, f1, f2, E);
            add3(lhF + 3 * i, f1);
            tail_receive(sys, st, j, f2, F, T, tF);
        }
    }
    /* globular head sites (rigid with their core) */
    /* LHG-tail, 2403-2476 */
    int k2 = 0;
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        if (sys->nc_lh_flag[k2] == 1) {
            for (int m = 0; m < n_lh_g; m++) {
                double g[3];
                body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, sys->LH_g_pos + 3 * m, g);
                for (int j = 0; j < n_tail; j++) {
                    pr_dhev(sys->LH_q[m], sys->t_q[j], eps, kap, kex, sys->LH_vdw_ht[m], g,
                            st->r_t + 3 * j, f1, f2, E);
                    add3(F + 3 * i, f1);
                    double tq[3];
                    torque_rel(f1, sys->LH_g_pos + 3 * m, a + 3 * i, b + 3 * i, c + 3 * i, tq);
                    add3(T + 3 * i, tq);
                    tail_receive(sys, st, j, f2, F, T, tF);
                }
            }
        }
        k2++;
    }
    /* LHG-linker DNA, 2481-2522 */
    k2 = 0;
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        if (sys->nc_lh_flag[k2] == 1) {
            for (int j = 0; j < n_lh_g; j++) {
                double g[3];
                body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, sys->LH_g_pos + 3 * j, g);
                for (int l = 0; l < n; l++) {
                    if (type[l] != 0) continue;
                    pr_dhev(sys->LH_q[j], sys->q_l, eps, kap, kex, sys->LH_vdw_hl[j], g, r + 3 * l,
                            f1, f2, E);
                    add3(F + 3 * i, f1);
                    add3(F + 3 * l, f2);
                    double tq[3];
                    torque_rel(f1, sys->LH_g_pos + 3 * j, a + 3 * i, b + 3 * i, c + 3 * i, tq);
                    add3(T + 3 * i, tq);
                }
            }
        }
        k2++;
    }
    /* LHG-charges of every other core, 2526-2586 */
    k2 = 0;
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        if (sys->nc_lh_flag[k2] == 1) {
            for (int j = 0; j < n_lh_g; j++) {
                double g[3];
                body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, sys->LH_g_pos + 3 * j, g);
                for (int l = 0; l < n; l++) {
                    if (l == i || type[l] != 1) continue;
                    for (int m = 0; m < sys->Nq; m++) {
                        double z[3];
                        body_to_world(r + 3 * l, a + 3 * l, b + 3 * l, c + 3 * l,
                                      sys->core_pos + 3 * m, z);
                        pr_dhev(sys->LH_q[j], sys->core_q[m], eps, kap, kex, sys->LH_vdw_hc[j], g,
                                z, f1, f2, E);
                        add3(F + 3 * i, f1);
                        add3(F + 3 * l, f2);
                        double tq[3];
                        torque_rel(f1, sys->LH_g_pos + 3 * j, a + 3 * i, b + 3 * i, c + 3 * i, tq);
                        add3(T + 3 * i, tq);
                        torque_rel(f2, sys->core_pos + 3 * m, a + 3 * l, b + 3 * l, c + 3 * l, tq);
                        add3(T + 3 * l, tq);
                    }
                }
            }
        }
        k2++;
    }
    /* LHG-free LH beads of other cores, 2590-2644 */
    k2 = 0;
    int count = 0;
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        if (sys->nc_lh_flag[k2] == 1) {
            for (int m = 0; m < n_lh_g; m++) {
                double g[3];
                body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, sys->LH_g_pos + 3 * m, g);
                for (int j = 0; j < n_LH; j++) {
                    if (j / nlh == count) continue;
                    int i2 = lh_tpl(sys, j);
                    pr_dhev(sys->LH_q[m], sys->LH_q[i2], eps, kap, kex,
                            (sys->LH_vdw_hh[m] + sys->LH_vdw_hh[i2]) / 2, g, rl + 3 * j, f1, f2, E);
                    add3(F + 3 * i, f1);
                    add3(lhF + 3 * j, f2);
                    double tq[3];
                    torque_rel(f1, sys->LH_g_pos + 3 * m, a + 3 * i, b + 3 * i, c + 3 * i, tq);
                    add3(T + 3 * i, tq);
                }
            }
            count++;
        }
        k2++;
    }
    /* LHG-LHG between two LH-carrying cores i < j, 2649-2719 */
    k2 = 0;
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        if (sys->nc_lh_flag[k2] == 1) {
            for (int m = 0; m < n_lh_g; m++) {
                double g[3];
                body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, sys->LH_g_pos + 3 * m, g);
                int kk = k2 + 1;
                for (int j = i + 1; j < n; j++) {
                    if (type[j] != 1) continue;
                    if (sys->nc_lh_flag[kk] == 1) {
                        for (int m1 = 0; m1 < n_lh_g; m1++) {
                            double g2[3];
                            body_to_world(r + 3 * j, a + 3 * j, b + 3 * j, c + 3 * j,
                                          sys->LH_g_pos + 3 * m1, g2);
                            pr_dhev(sys->LH_q[m], sys->LH_q[m1], eps, kap, kex,
                                    (sys->LH_vdw_hh[m] + sys->LH_vdw_hh[m1]) / 2, g, g2, f1, f2, E);
                            add3(F + 3 * i, f1);
                            add3(F + 3 * j, f2);
                            double tq[3];
                            torque_rel(f1, sys->LH_g_pos + 3 * m, a + 3 * i, b + 3 * i, c + 3 * i, tq);
                            add3(T + 3 * i, tq);
                            torque_rel(f2, sys->LH_g_pos + 3 * m1, a + 3 * j, b + 3 * j, c + 3 * j, tq);
                            add3(T + 3 * j, tq);
                        }
                    }
                    kk++;
                }
            }
        }
        k2++;
    }
    /* globular head - first C-term bead spring: force x10, energy x1,
     * torque without the x10, 2727-2777 */
    int kl = 0;
    k2 = 0;
    for (int i = 0; i < n; i++) {
        if (type[i] != 1) continue;
        if (sys->nc_lh_flag[k2] == 1) {
            for (int j = 0; j < n_lh_g; j++) {
                int ti = j + sys->n_lh_n;
                if (sys->LH_conn[ti] != 1) continue;
                double g[3];
                body_to_world(r + 3 * i, a + 3 * i, b + 3 * i, c + 3 * i, sys->LH_g_pos + 3 * j, g);
                const double *p = rl + 3 * kl * nlh;
                double dv[3] = {p[0] - g[0], p[1] - g[1], p[2] - g[2]};
                double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
                double ks = sys->LH_kstr[ti], l0 = sys->LH_streq[ti];
                *E += ks * (d - l0) * (d - l0) / 2;
                double S[3];
                for (int q = 0; q < 3; q++) S[q] = (d - l0) * dv[q] / d;
                for (int q = 0; q < 3; q++) {
                    lhF[3 * kl * nlh + q] -= ks * S[q] * 10;
                    F[3 * i + q] += ks * S[q] * 10;
                }
                double ksS[3] = {ks * S[0], ks * S[1], ks * S[2]};
                double tq[3];
                torque_rel(ksS, sys->LH_g_pos + 3 * j, a + 3 * i, b + 3 * i, c + 3 * i, tq);
                add3(T + 3 * i, tq);
            }
            kl++;
        }
        k2++;
    }
}

void or_forces(const or_system *sys, const or_state *st, double *F, double *T, double *tF,
               double *lhF, double *E) {
    memset(F, 0, sizeof(double) * 3 * sys->n);
    memset(T, 0, sizeof(double) * 3 * sys->n);
    if (sys->n_tail > 0) memset(tF, 0, sizeof(double) * 3 * sys->n_tail);
    if (sys->n_LH > 0) memset(lhF, 0, sizeof(double) * 3 * sys->n_LH);
    *E = 0.0;
    /* order of func_cuda.cu:3566-3603: mechanics first (it assigns), then
     * the accumulating terms */
    or_mech(sys, st, F, T, E);
    or_elec(sys, st, F, T, E);
    if (sys->n_tail > 0) or_tail(sys, st, F, T, tF, E);
    if (sys->n_LH > 0) or_lh(sys, st, F, T, tF, lhF, E);
}

/* ------------------------------------------------------------------ */
/* integrator                                                           */
/* ------------------------------------------------------------------ */

/* rotate, func.cpp:1607-1670 (Rodrigues update of the body frame) */
void or_rotate(int n, const double *a, const double *b, const double *c, double *a_n,
               double *b_n, double *c_n, const double *d_theta, double dt) {
    for (int i = 0; i < n; i++) {
        const int i1 = 3 * i;
        double wa = d_theta[i1], wb = d_theta[i1 + 1], wc = d_theta[i1 + 2];
        double wa2 = wa * wa, wb2 = wb * wb, wc2 = wc * wc;
        double z2 = wa2 + wb2 + wc2, z = sqrt(z2);
        double czt = cos(z * dt), szt = sin(z * dt), om = 1.0 - czt;
        if (!(z2 > 0.0)) {
            for (int k = 0; k < 3; k++) {
                a_n[i1 + k] = a[i1 + k]; b_n[i1 + k] = b[i1 + k]; c_n[i1 + k] = c[i1 + k];
            }
            continue;
        }
        double ga[3] = {((wb2 + wc2) * czt + wa2) / z2, wa * wb * om / z2 + wc * szt / z,
                        wa * wc * om / z2 - wb * szt / z};
        double gb[3] = {wa * wb * om / z2 - wc * szt / z, ((wa2 + wc2) * czt + wb2) / z2,
                        wb * wc * om / z2 + wa * szt / z};
        double gc[3] = {wa * wc * om / z2 + wb * szt / z, wb * wc * om / z2 - wa * szt / z,
                        ((wa2 + wb2) * czt + wc2) / z2};
        for (int k = 0; k < 3; k++) {
            double ak = a[i1 + k], bk = b[i1 + k], ck = c[i1 + k];
            a_n[i1 + k] = ga[0] * ak + ga[1] * bk + ga[2] * ck;
            b_n[i1 + k] = gb[0] * ak + gb[1] * bk + gb[2] * ck;
            c_n[i1 + k] = gc[0] * ak + gc[1] * bk + gc[2] * ck;
        }
    }
}

/* rotation_cal, func_cuda.cu:681-700: d_theta = dt*torque/zeta + rr, cores
 * use all three components, linker beads only the twist (a) component */
static void rot_increment(const or_system *sys, const double *T, const double *rr,
                          double *dth) {
    const double zc = 8 * OR_PI * OR_ETA * 125.0;
    const double zl = 4 * OR_PI * OR_ETA * OR_RH * OR_RH * OR_LO;
    for (int j = 0; j < sys->n; j++) {
        if (sys->type[j] != 0) {
            for (int k = 0; k < 3; k++)
                dth[3 * j + k] = sys->time_step * T[3 * j + k] / zc + rr[3 * j + k];
        } else {
            dth[3 * j] = sys->time_step * T[3 * j] / zl + rr[3 * j];
            dth[3 * j + 1] = 0.0;
            dth[3 * j + 2] = 0.0;
        }
    }
}

/* translation_cal, func_cuda.cu:650-679: x' = x + rd + del * D_blk * F, where
 * D_blk restricts the drift to the bead's own species block (main / tail /
 * LH).  With hi_all == 0 only the main block is hydrodynamically coupled and
 * tails / LH beads see their self-mobility a1/rad. */
static void translate(const or_system *sys, const double *D, int m_hi, const double *x0,
                      const double *rd, const double *F, int off, int cnt3, const double *rad,
                      double *x_out) {
    for (int q = 0; q < cnt3; q++) {
        int row = off + q;
        double v = x0[q] + rd[row];
        if (row < m_hi) {
            const double *Dr = D + (size_t)row * m_hi + off;
            for (int k = 0; k < cnt3; k++) v = v + sys->del * Dr[k] * F[k];
        } else {
            v = v + sys->del * (sys->a1 / rad[q / 3]) * F[q];
        }
        x_out[q] = v;
    }
}

static int step_impl(const or_system *sys, or_state *st, const double *rr, const double *p,
                     const double *L_reuse, double *energy, or_step_out *out) {
    const int n = sys->n, n3 = 3 * n, nt3 = 3 * sys->n_tail, nl3 = 3 * sys->n_LH;
    const int n_D = n + sys->n_tail + sys->n_LH, n_D3 = 3 * n_D;
    const int n_hi = sys->hi_all ? n_D : n, m = 3 * n_hi;
    const int nc = sys->n_c;
    int ret = 0;

    /* gather r_all / rad_all (main.cpp:301-322, step_init func_cuda.cu:1663-1675) */
    double *r_all = malloc(sizeof(double) * (size_t)n_D3);
    double *rad_all = malloc(sizeof(double) * (size_t)n_D);
    memcpy(r_all, st->r, sizeof(double) * n3);
    if (nt3) memcpy(r_all + n3, st->r_t, sizeof(double) * nt3);
    if (nl3) memcpy(r_all + n3 + nt3, st->r_lh, sizeof(double) * nl3);
    memcpy(rad_all, sys->rad, sizeof(double) * n);
    if (sys->n_tail) memcpy(rad_all + n, sys->t_rad, sizeof(double) * sys->n_tail);
    if (sys->n_LH) memcpy(rad_all + n + sys->n_tail, sys->lh_rad, sizeof(double) * sys->n_LH);

    /* LH bending angles at the current state (func_cuda.cu:3497) */
    if (sys->n_LH) or_update_lh_beta(sys->n_LH, sys->n_lh_c, st->r_lh, st->beta_lh);

    /* D and its Cholesky factor (func_cuda.cu:3504-3535) */
    double *D = malloc(sizeof(double) * (size_t)m * m);
    double *L = malloc(sizeof(double) * (size_t)m * m);
    or_diffusion_tensor(n_hi, r_all, rad_all, sys->a1, sys->a2, D);
    int neg = 0;
    if (L_reuse) memcpy(L, L_reuse, sizeof(double) * (size_t)m * m);
    else neg = or_cholesky(m, D, L);

    /* rd = s2dt * L * p (rd_cal, func_cuda.cu:634-647); beads outside the
     * HI set: s2dt * sqrt(a1/rad) * p */
    double *rd = malloc(sizeof(double) * (size_t)n_D3);
    for (int row = 0; row < n_D3; row++) {
        double v = 0.0;
        if (row < m) {
            const double *Lr = L + (size_t)row * m;
            for (int k = 0; k <= row; k++) v += sys->s2dt * Lr[k] * p[k];
        } else {
            v = sys->s2dt * sqrt(sys->a1 / rad_all[row / 3]) * p[row];
        }
        rd[row] = v;
    }

    /* work arrays */
    size_t s3n = sizeof(double) * n3, s1n = sizeof(double) * n, s3c = sizeof(double) * 3 * nc,
           s1c = sizeof(double) * nc;
    double *F0 = calloc(n3, sizeof(double)), *T0 = calloc(n3, sizeof(double));
    double *F1 = calloc(n3, sizeof(double)), *T1 = calloc(n3, sizeof(double));
    double *tF0 = calloc(nt3 + 3, sizeof(double)), *tF1 = calloc(nt3 + 3, sizeof(double));
    double *lF0 = calloc(nl3 + 3, sizeof(double)), *lF1 = calloc(nl3 + 3, sizeof(double));
    double *dth = malloc(s3n);
    or_state pr;
    pr.r = malloc(s3n); pr.a = malloc(s3n); pr.b = malloc(s3n); pr.c = malloc(s3n);
    pr.alpha = malloc(s1n); pr.beta = malloc(s1n); pr.gamma = malloc(s1n); pr.length = malloc(s1n);
    pr.a_dna = malloc(s3c); pr.b_dna = malloc(s3c); pr.c_dna = malloc(s3c);
    pr.alpha_p = malloc(s1c); pr.beta_p = malloc(s1c); pr.gamma_p = malloc(s1c);
    pr.r_t = malloc(sizeof(double) * (nt3 + 3)); pr.beta_t = malloc(sizeof(double) * (sys->n_tail + 1));
    pr.r_lh = malloc(sizeof(double) * (nl3 + 3)); pr.beta_lh = malloc(sizeof(double) * (sys->n_LH + 1));

    /* force pass 1 at x0 (func_cuda.cu:3556-3615) */
    double E0 = 0.0, E1 = 0.0;
    or_forces(sys, st, F0, T0, tF0, lF0, &E0);

    /* predictor (3617-3619) */
    rot_increment(sys, T0, rr, dth);
    or_rotate(n, st->a, st->b, st->c, pr.a, pr.b, pr.c, dth, 1.0);
    translate(sys, D, m, st->r, rd, F0, 0, n3, sys->rad, pr.r);
    if (nt3) translate(sys, D, m, st->r_t, rd, tF0, n3, nt3, sys->t_rad, pr.r_t);
    if (nl3) translate(sys, D, m, st->r_lh, rd, lF0, n3 + nt3, nl3, sys->lh_rad, pr.r_lh);

    /* rebuild derived quantities at the predicted state (3621-3625) */
    if (sys->n_tail) or_update_tail_beta(sys->n_tail, pr.r_t, pr.beta_t, sys->t_grp);
    if (sys->n_LH) or_update_lh_beta(sys->n_LH, sys->n_lh_c, pr.r_lh, pr.beta_lh);
    or_update_euler(sys, pr.r, pr.a, pr.b, pr.c, pr.alpha, pr.beta, pr.gamma, pr.length,
                    pr.a_dna, pr.b_dna, pr.c_dna, pr.alpha_p, pr.beta_p, pr.gamma_p);

    /* force pass 2 at the predicted state (3628-3646) */
    or_forces(sys, &pr, F1, T1, tF1, lF1, &E1);

    /* corrector with the averaged forces, same rd and rr (3648-3652) */
    double *Fa = malloc(s3n), *Ta = malloc(s3n);
    double *tFa = malloc(sizeof(double) * (nt3 + 3)), *lFa = malloc(sizeof(double) * (nl3 + 3));
    for (int q = 0; q < n3; q++) { Fa[q] = 0.5 * (F0[q] + F1[q]); Ta[q] = 0.5 * (T0[q] + T1[q]); }
    for (int q = 0; q < nt3; q++) tFa[q] = 0.5 * (tF0[q] + tF1[q]);
    for (int q = 0; q < nl3; q++) lFa[q] = 0.5 * (lF0[q] + lF1[q]);
    rot_increment(sys, Ta, rr, dth);
    or_rotate(n, st->a, st->b, st->c, pr.a, pr.b, pr.c, dth, 1.0);
    translate(sys, D, m, st->r, rd, Fa, 0, n3, sys->rad, pr.r);
    if (nt3) translate(sys, D, m, st->r_t, rd, tFa, n3, nt3, sys->t_rad, pr.r_t);
    if (nl3) translate(sys, D, m, st->r_lh, rd, lFa, n3 + nt3, nl3, sys->lh_rad, pr.r_lh);

    /* commit and rebuild (3654-3660) */
    memcpy(st->r, pr.r, s3n); memcpy(st->a, pr.a, s3n);
    memcpy(st->b, pr.b, s3n); memcpy(st->c, pr.c, s3n);
    if (nt3) memcpy(st->r_t, pr.r_t, sizeof(double) * nt3);
    if (nl3) memcpy(st->r_lh, pr.r_lh, sizeof(double) * nl3);
    or_update_euler(sys, st->r, st->a, st->b, st->c, st->alpha, st->beta, st->gamma, st->length,
                    st->a_dna, st->b_dna, st->c_dna, st->alpha_p, st->beta_p, st->gamma_p);
    if (sys->n_tail) or_update_tail_beta(sys->n_tail, st->r_t, st->beta_t, sys->t_grp);
    if (sys->n_LH) or_update_lh_beta(sys->n_LH, sys->n_lh_c, st->r_lh, st->beta_lh);

    if (energy) *energy = E0;
    if (out) {
        if (out->D) memcpy(out->D, D, sizeof(double) * (size_t)m * m);
        if (out->L) memcpy(out->L, L, sizeof(double) * (size_t)m * m);
        if (out->rd) memcpy(out->rd, rd, sizeof(double) * n_D3);
        if (out->F0) memcpy(out->F0, F0, s3n);
        if (out->T0) memcpy(out->T0, T0, s3n);
        if (out->F1) memcpy(out->F1, F1, s3n);
        if (out->T1) memcpy(out->T1, T1, s3n);
        if (out->tF0 && nt3) memcpy(out->tF0, tF0, sizeof(double) * nt3);
        if (out->tF1 && nt3) memcpy(out->tF1, tF1, sizeof(double) * nt3);
        if (out->lhF0 && nl3) memcpy(out->lhF0, lF0, sizeof(double) * nl3);
        if (out->lhF1 && nl3) memcpy(out->lhF1, lF1, sizeof(double) * nl3);
        out->neg_pivots = neg;
    }

    free(r_all); free(rad_all); free(D); free(L); free(rd);
    free(F0); free(T0); free(F1); free(T1); free(tF0); free(tF1); free(lF0); free(lF1);
    free(dth); free(Fa); free(Ta); free(tFa); free(lFa);
    free(pr.r); free(pr.a); free(pr.b); free(pr.c); free(pr.alpha); free(pr.beta);
    free(pr.gamma); free(pr.length); free(pr.a_dna); free(pr.b_dna); free(pr.c_dna);
    free(pr.alpha_p); free(pr.beta_p); free(pr.gamma_p); free(pr.r_t); free(pr.beta_t);
    free(pr.r_lh); free(pr.beta_lh);
    return ret;
}

int or_step(const or_system *sys, or_state *st, const double *rr, const double *p,
            double *energy, or_step_out *out) {
    return step_impl(sys, st, rr, p, NULL, energy, out);
}

int or_step_reuse(const or_system *sys, or_state *st, const double *rr, const double *p,
                  const double *L_reuse, double *energy, or_step_out *out) {
    return step_impl(sys, st, rr, p, L_reuse, energy, out);
}

/* ------------------------------------------------------------------ */
/* host RNG                                                             */
/* ------------------------------------------------------------------ */

/* init_genrand, mt.cpp:48-61 */
void or_rng_seed(or_rng *g, uint32_t s) {
    g->mt[0] = s;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->mti = 624;
    g->cached = 0;
    g->n2 = 0.0;
}

/* genrand_int32, mt.cpp:93-128 */
uint32_t or_rng_u32(or_rng *g) {
    if (g->mti >= 624) {
        for (int k = 0; k < 624; k++) {
            uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->mti = 0;
    }
    uint32_t y = g->mt[g->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* genrand_real1, mt.cpp:137-141: [0,1] with divisor 2^32-1 */
double or_rng_real1(or_rng *g) { return or_rng_u32(g) * (1.0 / 4294967295.0); }

/* rand_normal, utilities.cpp:304-336 (Marsaglia polar, second value cached
 * and scaled by the stddev of the call that returns it) */
double or_rng_normal(or_rng *g, double mean, double stddev) {
    if (g->cached) {
        g->cached = 0;
        return g->n2 * stddev + mean;
    }
    double x, y, r;
    do {
        x = 2.0 * or_rng_real1(g) - 1;
        y = 2.0 * or_rng_real1(g) - 1;
        r = x * x + y * y;
    } while (r == 0.0 || r > 1.0);
    double d = sqrt(-2.0 * log(r) / r);
    g->n2 = y * d;
    g->cached = 1;
    return x * d * stddev + mean;
}

void or_rng_step_noise(or_rng *g, int step, int n3, const double *r_var, int n_D3, double *rr,
                       double *p) {
    if (step % 100 == 0) {
        (void)or_rng_real1(g); /* LH bind/unbind draws, main.cpp:387-388 */
        (void)or_rng_real1(g);
    }
    for (int i = 0; i < n3; i++) rr[i] = or_rng_normal(g, 0, r_var[i]);
    for (int i = 0; i < n_D3; i++) p[i] = or_rng_normal(g, 0, 1);
}
