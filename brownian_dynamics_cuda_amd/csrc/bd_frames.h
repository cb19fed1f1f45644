// bd_frames.h -- the per-bead bodies of the frame passes (update_Euler_Angle)
// and of the tail / LH bending angles, shared by the frame kernels and the
// predicted state's fused frame-and-sites launch.
//
// They keep the reference's operation order with fused multiply-add
// contraction switched off: the Euler angles come out of acos() of dot
// products that sit at +-1 for straight linkers, where any rounding change
// is amplified to ~1e-8.
#pragma once
#include "bd_dev.h"

namespace bd {

__device__ __forceinline__ double clamp1(double v) {
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    return v;
}

// update_Euler_Angle pass 1 (func.cpp:448-505): linker frames and lengths.
// Only linker frames are written and only core frames are read across
// beads, so one thread per bead has no race (the reference GPU kernel
// fuses both passes and races, func_cuda.cu:722-957).
__device__ __forceinline__ void euler1_bead(const Stat &s, const Geo &g, int i) {
#pragma clang fp contract(off)
    const double si = sin(kTHETA), co = cos(kTHETA);
    double *ri = g.r + 3 * i, *rf = ri + 3, *ai = g.a + 3 * i, *bi = g.b + 3 * i, *ci = g.c + 3 * i;
    const double *af = ai + 3, *bf = bi + 3, *cf = ci + 3;
    double fw[3];
    if (s.type[i] == 0) {
        double ao[3] = {ai[0], ai[1], ai[2]};
        if (s.type[i + 1] == 0) {
            for (int k = 0; k < 3; k++) fw[k] = rf[k] - ri[k];
        } else {
            for (int k = 0; k < 3; k++) {
                double bm = -si * af[k] + co * bf[k];
                fw[k] = rf[k] - kRO * bm + kD1 * cf[k] - ri[k];
            }
        }
        double len = sqrt(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
        g.length[i] = len;
        double mi = 1.0 / len;
        double an[3] = {mi * fw[0], mi * fw[1], mi * fw[2]};
        double da[3] = {an[0] - ao[0], an[1] - ao[1], an[2] - ao[2]};
        double b0 = bi[0], b1 = bi[1], b2 = bi[2];
        double bda = b0 * da[0] + b1 * da[1] + b2 * da[2];
        b0 = b0 - bda * ao[0]; b1 = b1 - bda * ao[1]; b2 = b2 - bda * ao[2];
        bda = b0 * an[0] + b1 * an[1] + b2 * an[2];
        b0 = b0 - bda * an[0]; b1 = b1 - bda * an[1]; b2 = b2 - bda * an[2];
        mi = 1.0 / sqrt(b0 * b0 + b1 * b1 + b2 * b2);
        b0 = mi * b0; b1 = mi * b1; b2 = mi * b2;
        ai[0] = an[0]; ai[1] = an[1]; ai[2] = an[2];
        bi[0] = b0; bi[1] = b1; bi[2] = b2;
        ci[0] = an[1] * b2 - an[2] * b1;
        ci[1] = an[2] * b0 - an[0] * b2;
        ci[2] = an[0] * b1 - an[1] * b0;
    } else {
        for (int k = 0; k < 3; k++) fw[k] = rf[k] - (ri[k] - kRO * bi[k] - kD1 * ci[k]);
        g.length[i] = sqrt(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
    }
}

__device__ __forceinline__ double d3(const double *u, const double *v) {
#pragma clang fp contract(off)
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}
__device__ __forceinline__ double d6(const double *u, const double *v, const double *w, const double *x) {
#pragma clang fp contract(off)
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2] + w[0] * x[0] + w[1] * x[1] + w[2] * x[2];
}

// update_Euler_Angle pass 2 (func.cpp:507-681): Euler angles and the exit
// frame of every core.
__device__ __forceinline__ void euler2_bead(const Stat &s, const Geo &g, int i) {
#pragma clang fp contract(off)
    const double si = sin(kTHETA), co = cos(kTHETA);
    const double *ri = g.r + 3 * i, *rf = ri + 3, *ai = g.a + 3 * i, *bi = g.b + 3 * i, *ci = g.c + 3 * i;
    const double *af = ai + 3, *bf = bi + 3, *cf = ci + 3;
    double ada, f1, f2, Ac, apg, sb;
    if (s.type[i] == 0) {
        double am[3], bm[3];
        if (s.type[i + 1] == 0) {
            for (int k = 0; k < 3; k++) { am[k] = af[k]; bm[k] = bf[k]; }
        } else {
            for (int k = 0; k < 3; k++) {
                am[k] = co * af[k] + si * bf[k];
                bm[k] = -si * af[k] + co * bf[k];
            }
        }
        ada = clamp1(d3(ai, am));
        double be = acos(ada);
        g.beta[i] = be;
        sb = sin(be);
        f1 = (be > 1e-10) ? d3(am, bi) / sb : d3(bm, bi);
        Ac = acos(clamp1(f1));
        f2 = d3(am, ci);
        double al = (f2 >= 0) ? Ac : -Ac;
        g.alpha[i] = al;
        f1 = clamp1(d6(bi, bm, ci, cf) / (1.0 + ada));
        apg = acos(f1);
        f2 = (d3(ci, bm) - d3(bi, cf)) / (1.0 + ada);
        g.gamma[i] = (f2 >= 0.0) ? apg - al : -apg - al;
    } else {
        const int ic = s.ic_of[i];
        double *ad = g.a_dna + 3 * ic, *bd = g.b_dna + 3 * ic, *cd = g.c_dna + 3 * ic;
        const double len = g.length[i];
        double adv[3];
        for (int k = 0; k < 3; k++) adv[k] = (rf[k] - (ri[k] - kRO * bi[k] - kD1 * ci[k])) / len;
        ad[0] = adv[0]; ad[1] = adv[1]; ad[2] = adv[2];
        double cb = clamp1(d3(ai, adv));
        double bp = acos(cb);
        g.beta_p[ic] = bp;
        sb = sin(bp);
        double bdv[3], cdv[3];
        if (bp >= 1e-10) {
            double bm[3];
            for (int k = 0; k < 3; k++) bm[k] = (adv[k] - cb * ai[k]) / sb;
            double ca = clamp1(d3(bm, bi));
            Ac = acos(ca);
            f1 = d3(adv, ci);
            double alp = (f1 >= 0) ? Ac : -Ac;
            double gap = -alp;
            g.alpha_p[ic] = alp;
            g.gamma_p[ic] = gap;
            double sa = sin(alp), sg = sin(gap), cg = cos(gap);
            double R21 = -cg * sb, R22 = cg * cb * ca - sg * sa, R23 = cg * cb * sa + sg * ca;
            double R31 = sg * sb, R32 = -sg * cb * ca - cg * sa, R33 = -sg * cb * sa + cg * ca;
            for (int k = 0; k < 3; k++) {
                bdv[k] = R21 * ai[k] + R22 * bi[k] + R23 * ci[k];
                cdv[k] = R31 * ai[k] + R32 * bi[k] + R33 * ci[k];
            }
        } else {
            for (int k = 0; k < 3; k++) { bdv[k] = bi[k]; cdv[k] = ci[k]; }
        }
        for (int k = 0; k < 3; k++) { bd[k] = bdv[k]; cd[k] = cdv[k]; }
        ada = clamp1(d3(adv, af));
        double be = acos(ada);
        g.beta[i] = be;
        sb = sin(be);
        f1 = (be >= 1e-10) ? d3(af, bdv) / sb : d3(bf, bdv);
        Ac = acos(clamp1(f1));
        f2 = d3(af, cdv);
        double al = (f2 >= 0) ? Ac : -Ac;
        g.alpha[i] = al;
        f1 = clamp1(d6(bdv, bf, cdv, cf) / (1.0 + ada));
        apg = acos(f1);
        f2 = (d3(cdv, bf) - d3(bdv, cf)) / (1.0 + ada);
        g.gamma[i] = (f2 >= 0.0) ? apg - al : -apg - al;
    }
}

// update_tail_beta (func.cpp:1763-1790) and update_LH_beta (2062-2089)
__device__ __forceinline__ void beta_site(const Stat &s, const Geo &g, int i, int do_t, int do_l) {
#pragma clang fp contract(off)
    const double *p1, *p2, *p3;
    double *out;
    bool on;
    if (i < s.n_tail) {
        if (!do_t) return;
        on = (i < s.n_tail - 2) && s.t_grp[i] == s.t_grp[i + 1] && s.t_grp[i] == s.t_grp[i + 2];
        p1 = g.r_t + 3 * i; p2 = p1 + 3; p3 = p1 + 6;
        out = g.beta_t + i;
    } else {
        i -= s.n_tail;
        if (i >= s.n_LH || !do_l) return;
        on = (i % s.n_lh_c != 0) && (i % s.n_lh_c != s.n_lh_c - 1);
        p1 = g.r_lh + 3 * i - 3; p2 = g.r_lh + 3 * i; p3 = g.r_lh + 3 * i + 3;
        out = g.beta_lh + i;
    }
    if (!on) { *out = 0.0; return; }
    double x12 = p1[0] - p2[0], y12 = p1[1] - p2[1], z12 = p1[2] - p2[2];
    double r12 = sqrt(x12 * x12 + y12 * y12 + z12 * z12);
    double x32 = p3[0] - p2[0], y32 = p3[1] - p2[1], z32 = p3[2] - p2[2];
    double r32 = sqrt(x32 * x32 + y32 * y32 + z32 * z32);
    *out = acos((x12 * x32 + y12 * y32 + z12 * z32) / (r12 * r32));
}

}  // namespace bd
