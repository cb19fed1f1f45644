// bd_dev.h -- device-side data layout and inline physics primitives.
//
// Layout in HBM (one context = one fiber):
//   * bead state in the reference AoS layout x[3i+k] (r, a, b, c, r_t, r_lh),
//     so host buffers of the reference driver copy 1:1;
//   * core charges and globular-head (LHG) sites are expanded to world
//     coordinates once per force pass (SoA-free AoS double triples);
//   * per-bead outputs are written by exactly one thread (gather form), so
//     every force is a deterministic fixed-order sum -- no atomics;
//   * energies go to per-thread partial slots reduced in a fixed order.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace bd {

constexpr double kPI = 3.141592653589790;  // constants.h:6
constexpr double kLO = 3.0;
constexpr double kRO = 4.8;
constexpr double kD1 = 1.8;
constexpr double kTHETA = 1.2 * kPI / 2;
constexpr double kETA = 1.137076e-12;
constexpr double kRH = 1.2;
constexpr int kMaxGSites = 8;  // globular-head sites per core held in registers

enum Terms { TERM_MECH = 1, TERM_ELEC = 2, TERM_TAIL = 4, TERM_LH = 8 };

// Positions, frames and derived quantities of one configuration.
struct Geo {
    double *r, *a, *b, *c;                 // 3n
    double *alpha, *beta, *gamma, *length; // n
    double *a_dna, *b_dna, *c_dna;         // 3 n_c
    double *alpha_p, *beta_p, *gamma_p;    // n_c
    double *r_t, *beta_t;                  // 3 n_tail, n_tail
    double *r_lh, *beta_lh;                // 3 n_LH, n_LH
};

// Per-bead force outputs of one force pass.
struct Frc {
    double *F, *T;  // 3n
    double *tF;     // 3 n_tail
    double *lF;     // 3 n_LH
};

// Static device data (everything the step reads but never writes).
struct Stat {
    int n, n_c, n_tail, n_t, n_lh_n, n_lh_g, n_lh_c, nlh, n_LH, n_lhc, Nq, n_lnk;
    // the Taylor coefficients of exp_neg (kExpCoef), carried in the kernel
    // argument so that the pair kernels read them into scalar registers
    struct {
        double c[11];
    } ec;
    const int *type;      // n
    const int *ic_of;     // n: ordinal of the last core at or before bead i
    const int *core_bead; // n_c: bead index of core k
    const int *elig;      // n: bead takes part in the 25 nm core/linker terms
    const int *lnk;       // n_lnk: bead indices of linker beads
    // chain segments of the frame passes: nfseg + 1 bead indices, segment k
    // = beads fseg[k] .. fseg[k+1]-1 (at most 64), each starting at bead 0 or
    // at a core, the last ending at n - 1; nfseg = 0: one segment
    const int *fseg;
    int nfseg;
    const double *phi_o, *rad;
    const double *cpos, *cq;  // 3Nq, Nq
    const double *tail_pos;   // 3 n_t
    const double *h_t, *g_t, *lo_t, *beta_o_t, *t_q, *t_rad;
    const int *t_grp, *t_fix;
    const int *t_rcore;       // n_tail: core bead that takes LH forces of an anchored tail
    const int *rt_off, *rt_list;  // CSR per core ordinal: anchored tails redirected to it
    const double *LH_g_pos;   // 3 n_lh_g
    const int *LH_conn;       // template
    const double *LH_q, *LH_hh, *LH_hc, *LH_hl, *LH_ht, *LH_kstr, *LH_kben, *LH_streq,
        *LH_betaeq;
    const int *lhc_bead;      // n_lhc: core bead of LH ordinal k
    const int *lh_of_core;    // n_c: LH ordinal of core k or -1
    double h, g, s, debyell, debye, q_l, k_e, k_ex;
    double sig_dd, sig_dc, sig_cc, sig_tt, sig_tl, sig_tc;
    double a1, a2, s2dt, del, time_step;
    double c4pe;      // 4*PI*(1/(4*PI*k_e)) as the reference forms it
    double rmax_q;    // max |core_pos|
    double rmax_g;    // max |LH_g_pos|
};

// partners per bounding sphere of the culled partner loops
constexpr int kChunk = 16;
__host__ __device__ inline int nchunks(int n) { return (n + kChunk - 1) / kChunk; }

// Cell grid of one partner set, rebuilt per pass on the GPU: cubic cells
// of edge at least the 7 nm reach over the set's bounding box (the edge
// grows when the box would need more than kCellMax cells); the members of
// each cell are the partner indices in increasing order, cells in x-fastest
// order, so a run of cells along x is one contiguous range of members.
constexpr int kCellMax = 1 << 17;
struct CellGrid {
    // the set's points for the build, set per pass by launch_sites:
    // point k is at pos + 3 idx[k] (idx NULL: pos + 3 k)
    const double *pos;
    const int *idx;
    int n;
    double *hdr;   // 4: origin x, y, z and 1 / edge
    int *dim;      // 3: nx, ny, nz (0 for an empty set)
    int *start;    // kCellMax + 1: first member of each cell
    int *cursor;   // kCellMax: fill cursor of the build
    int *members;  // partner indices, cell by cell
    int *cid;      // cell of each partner
};
// partner sets of the culled loops
enum { SET_TAIL = 0, SET_LH = 1, SET_HEAD = 2, SET_LINKER = 3, NSETS = 4 };

struct Sites {
    double *Q;  // 3 * n_c * Nq world positions of core charges
    // 4 * n_c * Nq: per core charge {world position (3), charge}, the
    // records of the core-core sums
    double *Qp;
    double *G;  // 3 * n_lhc * n_lh_g world positions of globular-head sites
    // bounding spheres (centre x, y, z, radius) of kChunk consecutive
    // partners, rebuilt per pass: tail beads, free LH beads, head sites, and
    // linker beads in the order of the linker list
    double4 *ch_t, *ch_lh, *ch_g, *ch_l;
    // the same partner sets as cell grids; use_cells selects them
    CellGrid cg[NSETS];
    int use_cells;
};

__device__ __forceinline__ double3 ld3(const double *p) { return make_double3(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(double *p, double3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
__device__ __forceinline__ double3 operator+(double3 u, double3 v) { return make_double3(u.x + v.x, u.y + v.y, u.z + v.z); }
__device__ __forceinline__ double3 operator-(double3 u, double3 v) { return make_double3(u.x - v.x, u.y - v.y, u.z - v.z); }
__device__ __forceinline__ double3 operator*(double s, double3 v) { return make_double3(s * v.x, s * v.y, s * v.z); }
__device__ __forceinline__ void operator+=(double3 &u, double3 v) { u.x += v.x; u.y += v.y; u.z += v.z; }
__device__ __forceinline__ void operator-=(double3 &u, double3 v) { u.x -= v.x; u.y -= v.y; u.z -= v.z; }
__device__ __forceinline__ double dot(double3 u, double3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
__device__ __forceinline__ double3 cross(double3 u, double3 v) {
    return make_double3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
// 1/sqrt(f) to full FP64 accuracy: v_rsq_f64 (relative error up to
// 5.2e-8 on gfx950) refined by one third-order step y (1 + r/2 + 3r^2/8),
// r = 1 - f y^2; measured maximum relative error 1.4e-16 over f in
// [e^-12, e^12] (two Newton steps: 2.4e-16, with a longer dependent chain).
__device__ __forceinline__ double rsqrt_f64(double f) {
    const double y = __builtin_amdgcn_rsq(f);
    const double r = __builtin_fma(-f * y, y, 1.0);
    return __builtin_fma(y * r, __builtin_fma(r, 0.375, 0.5), y);
}
// e^x for -700 < x <= 0 without range checks: x = n ln2 + r with
// |r| <= ln2/2 (n rounded to nearest; ln2 split into a high part with
// trailing zero bits and the low remainder, the published fdlibm constants,
// so that n ln2_hi is exact), e^r by its Taylor series to r^12 (truncation
// below 2e-16 relative) in Horner form, scaled by 2^n.
// Taylor coefficients 1/k! of e^r for k = 12 .. 2.
struct ExpCoef {
    double c[11];
};
constexpr ExpCoef kExpCoef = {{1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0,
                               1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0,
                               0.5}};
template <typename C>
__device__ __forceinline__ double exp_neg_c(double x, const C &k) {
    const double n = rint(x * 1.4426950408889634);
    double r = __builtin_fma(-n, 6.93147180369123816490e-01, x);
    r = __builtin_fma(-n, 1.90821492927058770002e-10, r);
    double p = k.c[0];
#pragma unroll
    for (int i = 1; i < 11; i++) p = __builtin_fma(p, r, k.c[i]);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)n);
}
__device__ __forceinline__ double exp_neg(double x) { return exp_neg_c(x, kExpCoef); }
// 2^(j/32) for j = 0..31, each the double nearest to the exact value
struct Exp2Tab {
    double t[32];
};
constexpr Exp2Tab kExp2Tab = {{
                              1.0, 1.0218971486541166, 1.0442737824274138, 1.0671404006768237,
                              1.0905077326652577, 1.1143867425958924, 1.1387886347566916, 1.1637248587775775,
                              1.189207115002721, 1.215247359980469, 1.241857812073484, 1.2690509571917332,
                              1.2968395546510096, 1.3252366431597413, 1.3542555469368927, 1.383909881963832,
                              1.4142135623730951, 1.4451808069770467, 1.4768261459394993, 1.5091644275934228,
                              1.5422108254079407, 1.5759808451078865, 1.6104903319492543, 1.645755478153965,
                              1.681792830507429, 1.718619298122478, 1.7562521603732995, 1.7947090750031072,
                              1.8340080864093424, 1.8741676341103, 1.9152065613971474, 1.9571441241754002}};
// Taylor coefficients of e^r from r^6 down to r^2: 1/6!, 1/5!, 1/4!, 1/3!, 1/2
struct ExpCoef6 {
    double c[5];
};
constexpr ExpCoef6 kExpCoef6 = {{1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5}};
// e^x for -700 < x <= 0 with a 32-entry table (Tang's method): x = (32n + j)
// ln2/32 + r with m = 32n + j = rint(32x/ln2) and |r| <= ln2/64, the two-part
// ln2 above scaled by 1/32 (so m ln2_hi/32 is exact); e^r by its Taylor
// series to r^6 (truncation below 4e-18), times tab[j] = 2^(j/32) and 2^n.
// Seven FMAs instead of the twelve of exp_neg_c, for one table read.
template <typename C>
__host__ __device__ __forceinline__ double exp_neg_t(double x, const double *tab, const C &k) {
    const double m = rint(x * 46.166241308446828);
    double r = __builtin_fma(-m, 6.93147180369123816490e-01 / 32.0, x);
    r = __builtin_fma(-m, 1.90821492927058770002e-10 / 32.0, r);
    const int im = (int)m;
    double p = k.c[0];
#pragma unroll
    for (int i = 1; i < 5; i++) p = __builtin_fma(p, r, k.c[i]);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(tab[im & 31] * p, im >> 5);
}
// sin and cos of x for |x| < 2^19 (the frame angles here lie within
// [-2 pi, 2 pi]), branch-free: x = n pi/2 + r with n = rint(2x/pi) and pi/2
// in three parts (the published constants of fdlibm's e_rem_pio2.c: two
// 33-bit parts, so that n pio2_1 and n pio2_2 are exact, and a 53-bit
// tail), |r| <= pi/4; Taylor series of sin r to r^15 and cos r to r^16
// (truncation below 7e-17) in Horner form in r^2; then the quadrant n mod 4.
__host__ __device__ __forceinline__ void sincos_small(double x, double &sn, double &cs) {
    const double n = rint(x * 6.36619772367581382433e-01);
    double r = __builtin_fma(-n, 1.57079632673412561417e+00, x);
    r = __builtin_fma(-n, 6.07710050630396597660e-11, r);
    r = __builtin_fma(-n, 2.02226624879595063154e-21, r);
    const double z = r * r;
    double ps = -1.0 / 1307674368000.0;
    ps = __builtin_fma(ps, z, 1.0 / 6227020800.0);
    ps = __builtin_fma(ps, z, -1.0 / 39916800.0);
    ps = __builtin_fma(ps, z, 1.0 / 362880.0);
    ps = __builtin_fma(ps, z, -1.0 / 5040.0);
    ps = __builtin_fma(ps, z, 1.0 / 120.0);
    ps = __builtin_fma(ps, z, -1.0 / 6.0);
    const double sr = __builtin_fma(r * z, ps, r);
    double pc = 1.0 / 20922789888000.0;
    pc = __builtin_fma(pc, z, -1.0 / 87178291200.0);
    pc = __builtin_fma(pc, z, 1.0 / 479001600.0);
    pc = __builtin_fma(pc, z, -1.0 / 3628800.0);
    pc = __builtin_fma(pc, z, 1.0 / 40320.0);
    pc = __builtin_fma(pc, z, -1.0 / 720.0);
    pc = __builtin_fma(pc, z, 1.0 / 24.0);
    pc = __builtin_fma(pc, z, -0.5);
    const double cr = __builtin_fma(z, pc, 1.0);
    const int q = (int)n & 3;
    const double s0 = (q & 1) ? cr : sr, c0 = (q & 1) ? sr : cr;
    sn = (q & 2) ? -s0 : s0;
    cs = ((q + 1) & 2) ? -c0 : c0;
}
__device__ __forceinline__ double tan_small(double x) {
    double sn, cs;
    sincos_small(x, sn, cs);
    return sn / cs;
}
__device__ __forceinline__ double3 zero3() { return make_double3(0.0, 0.0, 0.0); }

// r + a*p.x + b*p.y + c*p.z (body point p of a frame at r)
__device__ __forceinline__ double3 body_to_world(double3 r, double3 a, double3 b, double3 c, double3 p) {
    return make_double3(r.x + a.x * p.x + b.x * p.y + c.x * p.z, r.y + a.y * p.x + b.y * p.y + c.y * p.z,
                        r.z + a.z * p.x + b.z * p.y + c.z * p.z);
}

// torque_due_to_force_relative (func.cpp:396-410): comp x (body components of f)
__device__ __forceinline__ double3 torque_rel(double3 f, double3 comp, double3 a, double3 b, double3 c) {
    double fa = dot(a, f), fb = dot(b, f), fc = dot(c, f);
    return make_double3(fc * comp.y - fb * comp.z, fa * comp.z - fc * comp.x, fb * comp.x - fa * comp.y);
}

__device__ __forceinline__ double clamp09(double v) {
    if (v > 0.9) v = 0.9;
    if (v < -0.9) v = -0.9;
    return v;
}

// Pair primitives of func.cpp:278-371 written from the receiver's side:
// the force on the particle at p from a partner at q is
// clamp(F * (p - q) / d), which equals the reference's fp1 when the receiver
// is coord1 and its fp2 when it is coord2 (the clamp is odd-symmetric).
// Pairs beyond both cut-offs contribute exactly zero and are skipped.

// Force_Ele_Vdw: DH below 7 nm plus excluded volume below 4 nm.
template <typename EC>
__device__ __forceinline__ bool dhev_recv(const EC &ec, double q1q2, double c4pe, double kappa, double kev,
                                          double sigma, double3 p, double3 q, double3 &f,
                                          double &E) {
    double3 dv = p - q;
    double d2 = dot(dv, dv);
    // The reference's cut-offs d < 7 and d < 4 on d = sqrt(d2), tested on d2:
    // sqrt is correctly rounded and monotonic, 49 and 16 are exact squares,
    // and the double below each has its root below 7 (4), so sqrt(d2) < 7
    // exactly when d2 < 49 (and < 4 when d2 < 16).  Beyond: exactly zero.
    if (!(d2 < 49.0)) return false;
    // the reference's quotients by d, d^2, d^7 and d^13 as products with
    // 1/d from the refined reciprocal square root (no IEEE root or division),
    // and its quotient by 4 pi eps as a product with the reciprocal, which
    // does not change along a partner loop
    const double id = rsqrt_f64(d2), d = d2 * id, qe = q1q2 * (1.0 / c4pe) * exp_neg_c(-kappa * d, ec);
    double fs = qe * (kappa * d + 1) * (id * id);
    E += qe * id;
    if (d2 < 16.0) {
        const double x = sigma * id, x2 = x * x, x6 = x2 * x2 * x2;
        const double s6 = sigma * sigma, s6b = s6 * s6 * s6;
        const double i2 = id * id, i7 = i2 * i2 * i2 * id;  // d^-7
        fs += -kev * (6 * s6b * i7 - 12 * (s6b * s6b) * (i7 * i2 * i2 * i2));
        E += kev * (x6 * x6 - x6);
    }
    const double fd = fs * id;
    f.x += clamp09(fd * dv.x);
    f.y += clamp09(fd * dv.y);
    f.z += clamp09(fd * dv.z);
    return true;
}

// Squared forms of the reference's other cut-offs on d = sqrt(d2) (the root
// correctly rounded, so monotonic): d <= 8 exactly when d2 <= kD8Sq, the
// double after 64 (its root rounds down to 8); d > 25 exactly when d2 > 625.
constexpr double kD8Sq = 0x1.0000000000001p+6;
constexpr double kD25Sq = 625.0;

// Force_Electrostatics: DH below 7 nm only.
template <typename EC>
__device__ __forceinline__ bool dh_recv(const EC &ec, double q1q2, double c4pe, double kappa, double3 p,
                                        double3 q, double3 &f, double &E) {
    double3 dv = p - q;
    double d2 = dot(dv, dv);
    if (!(d2 < 49.0)) return false;  // d < 7, as in dhev_recv
    const double id = rsqrt_f64(d2), d = d2 * id, qe = q1q2 * (1.0 / c4pe) * exp_neg_c(-kappa * d, ec);
    const double fd = qe * (kappa * d + 1) * (id * id * id);
    E += qe * id;
    f.x += clamp09(fd * dv.x);
    f.y += clamp09(fd * dv.y);
    f.z += clamp09(fd * dv.z);
    return true;
}

// Force_Exclude_Volume: below 4 nm only.
__device__ __forceinline__ bool ev_recv(double kev, double sigma, double3 p, double3 q,
                                        double3 &f, double &E) {
    double3 dv = p - q;
    double d2 = dot(dv, dv);
    if (!(d2 < 16.0)) return false;  // d < 4, as in dhev_recv
    const double id = rsqrt_f64(d2);
    const double x = sigma * id, x2 = x * x, x6 = x2 * x2 * x2;
    const double s6 = sigma * sigma, s6b = s6 * s6 * s6;
    const double i2 = id * id, i7 = i2 * i2 * i2 * id;  // d^-7
    const double fs = -kev * (6 * s6b * i7 - 12 * (s6b * s6b) * (i7 * i2 * i2 * i2));
    E += kev * (x6 * x6 - x6);
    const double fd = fs * id;
    f.x += clamp09(fd * dv.x);
    f.y += clamp09(fd * dv.y);
    f.z += clamp09(fd * dv.z);
    return true;
}

// Force_Stretching (func.cpp:134-151): force on p1 (f1); f2 = -f1.
__device__ __forceinline__ double3 stretch_f1(double k, double3 p1, double3 p2, double l0,
                                              double &E) {
    double3 dv = p1 - p2;
    double d = sqrt(dot(dv, dv));
    E += k * (d - l0) * (d - l0) / 2;
    double f = -k * (d - l0);
    return make_double3(f * dv.x / d, f * dv.y / d, f * dv.z / d);
}

// Force_Bending (func.cpp:153-196): the three projections.
__device__ __forceinline__ void bend3(double kb, double beta, double beta0, double3 p1,
                                     double3 p2, double3 p3, double3 &f1, double3 &f2,
                                     double3 &f3, double &E) {
    E += kb * (beta - beta0) * (beta - beta0) / 2;
    double f = -kb * (beta - beta0);
    double3 ji = p2 - p1, jk = p2 - p3, kj = p3 - p2;
    double d1 = sqrt(dot(ji, ji)), d2 = sqrt(dot(kj, kj));
    double3 n0 = cross(ji, jk);
    double3 ri = cross(ji, n0);
    double nri = sqrt(dot(ri, ri));
    double3 rk = cross(kj, n0);
    double nrk = sqrt(dot(rk, rk));
    double a1 = f / d1, a2 = f / d2;
    f1 = make_double3(a1 * ri.x / nri, a1 * ri.y / nri, a1 * ri.z / nri);
    double3 g = make_double3(a2 * rk.x / nrk, a2 * rk.y / nrk, a2 * rk.z / nrk);
    f2 = make_double3(-a1 * ri.x / nri - g.x, -a1 * ri.y / nri - g.y, -a1 * ri.z / nri - g.z);
    f3 = g;
}

// Block-wide sum of a double over blockDim.x threads (fixed order).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (NT + 63) / 64; i++) t += sh[i];
    }
    __syncthreads();
    return t;  // valid on thread 0
}

// Block-wide sums of K doubles at once (fixed order); results valid on
// thread 0.  sh must hold K * (NT / 64) doubles.
template <int NT, int K>
__device__ __forceinline__ void block_sum_n(double (&v)[K], double *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < K; q++)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < K; q++) sh[q * (NT / 64) + w] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < K; q++) {
            double t = 0.0;
            for (int i = 0; i < NT / 64; i++) t += sh[q * (NT / 64) + i];
            v[q] = t;
        }
    __syncthreads();
}

}  // namespace bd
