// bd_integrate.hip -- frames, Euler angles, rotation, translation, Heun
// average and commit of the BD step (gfx950).  The per-bead frame and
// bending-angle bodies are in bd_frames.h.
#include "bd_dev.h"
#include "bd_kernels.h"
#include "bd_frames.h"

namespace bd {

static inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

__global__ void k_euler1(Stat s, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.n - 1) euler1_bead(s, g, i);
}
__global__ void k_euler2(Stat s, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.n - 1) euler2_bead(s, g, i);
}
__global__ void k_betas(Stat s, Geo g, int do_t, int do_l) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.n_tail + s.n_LH) beta_site(s, g, i, do_t, do_l);
}
// Both frame passes and the tail/LH bending angles in one launch when the
// chain has at most kFrameThreads + 1 beads: workgroup 0 runs pass 1 and,
// after a barrier, pass 2 (pass 2 reads the linker frames of pass 1); the
// other workgroups take the bending angles, which read positions only.
constexpr int kFrameThreads = 1024;
__global__ void __launch_bounds__(kFrameThreads) k_frames_betas(Stat s, Geo g, int do_t, int do_l) {
    if (blockIdx.x == 0) {
        const int i = threadIdx.x;
        if (i < s.n - 1) euler1_bead(s, g, i);
        __syncthreads();
        if (i < s.n - 1) euler2_bead(s, g, i);
        return;
    }
    const int i = (blockIdx.x - 1) * kFrameThreads + threadIdx.x;
    if (i < s.n_tail + s.n_LH) beta_site(s, g, i, do_t, do_l);
}

// The same with the chain cut into segments that start at cores (Stat
// fseg): a one-wavefront workgroup per segment runs pass 1 and, after a
// barrier, pass 2 of its beads; nothing one segment writes is read by
// another (pass 1 reads core frames only across beads, and the bead after
// a segment is a core).  The bending angles follow, 64 sites a workgroup.
constexpr int kSegThreads = 64;
__global__ void __launch_bounds__(kSegThreads) k_frames_seg(Stat s, Geo g, int do_t, int do_l) {
    if ((int)blockIdx.x < s.nfseg) {
        const int i = s.fseg[blockIdx.x] + threadIdx.x;
        const bool mine = i < s.fseg[blockIdx.x + 1];
        if (mine) euler1_bead(s, g, i);
        __syncthreads();
        if (mine) euler2_bead(s, g, i);
        return;
    }
    const int i = (blockIdx.x - s.nfseg) * kSegThreads + threadIdx.x;
    if (i < s.n_tail + s.n_LH) beta_site(s, g, i, do_t, do_l);
}

// rotation_cal + rotate (func_cuda.cu:681-704, func.cpp:1607-1670): one
// thread per bead, the increment of every bead formed before any frame is
// written (the reference GPU kernel races here, func_cuda.cu:702).
__device__ __forceinline__ void rotate_bead(const Stat &s, const Geo &src, const Geo &dst, int j, const double (&T)[3],
                                            const double *rr) {
    const double zc = 8 * kPI * kETA * 125.0, zl = 4 * kPI * kETA * kRH * kRH * kLO;
    double wa, wb, wc;
    if (s.type[j] != 0) {
        wa = s.time_step * T[0] / zc + rr[3 * j];
        wb = s.time_step * T[1] / zc + rr[3 * j + 1];
        wc = s.time_step * T[2] / zc + rr[3 * j + 2];
    } else {
        wa = s.time_step * T[0] / zl + rr[3 * j];
        wb = 0.0;
        wc = 0.0;
    }
    const double *a = src.a + 3 * j, *b = src.b + 3 * j, *c = src.c + 3 * j;
    double *an = dst.a + 3 * j, *bn = dst.b + 3 * j, *cn = dst.c + 3 * j;
    double wa2 = wa * wa, wb2 = wb * wb, wc2 = wc * wc;
    double z2 = wa2 + wb2 + wc2, z = sqrt(z2);
    double czt, szt;
    sincos_small(z, szt, czt);
    const double om = 1.0 - czt;
    if (!(z2 > 0.0)) {
        for (int k = 0; k < 3; k++) { an[k] = a[k]; bn[k] = b[k]; cn[k] = c[k]; }
        return;
    }
    double ga0 = ((wb2 + wc2) * czt + wa2) / z2, ga1 = wa * wb * om / z2 + wc * szt / z,
           ga2 = wa * wc * om / z2 - wb * szt / z;
    double gb0 = wa * wb * om / z2 - wc * szt / z, gb1 = ((wa2 + wc2) * czt + wb2) / z2,
           gb2 = wb * wc * om / z2 + wa * szt / z;
    double gc0 = wa * wc * om / z2 + wb * szt / z, gc1 = wb * wc * om / z2 - wa * szt / z,
           gc2 = ((wa2 + wb2) * czt + wc2) / z2;
    double A[3] = {a[0], a[1], a[2]}, B[3] = {b[0], b[1], b[2]}, C[3] = {c[0], c[1], c[2]};
    for (int k = 0; k < 3; k++) {
        an[k] = ga0 * A[k] + ga1 * B[k] + ga2 * C[k];
        bn[k] = gb0 * A[k] + gb1 * B[k] + gb2 * C[k];
        cn[k] = gc0 * A[k] + gc1 * B[k] + gc2 * C[k];
    }
}

// RPY 3x3 block between beads at ri and rj (func.cpp:752-799); rows are
// the components of bead i.  Identical value for (j, i): only products of
// pairs of components and squares of r_i - r_j enter.  The far-field
// quotients by s and s^2 as products with one reciprocal of s.
__device__ __forceinline__ void rpy_block(double3 ri, double3 rj, double radi, double radj, double a1,
                                          double a2, double M[3][3]) {
    double sij[3] = {ri.x - rj.x, ri.y - rj.y, ri.z - rj.z};
    double ssq = sij[0] * sij[0] + sij[1] * sij[1] + sij[2] * sij[2];
    double s = sqrt(ssq);
    if (s >= radi + radj) {
        const double is = 1.0 / s, iss = is * is;
        const double f = (radi * radi + radj * radj) * iss;
        const double f1 = 1.0 + f / 3, f2i = (1.0 - f) * iss;
        const double pre = a2 * is;
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
                M[k][l] = (k == l) ? pre * (f1 + f2i * sij[k] * sij[k]) : pre * (f2i * sij[k] * sij[l]);
    } else {
        double se = pow((radi * radi * radi + radj * radj * radj) / 2.0, 1.0 / 3);
        double pre = a1 / se, sq = sqrt(ssq);
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
                M[k][l] = (k == l) ? pre * (1.0 - 9.0 * sq / (32 * se) + 3.0 * sij[k] * sij[k] / (32.0 * se * sq))
                                   : pre * (3.0 * sij[k] * sij[l] / (32.0 * se * sq));
    }
}

// translation_cal (func_cuda.cu:650-679) without a stored tensor: the drift
// del * D_blk * F of a bead is summed over the beads of its species block
// (main / tail / LH) with the RPY blocks recomputed in closed form -- the
// same entries a row of the stored D holds.  One workgroup per bead of
// the hydrodynamic set, threads over partners, fixed-order reduction; the
// workgroups past n_hi take kTrThreads beads outside the set each
// (self-mobility a1/rad only).
constexpr int kTrThreads = 256;

__device__ __forceinline__ void species(const Stat &s, const Geo &src, const Geo &dst, const Frc &f, const Frc &f2,
                                        int b, const double *&x0, const double *&F, const double *&F2,
                                        double *&out, int &lo, int &cnt) {
    if (b < s.n) { x0 = src.r; F = f.F; F2 = f2.F; out = dst.r; lo = 0; cnt = s.n; }
    else if (b < s.n + s.n_tail) { x0 = src.r_t; F = f.tF; F2 = f2.tF; out = dst.r_t; lo = s.n; cnt = s.n_tail; }
    else { x0 = src.r_lh; F = f.lF; F2 = f2.lF; out = dst.r_lh; lo = s.n + s.n_tail; cnt = s.n_LH; }
}
// force component i: F's, or the Heun mean of F and F2 (force_torque_tmp,
// func_cuda.cu:2626-2641) formed where it is read
__device__ __forceinline__ double fcomp(const double *F, const double *F2, int i) {
    return F2 ? 0.5 * (F[i] + F2[i]) : F[i];
}

__global__ void __launch_bounds__(kTrThreads) k_translate(Stat s, int n_hi, int ntr, Geo src, Geo dst, Frc f,
                                                         Frc f2, const double *rd, const double *rad_all,
                                                         const double *rr) {
    __shared__ double red[3 * (kTrThreads / 64)];
    const int b = blockIdx.x;
    const double *x0, *F, *F2;
    double *out;
    int lo, cnt;
    if (b >= ntr) {
        // the rotation of the main beads (frames only; the translation
        // writes positions only)
        const int j = (b - ntr) * kTrThreads + threadIdx.x;
        if (j >= s.n) return;
        const double Tj[3] = {fcomp(f.T, f2.T, 3 * j), fcomp(f.T, f2.T, 3 * j + 1), fcomp(f.T, f2.T, 3 * j + 2)};
        rotate_bead(s, src, dst, j, Tj, rr);
        return;
    }
    if (b >= n_hi) {
        const int bd = n_hi + (b - n_hi) * kTrThreads + threadIdx.x;
        if (bd >= s.n + s.n_tail + s.n_LH) return;
        species(s, src, dst, f, f2, bd, x0, F, F2, out, lo, cnt);
        const int i = bd - lo;
        const double dself = s.a1 / rad_all[bd];
        for (int k = 0; k < 3; k++)
            out[3 * i + k] = x0[3 * i + k] + rd[3 * bd + k] + s.del * dself * fcomp(F, F2, 3 * i + k);
        return;
    }
    species(s, src, dst, f, f2, b, x0, F, F2, out, lo, cnt);
    const int i = b - lo;
    const double radi = rad_all[b];
    const double3 ri = ld3(x0 + 3 * i);
    double v[3] = {0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
        const double Fj[3] = {fcomp(F, F2, 3 * j), fcomp(F, F2, 3 * j + 1), fcomp(F, F2, 3 * j + 2)};
        if (j == i) {
            const double dself = s.a1 / radi;
            for (int k = 0; k < 3; k++) v[k] += s.del * dself * Fj[k];
            continue;
        }
        double M[3][3];
        rpy_block(ri, ld3(x0 + 3 * j), radi, rad_all[lo + j], s.a1, s.a2, M);
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) v[k] += s.del * M[k][l] * Fj[l];
    }
    block_sum_n<kTrThreads, 3>(v, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[3 * i + k] = (x0[3 * i + k] + rd[3 * b + k]) + v[k];
}

// final_updates (func_cuda.cu:2643-2662)
__global__ void k_commit(int n3, int nt3, int nl3, Geo from, Geo to) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) { to.r[i] = from.r[i]; to.a[i] = from.a[i]; to.b[i] = from.b[i]; to.c[i] = from.c[i]; }
    if (i < nt3) to.r_t[i] = from.r_t[i];
    if (i < nl3) to.r_lh[i] = from.r_lh[i];
}

void launch_frames(const Stat &s, const Geo &g, hipStream_t q) {
    if (s.n < 2) return;
    k_euler1<<<nblk(s.n - 1, 64), 64, 0, q>>>(s, g);
    k_euler2<<<nblk(s.n - 1, 64), 64, 0, q>>>(s, g);
}

bool launch_frames_betas(const Stat &s, const Geo &g, int tails, int lh, hipStream_t q) {
    if (s.n < 2) return false;
    const long long tot = (tails || lh) ? (long long)s.n_tail + s.n_LH : 0;
    if (s.nfseg > 0) {
        k_frames_seg<<<s.nfseg + nblk(tot, kSegThreads), kSegThreads, 0, q>>>(s, g, tails, lh);
        return true;
    }
    if (s.n - 1 > kFrameThreads) return false;
    k_frames_betas<<<1 + nblk(tot, kFrameThreads), kFrameThreads, 0, q>>>(s, g, tails, lh);
    return true;
}

void launch_betas(const Stat &s, const Geo &g, int tails, int lh, hipStream_t q) {
    long long tot = (long long)s.n_tail + s.n_LH;
    if (tot > 0 && (tails || lh)) k_betas<<<nblk(tot, 128), 128, 0, q>>>(s, g, tails, lh);
}

void launch_translate(const Stat &s, int hi_all, const Geo &src, const Geo &dst, const Frc &f, const Frc &f2,
                      const double *rd, const double *rad_all, const double *rr, hipStream_t q) {
    const int tot = s.n + s.n_tail + s.n_LH;
    const int n_hi = hi_all ? tot : s.n;
    const int ntr = n_hi + nblk(tot - n_hi, kTrThreads), nrot = rr ? nblk(s.n, kTrThreads) : 0;
    k_translate<<<ntr + nrot, kTrThreads, 0, q>>>(s, n_hi, ntr, src, dst, f, f2, rd, rad_all, rr);
}

void launch_commit(const Stat &s, const Geo &from, const Geo &to, hipStream_t q) {
    int mx = 3 * s.n;
    if (3 * s.n_tail > mx) mx = 3 * s.n_tail;
    if (3 * s.n_LH > mx) mx = 3 * s.n_LH;
    k_commit<<<nblk(mx, 256), 256, 0, q>>>(3 * s.n, 3 * s.n_tail, 3 * s.n_LH, from, to);
}

}  // namespace bd
