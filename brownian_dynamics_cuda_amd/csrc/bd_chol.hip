// bd_chol.hip -- RPY diffusion tensor, blocked Cholesky and the noise
// displacement rd = s2dt * L * p on gfx950.
//
// Replaces Diffusion_Tensor_CUDA + cusolverDnDpotrf + Cholesky_Decomposition_mod
// + rd_cal (func_cuda.cu:500-647, 3504-3552).  The factor follows the CPU
// Cholesky of func.cpp:806-836, including |f| for a negative pivot.
//
// Matrix: row-major, lower triangle, padded to a multiple of 64 with an
// identity block (so the padded factor is [[L,0],[0,I]] and no kernel needs
// a bounds test).  Right-looking blocked algorithm with 64x64 tiles, one
// launch per block column kb (k_step): every workgroup forms the two panel
// tiles it needs as A_ib,kb inv(L_kb,kb)^T (FP64 matrix cores), updates its
// trailing tile A_ib,jb -= L_ib,kb L_jb,kb^T (FP64, FP32 or three-slice
// BF16 matrix-core products), and the workgroup owning tile (kb+1,kb+1)
// then factors it and inverts the factor (four wavefronts, FP64, the
// inverse built beside the pivots).  The solved off-diagonal tile L_ib,kb
// stays parked in the upper block (kb, ib) of the same storage.
//
// Large matrices take super-panels of NB block columns instead: the panel
// columns are factored as above with their updates restricted to the panel,
// and the trailing matrix then receives the whole panel in one launch
// (k_trail: each tile accumulates K = 64 NB in registers before its one
// read-modify-write), so the trailing matrix streams through HBM once per
// NB block columns rather than once per block column.  On large trailing
// matrices the panel tiles are first copied once into the operand order of
// the matrix cores (k_pack64; k_split3 splits them into BF16 slices) and
// the update takes 128x128 blocks fed by global-to-LDS copies (k_trail_pk,
// k_trail_bf).  In the FP32 and BF16 modes only that trailing update takes
// the reduced precision; the panel's own columns stay FP64.
#include <algorithm>

#include "bd_dev.h"
#include "bd_kernels.h"

namespace bd {

constexpr int TB = 64;        // tile edge
constexpr int LDP = TB + 1;   // LDS row pitch (doubles)
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

static inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// RPY tensor assembly (func.cpp:737-804): one workgroup per lower 64x64
// tile, a thread per bead pair whose 3x3 block meets the tile (the pair's
// distance terms evaluated once, each entry by the reference's expression),
// staged in LDS and stored row by row; zeros above the diagonal and the
// identity padding.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double3 hi_pos(const Stat &s, const Geo &g, int b) {
    if (b < s.n) return ld3(g.r + 3 * b);
    b -= s.n;
    if (b < s.n_tail) return ld3(g.r_t + 3 * b);
    return ld3(g.r_lh + 3 * (b - s.n_tail));
}

// M[k][l] for beads i != j at ri, rj (func.cpp:754-796)
__device__ __forceinline__ void rpy_block(double3 ri, double3 rj, double radi, double radj, double a1, double a2,
                                          double (&M)[3][3]) {
    const double sij[3] = {ri.x - rj.x, ri.y - rj.y, ri.z - rj.z};
    const double ssq = sij[0] * sij[0] + sij[1] * sij[1] + sij[2] * sij[2];
    const double s = sqrt(ssq);
    if (s >= radi + radj) {
        const double f = (radi * radi + radj * radj) / ssq;
        const double f1 = 1.0 + f / 3, f2 = 1.0 - f, c = a2 / s;
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int l = 0; l < 3; l++)
                M[k][l] = (k == l) ? c * (f1 + f2 * sij[k] * sij[k] / ssq) : c * (f2 * sij[k] * sij[l] / ssq);
    } else {
        const double se = pow((radi * radi * radi + radj * radj * radj) / 2.0, 1.0 / 3);
        const double sq = sqrt(ssq), c = a1 / se;
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int l = 0; l < 3; l++)
                M[k][l] = (k == l) ? c * (1.0 - 9.0 * sq / (32 * se) + 3.0 * sij[k] * sij[k] / (32.0 * se * sq))
                                   : c * (3.0 * sij[k] * sij[l] / (32.0 * se * sq));
    }
}

__device__ __forceinline__ void tri_index(int idx, int &i, int &j) {
    // idx -> (i, j), 0 <= j <= i, row-major over the lower triangle
    int r = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= idx) r++;
    while (r * (r + 1) / 2 > idx) r--;
    i = r;
    j = idx - r * (r + 1) / 2;
}

__device__ __forceinline__ void diag_factor_inv(double *T, double *LP, double *X, double *A, int ld, int d0,
                                                double *Linv, int *neg);

// FACTOR: workgroup 0 (tile (0,0)) then factors and inverts its tile in
// place of a separate first-tile launch, beside the other tiles' assembly
// (every workgroup then carries the factor's LDS).
template <bool FACTOR>
__global__ void __launch_bounds__(256) k_rpy(Stat s, Geo g, const double *rad_all, double *A, int m, int ld,
                                             double a1, double a2, int *neg, double *Linv) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *neg = 0;  // negative-pivot count of the factor that follows
    __shared__ double T[TB * LDP];
    __shared__ double LP[FACTOR ? TB * TB : 1], X[FACTOR ? TB * LDP : 1];
    int ti, tj;
    tri_index(blockIdx.x, ti, tj);
    const int r0 = ti * TB, c0 = tj * TB, nhi = m / 3;
    // beads whose rows / columns meet the tile
    const int bi0 = r0 / 3, bj0 = c0 / 3;
    const int nbi = max(0, min((r0 + TB - 1) / 3, nhi - 1) - bi0 + 1);
    const int nbj = max(0, min((c0 + TB - 1) / 3, nhi - 1) - bj0 + 1);
    for (int p = threadIdx.x; p < nbi * nbj; p += blockDim.x) {
        const int i = bi0 + p / nbj, j = bj0 + p % nbj;
        if (j > i) continue;
        double M[3][3];
        if (i == j) {
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int l = 0; l < 3; l++) M[k][l] = (k == l) ? a1 / rad_all[i] : 0.0;
        } else {
            rpy_block(hi_pos(s, g, i), hi_pos(s, g, j), rad_all[i], rad_all[j], a1, a2, M);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int row = 3 * i + k - r0;
            if (row < 0 || row >= TB) continue;
#pragma unroll
            for (int l = 0; l < 3; l++) {
                const int col = 3 * j + l - c0;
                if (col >= 0 && col < TB) T[row * LDP + col] = M[k][l];
            }
        }
    }
    __syncthreads();
    const bool fac = FACTOR && blockIdx.x == 0;
    for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
        const int row = r0 + e / TB, col = c0 + e % TB;
        double v;
        if (col > row) v = 0.0;
        else if (row >= m || col >= m) v = (row == col) ? 1.0 : 0.0;
        else v = T[(e / TB) * LDP + e % TB];
        if (fac) T[(e / TB) * LDP + e % TB] = v;  // the factor writes L here instead
        else A[(size_t)row * ld + col] = v;
    }
    if constexpr (FACTOR) {
        if (fac) {
            __syncthreads();
            diag_factor_inv(T, LP, X, A, ld, 0, Linv, neg);
        }
    }
}

// ---------------------------------------------------------------------------
// tile helpers
// ---------------------------------------------------------------------------
// Tile copies between global memory and LDS (pitch LDP) by the 256 threads
// of a workgroup, 16-byte accesses, fully unrolled: all eight global loads
// of a thread are in flight before the first LDS store waits on them.
constexpr int NTHR = 256;
__device__ __forceinline__ void load_tile(const double *A, int ld, int r0, int c0, double *S) {
    dbl2 v[TB * TB / 2 / NTHR];
#pragma unroll
    for (int i = 0; i < TB * TB / 2 / NTHR; i++) {
        const int e = threadIdx.x + i * NTHR, r = e / (TB / 2), c = 2 * (e % (TB / 2));
        v[i] = *reinterpret_cast<const dbl2 *>(A + (size_t)(r0 + r) * ld + c0 + c);
    }
#pragma unroll
    for (int i = 0; i < TB * TB / 2 / NTHR; i++) {
        const int e = threadIdx.x + i * NTHR, r = e / (TB / 2), c = 2 * (e % (TB / 2));
        S[r * LDP + c] = v[i].x;
        S[r * LDP + c + 1] = v[i].y;
    }
}
__device__ __forceinline__ void store_tile(double *A, int ld, int r0, int c0, const double *S, bool lower_only) {
#pragma unroll
    for (int i = 0; i < TB * TB / 2 / NTHR; i++) {
        const int e = threadIdx.x + i * NTHR, r = e / (TB / 2), c = 2 * (e % (TB / 2));
        dbl2 v;
        v.x = (lower_only && c > r) ? 0.0 : S[r * LDP + c];
        v.y = (lower_only && c + 1 > r) ? 0.0 : S[r * LDP + c + 1];
        if (!lower_only || c <= r) *reinterpret_cast<dbl2 *>(A + (size_t)(r0 + r) * ld + c0 + c) = v;
    }
}

// acc += As * Bs^T for a 64x64 tile; 4 waves, each a 32x32 quarter as 2x2
// tiles of v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][l/16] and
// B^T[l/16][l%16]; result register v of lane l is C[l/16 + 4v][l%16]
// (measured on gfx950; the FP32 16x16x4 form places it at C[4(l/16)+v]).
__device__ __forceinline__ void mma_f64(const double *As, const double *Bs, dbl4 acc[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, i16 = lane & 15, k4 = lane >> 4;
    // the operands of K step kk + 4 are read while the products of step kk run
    const double *pa0 = As + (wr + i16) * LDP + k4, *pa1 = pa0 + 16 * LDP;
    const double *pb0 = Bs + (wc + i16) * LDP + k4, *pb1 = pb0 + 16 * LDP;
    double na0 = pa0[0], na1 = pa1[0], nb0 = pb0[0], nb1 = pb1[0];
#pragma unroll
    for (int kk = 0; kk < TB; kk += 4) {
        const double a0 = na0, a1 = na1, b0 = nb0, b1 = nb1;
        if (kk + 4 < TB) {
            na0 = pa0[kk + 4];
            na1 = pa1[kk + 4];
            nb0 = pb0[kk + 4];
            nb1 = pb1[kk + 4];
        }
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// Same product with FP32 operands and FP32 accumulation
// (v_mfma_f32_16x16x4f32, identical lane layout).
__device__ __forceinline__ void mma_f32(const double *As, const double *Bs, flt4 acc[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, i16 = lane & 15, k4 = lane >> 4;
    const double *pa0 = As + (wr + i16) * LDP + k4, *pa1 = pa0 + 16 * LDP;
    const double *pb0 = Bs + (wc + i16) * LDP + k4, *pb1 = pb0 + 16 * LDP;
    double na0 = pa0[0], na1 = pa1[0], nb0 = pb0[0], nb1 = pb1[0];
#pragma unroll
    for (int kk = 0; kk < TB; kk += 4) {
        const float a0 = (float)na0, a1 = (float)na1, b0 = (float)nb0, b1 = (float)nb1;
        if (kk + 4 < TB) {
            na0 = pa0[kk + 4];
            na1 = pa1[kk + 4];
            nb0 = pb0[kk + 4];
            nb1 = pb1[kk + 4];
        }
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// Three-slice BF16 form of the same product.  Each operand x is rounded to
// FP32 and split exactly into three BF16 numbers, x_f32 = h + m + l
// (h = bf16(x), m = bf16(x - h), l = x - h - m, which has at most 8
// significant bits and is exact); the six products whose weight reaches
// 2^-24 -- hh, hm, mh, hl, mm, lh -- are accumulated in FP32 on
// v_mfma_f32_16x16x32_bf16, smallest first.  The result carries FP32-class
// accuracy from BF16 matrix-core products.  Lane l supplies row l%16 and
// the eight K entries 8(l/16)..8(l/16)+7 of a 32-wide K step (the same K
// mapping for both operands, so the sum is the same whatever the hardware's
// K order); results as in the FP32 form (row 4(l/16) + v).
__device__ __forceinline__ void split3(const double *src, bf8 &h, bf8 &m, bf8 &l) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const float f = (float)src[e];
        const __bf16 hh = (__bf16)f;
        const float r1 = f - (float)hh;
        const __bf16 mm = (__bf16)r1;
        h[e] = hh;
        m[e] = mm;
        l[e] = (__bf16)(r1 - (float)mm);
    }
}
__device__ __forceinline__ void mma_bf16x3(const double *As, const double *Bs, flt4 acc[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, i16 = lane & 15, k8 = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < TB; kk += 32) {
        bf8 ah[2], am[2], al[2], bh[2], bm[2], bl[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            split3(As + (wr + 16 * q + i16) * LDP + kk + 8 * k8, ah[q], am[q], al[q]);
            split3(Bs + (wc + 16 * q + i16) * LDP + kk + 8 * k8, bh[q], bm[q], bl[q]);
        }
#pragma unroll
        for (int mi = 0; mi < 2; mi++)
#pragma unroll
            for (int ni = 0; ni < 2; ni++) {
                flt4 c = acc[mi][ni];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mi], bh[ni], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[mi], bm[ni], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bl[ni], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[mi], bh[ni], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bm[ni], c, 0, 0, 0);
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bh[ni], c, 0, 0, 0);
            }
    }
}

// acc (-)= As Bs^T for one 64x64 tile in the arithmetic of MODE
// (BD_CHOL_FP64 / FP32 / BF16); Acc is dbl4 for FP64, flt4 otherwise.
template <int MODE, typename Acc>
__device__ __forceinline__ void tile_mma(const double *As, const double *Bs, Acc acc[2][2]) {
    if constexpr (MODE == 0) mma_f64(As, Bs, acc);
    else if constexpr (MODE == 1) mma_f32(As, Bs, acc);
    else mma_bf16x3(As, Bs, acc);
}
template <int MODE> struct AccOf { typedef flt4 T; };
template <> struct AccOf<0> { typedef dbl4 T; };

// C (LDS) += sign * acc, element mapping of mma_* (row of register v of
// lane l: l/16 + 4v for the FP64 MFMA, 4(l/16) + v for the FP32 one)
template <typename V>
__device__ __forceinline__ void acc_to_lds(double *C, V acc[2][2], double sign, bool overwrite) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, i16 = lane & 15, k4 = lane >> 4;
    constexpr bool f64 = sizeof(acc[0][0][0]) == 8;
    for (int mi = 0; mi < 2; mi++)
        for (int ni = 0; ni < 2; ni++)
            for (int v = 0; v < 4; v++) {
                int r = wr + 16 * mi + (f64 ? k4 + 4 * v : 4 * k4 + v), c = wc + 16 * ni + i16;
                double x = sign * (double)acc[mi][ni][v];
                if (overwrite) C[r * LDP + c] = x;
                else C[r * LDP + c] += x;
            }
}

// Copies up to four tiles from global memory into LDS with every global
// load of the thread in flight before the first LDS store (n: tiles used).
__device__ __forceinline__ void load_tiles(int n, const double *const (&src)[4], const int (&ld)[4],
                                           double *const (&dst)[4]) {
    constexpr int PER = TB * TB / 2 / NTHR;
    dbl2 v[4][PER];
#pragma unroll
    for (int t = 0; t < 4; t++)
        if (t < n)
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int e = threadIdx.x + i * NTHR, r = e / (TB / 2), c = 2 * (e % (TB / 2));
                v[t][i] = *reinterpret_cast<const dbl2 *>(src[t] + (size_t)r * ld[t] + c);
            }
#pragma unroll
    for (int t = 0; t < 4; t++)
        if (t < n)
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int e = threadIdx.x + i * NTHR, r = e / (TB / 2), c = 2 * (e % (TB / 2));
                dst[t][r * LDP + c] = v[t][i].x;
                dst[t][r * LDP + c + 1] = v[t][i].y;
            }
}

// A 16x16 output block of the FP64 matrix cores (mapping of mma_f64:
// register v of lane l is element (l/16 + 4v, l%16) of the block) into C.
__device__ __forceinline__ void blk_to_lds(double *C, int I, int J, dbl4 acc, double sign, bool overwrite) {
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        double *c = C + (16 * I + k4 + 4 * v) * LDP + 16 * J + i16;
        if (overwrite) *c = sign * acc[v];
        else *c += sign * acc[v];
    }
}

// Panel product P = A inv(L)^T with inv(L) lower triangular (zero above
// its diagonal 16x16 blocks): output block column J needs k < 16 (J + 1)
// only, and wavefront w takes the 16-row stripe w of P, so every
// wavefront issues the same 40 products.  Skipping the zero blocks adds
// nothing to any sum, so P is the full product bit for bit.
__device__ __forceinline__ void mma_panel(const double *As, const double *Li, dbl4 (&acc)[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i16 = lane & 15, k4 = lane >> 4;
    const double *pa = As + (16 * w + i16) * LDP + k4, *pl = Li + i16 * LDP + k4;
    double na = pa[0], nb[4];
#pragma unroll
    for (int J = 0; J < 4; J++) nb[J] = pl[16 * J * LDP];
#pragma unroll
    for (int kk = 0; kk < TB; kk += 4) {
        const double a = na;
        double b[4];
#pragma unroll
        for (int J = 0; J < 4; J++) b[J] = nb[J];
        if (kk + 4 < TB) {  // next K step's operands beside this step's products
            na = pa[kk + 4];
#pragma unroll
            for (int J = 0; J < 4; J++)
                if (kk + 4 < 16 * (J + 1)) nb[J] = pl[16 * J * LDP + kk + 4];
        }
#pragma unroll
        for (int J = 0; J < 4; J++)
            if (kk < 16 * (J + 1)) acc[J] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[J], acc[J], 0, 0, 0);
    }
}
__device__ __forceinline__ void panel_to_lds(double *C, const dbl4 (&acc)[4]) {
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int J = 0; J < 4; J++) blk_to_lds(C, w, J, acc[J], 1.0, true);
}

// Diagonal tile update C -= P P^T on its ten lower 16x16 blocks only (the
// factor never reads above the diagonal), three or two per wavefront, their
// products interleaved (independent accumulators) with the next K step's
// operands read beside them; each block sums K in increasing order from
// zero and is then subtracted from C.
template <int N, int I0, int J0, int I1, int J1, int I2, int J2>
__device__ __forceinline__ void syrk_blocks(const double *P, double *C) {
    constexpr int I[3] = {I0, I1, I2}, J[3] = {J0, J1, J2};
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    const double *p = P + i16 * LDP + k4;
    dbl4 acc[N];
    double na[N], nb[N];
#pragma unroll
    for (int b = 0; b < N; b++) {
        acc[b] = dbl4{0.0, 0.0, 0.0, 0.0};
        na[b] = p[16 * I[b] * LDP];
        nb[b] = p[16 * J[b] * LDP];
    }
#pragma unroll
    for (int kk = 0; kk < TB; kk += 4) {
        double a[N], bv[N];
#pragma unroll
        for (int b = 0; b < N; b++) {
            a[b] = na[b];
            bv[b] = nb[b];
        }
        if (kk + 4 < TB)
#pragma unroll
            for (int b = 0; b < N; b++) {
                na[b] = p[16 * I[b] * LDP + kk + 4];
                nb[b] = p[16 * J[b] * LDP + kk + 4];
            }
#pragma unroll
        for (int b = 0; b < N; b++) acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[b], bv[b], acc[b], 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < N; b++) blk_to_lds(C, I[b], J[b], acc[b], -1.0, false);
}
__device__ __forceinline__ void syrk_lower(const double *P, double *C) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // (I, J) blocks of wavefront w: {00 10 11} {20 21 22} {30 31} {32 33}
    if (w == 0) syrk_blocks<3, 0, 0, 1, 0, 1, 1>(P, C);
    else if (w == 1) syrk_blocks<3, 2, 0, 2, 1, 2, 2>(P, C);
    else if (w == 2) syrk_blocks<2, 3, 0, 3, 1, 0, 0>(P, C);
    else syrk_blocks<2, 3, 2, 3, 3, 0, 0>(P, C);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------
// Diagonal tile: factor and inverse by the four wavefronts of a workgroup.
//
// Wavefront w holds 16 columns of every row (lane = row) in g[]: panels w
// and w + 4 of the 8 panels of 8 columns, which it factors in turn: the CPU
// algorithm of func.cpp:806-836 in right-looking order -- pivot f (|f| and
// a count if f < 0), column scaled by 1/sqrt(f) -- with the updates inside
// a panel taken from the owning lanes by readlane.  A finished panel goes
// to LP in panel order, LP[p][row][q] = L[row][8p + q], and the other
// wavefronts apply it to their columns (rank-8 updates that read the 8
// entries of a column as broadcast 16-byte LDS loads).  With the panels
// dealt round-robin, the owner of the next panel has only that panel's 8
// columns to bring up to date before it pivots: that catch-up and the
// pivots are the critical path.  The wavefronts with time left build
// X = inv(L) from 16x16 blocks: the diagonal blocks from 8x8 substitutions,
// the off-diagonal blocks on the matrix cores,
//   X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj.
// ---------------------------------------------------------------------------

// g[C0 .. C0+NC) -= sum_q x[q] * lp[col*8 + q] for the columns col =
// colbase + C0 + c, each column's sum in increasing q.  Four columns at a
// time, their FMA chains interleaved (with two chains the dependent FP64
// FMA latency, not the issue rate, set the pace); the entries of the next
// four columns (broadcast 16-byte LDS loads) are loaded one group ahead (a
// scheduling barrier per group keeps that distance).
template <int C0, int NC>
__device__ __forceinline__ void panel_update16(double (&g)[16], const double (&x)[8], const double *lp,
                                               int colbase) {
    constexpr int NG = NC / 4;
    const double2 *l2 = reinterpret_cast<const double2 *>(lp) + (colbase + C0) * 4;
    double2 e[NG][16];
#pragma unroll
    for (int k = 0; k < 16; k++) e[0][k] = l2[k];
#pragma unroll
    for (int i = 0; i < NG; i++) {
        if (i + 1 < NG)
#pragma unroll
            for (int k = 0; k < 16; k++) e[i + 1][k] = l2[(i + 1) * 16 + k];
        const int c = C0 + 4 * i;
        double s[4] = {g[c], g[c + 1], g[c + 2], g[c + 3]};
#pragma unroll
        for (int k = 0; k < 4; k++) {
#pragma unroll
            for (int m = 0; m < 4; m++) s[m] -= x[2 * k] * e[i][4 * m + k].x;
#pragma unroll
            for (int m = 0; m < 4; m++) s[m] -= x[2 * k + 1] * e[i][4 * m + k].y;
        }
#pragma unroll
        for (int m = 0; m < 4; m++) g[c + m] = s[m];
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Pivots of panel P (columns 8P..8P+7) held in g[L0 .. L0+7]; publishes the
// panel to LP and the pivot inverses 1/L_jj of its lanes to ivs.
// Branch-free, so the row stays in registers.
template <int P, int L0>
__device__ __forceinline__ void pivot_panel(double (&g)[16], double *LP, double *ivs, double &myinv, int &nneg) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int jc = 8 * P + j;
        const double f0 = readlane_f64(g[L0 + j], jc);
        nneg += f0 < 0 ? 1 : 0;
        const double f = fabs(f0);
        const double inv = rsqrt_f64(f);
        const double lj = lane > jc ? g[L0 + j] * inv : (lane == jc ? f * inv : 0.0);
        myinv = lane == jc ? inv : myinv;
        g[L0 + j] = lj;
#pragma unroll
        for (int c = j + 1; c < 8; c++) g[L0 + c] -= lj * readlane_f64(lj, 8 * P + c);
    }
    double *lp = LP + P * (TB * 8);
#pragma unroll
    for (int q = 0; q < 8; q += 2) {
        double2 v;
        v.x = g[L0 + q];
        v.y = g[L0 + q + 1];
        *reinterpret_cast<double2 *>(lp + lane * 8 + q) = v;
    }
    if (lane >= 8 * P && lane < 8 * P + 8) ivs[lane] = myinv;
}

// this lane's row of panel P, from LP
template <int P>
__device__ __forceinline__ void panel_row(const double *LP, double (&x)[8]) {
    const double2 *r = reinterpret_cast<const double2 *>(LP + P * (TB * 8) + (threadIdx.x & 63) * 8);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double2 d = r[k];
        x[2 * k] = d.x;
        x[2 * k + 1] = d.y;
    }
}

// X_ww = inv(L_ww) for the 16x16 diagonal block w, by wavefront w, in 8x8
// blocks: L_ww = [[A, 0], [B, C]] gives X_ww = [[A^-1, 0], [-C^-1 B A^-1, C^-1]].
// Lanes 0-7 solve A y = e_r and lanes 8-15 C y = e_r by substitution, side
// by side (y_c = (e_r[c] - sum_{j<c} L_cj y_j) / L_cc, eight steps); then
// every lane forms one entry of Y = B A^-1 and one of -C^-1 Y.  The upper
// triangle and the upper right block are written as zeros (the products
// that read X_ww take whole 16x16 blocks).  S: the wavefront's LDS scratch
// (8 rows of pitch 17).
__device__ __forceinline__ void diag_block_inverse(const double *LP, const double *ivs, double *X, double *S, int w) {
    const int lane = threadIdx.x & 63;
    {
        const int h = (lane >> 3) & 1, r = lane & 7;
        const int b0 = 16 * w + 8 * h;  // first row and column of block A (h = 0) or C (h = 1)
        // L[b0 + c][b0 + j] = lh[c*8 + j]: panel 2w + h, rows b0..b0+7
        const double *lh = LP + (2 * w + h) * (TB * 8) + b0 * 8;
        const double2 *i2 = reinterpret_cast<const double2 *>(ivs + b0);
        double iw[8];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double2 v = i2[k];
            iw[2 * k] = v.x;
            iw[2 * k + 1] = v.y;
        }
        double y[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            double l[8];
#pragma unroll
            for (int j = 0; j < c; j += 2) {
                const double2 v = *reinterpret_cast<const double2 *>(lh + c * 8 + j);
                l[j] = v.x;
                l[j + 1] = v.y;
            }
            double sum = (c == r) ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < c; j++) sum -= y[j] * l[j];
            y[c] = sum * iw[c];
        }
        if (lane < 16) {
#pragma unroll
            for (int c = 0; c < 8; c++) X[(b0 + c) * LDP + b0 + r] = y[c];
            if (h)
#pragma unroll
                for (int c = 0; c < 8; c++) X[(16 * w + c) * LDP + b0 + r] = 0.0;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int a = lane >> 3, b = lane & 7, d0 = 16 * w;
    // Y[a][b] = sum_k B[a][k] A^-1[k][b], B[a][k] = L[d0 + 8 + a][d0 + k] (panel 2w)
    const double *lb = LP + 2 * w * (TB * 8) + (d0 + 8 + a) * 8;
    double yv = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) yv += lb[k] * X[(d0 + k) * LDP + d0 + b];
    S[a * 17 + b] = yv;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double z = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) z += X[(d0 + 8 + a) * LDP + d0 + 8 + k] * S[k * 17 + b];
    X[(d0 + 8 + a) * LDP + d0 + b] = -z;
}

// 16x16 block products on the matrix cores (v_mfma_f64_16x16x4f64; result
// register v of lane l is element (l/16 + 4v, l%16)), one wavefront each.
// acc += L_ik X_kj, with L in panel order (LP) and X of pitch LDP.
__device__ __forceinline__ dbl4 mm_lx(dbl4 acc, const double *LP, const double *X, int i, int k, int j) {
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
        const int col = 16 * k + kk + k4;
        const double a = LP[(col >> 3) * (TB * 8) + (16 * i + i16) * 8 + (col & 7)];
        const double b = X[(16 * k + kk + k4) * LDP + 16 * j + i16];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}
// acc += X_ii S, S a 16x16 block of pitch 17
__device__ __forceinline__ dbl4 mm_xs(dbl4 acc, const double *X, const double *S, int i) {
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
        const double a = X[(16 * i + i16) * LDP + 16 * i + kk + k4];
        const double b = S[(kk + k4) * 17 + i16];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}
// D (a 16x16 block of the given pitch) = sign * acc; the wavefront's later
// LDS reads of D are ordered after these stores
__device__ __forceinline__ void put_block(double *D, int pitch, dbl4 acc, double sign) {
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int v = 0; v < 4; v++) D[(k4 + 4 * v) * pitch + i16] = sign * acc[v];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
// S = sum_{k=j}^{i-1} L_ik X_kj (the sum of X_ij = -X_ii S without X_ii)
__device__ __forceinline__ void x_sum(const double *LP, const double *X, double *S, int i, int j) {
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k = j; k < i; k++) acc = mm_lx(acc, LP, X, i, k, j);
    put_block(S, 17, acc, 1.0);
}
// X_ij = -X_ii S from a sum S formed by x_sum (possibly by another wavefront)
__device__ __forceinline__ void x_final(double *X, const double *S, int i, int j) {
    const dbl4 z = {0.0, 0.0, 0.0, 0.0};
    put_block(X + 16 * i * LDP + 16 * j, LDP, mm_xs(z, X, S, i), -1.0);
}
// S (+)= sum_{k=k0}^{k1-1} L_3k X_kj: terms of the part of X_3j that does
// not need X_33, S_3j = sum_{k=j}^{2} L_3k X_kj; with ADD the sum continues
// from the partial sum already in S (the same order as in one pass).
__device__ __forceinline__ void s3_part(const double *LP, const double *X, double *S, int j, int k0, int k1,
                                        bool add) {
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    if (add)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[v] = S[(k4 + 4 * v) * 17 + i16];
    for (int k = k0; k < k1; k++) acc = mm_lx(acc, LP, X, 3, k, j);
    put_block(S, 17, acc, 1.0);
}
// Rows r0 .. r0+15 of inv(L) from X to global memory (row-major 64x64), with
// zeros in the 16x16 blocks above the diagonal; one wavefront.
__device__ __forceinline__ void store_xrows(const double *X, double *Linv, int r0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int e = lane; e < 16 * (TB / 2); e += 64) {
        const int r = r0 + e / (TB / 2), c = 2 * (e % (TB / 2));
        dbl2 v;
        const bool keep = (r >> 4) >= (c >> 4);
        v.x = keep ? X[r * LDP + c] : 0.0;
        v.y = keep ? X[r * LDP + c + 1] : 0.0;
        *reinterpret_cast<dbl2 *>(Linv + r * TB + c) = v;
    }
}

// Columns 8G..8G+7 of this lane's g (group G = 0 or 1) to row[0..7]: 8
// columns of row d0 + lane of L (zeros above the diagonal: a lane above a
// pivot column takes 0 at that pivot)
__device__ __forceinline__ void store_lrow8(const double (&g)[16], int G, double *row) {
#pragma unroll
    for (int c = 0; c < 8; c += 2) {
        dbl2 v;
        v.x = g[8 * G + c];
        v.y = g[8 * G + c + 1];
        *reinterpret_cast<dbl2 *>(row + c) = v;
    }
}

// One panel step of the factor (all four wavefronts; ends with a barrier).
// Wavefront w owns panels w and w + 4 (columns 8w.. in g[0..7], 32 + 8w..
// in g[8..15]).  The owner of panel P applies panel P-1 to the panel's 8
// columns (the earlier panels are already in them), pivots it and stores
// its 8 columns of L.  The other wavefronts apply panel P-1 to their columns
// not yet pivoted; the owner of the previous step, which skipped its other
// group while pivoting, applies panel P-2 to that group first, so every
// column receives the panels in increasing order.  Beside this, the
// wavefronts with time left build inv(L), each piece as soon as its inputs
// are final: X_00 (step 2), S_10 = L_10 X_00 (3), X_11 (4), X_10 = -X_11
// S_10, S_21 = L_21 X_11 and the k = 1 term of S_31 (5), X_22, S_20 and the
// k < 2 terms of S_30 (6), X_20, X_21 and S_32 = L_32 X_22 (7); rows 0-15
// (5) and 16-31 (7) of inv(L) go to global memory.  S: the wavefront's
// 16x17 scratch (the four lie SW apart, and a sum left in one wavefront's
// scratch is read by another in a later step), S3: the partial sums S_3j.
template <int P>
__device__ __forceinline__ void factor_step(double (&g)[16], double *LP, double *ivs, double *X, double *S,
                                            double *S3, double &myinv, int &nneg, int w, double *lrow,
                                            double *Linv) {
    constexpr int O = P & 3, LOC = P >> 2, SW = 16 * 17;
    if (w == O) {
        if constexpr (P > 0) {
            double x[8];
            panel_row<P - 1>(LP, x);
            __builtin_amdgcn_sched_barrier(0);
            panel_update16<8 * LOC, 8>(g, x, LP + (P - 1) * (TB * 8), 8 * P - 8 * LOC);
        }
        pivot_panel<P, 8 * LOC>(g, LP, ivs, myinv, nneg);
        store_lrow8(g, LOC, lrow + 8 * P);
    } else {
        if constexpr (P > 0) {
            double x[8];
            panel_row<P - 1>(LP, x);
            __builtin_amdgcn_sched_barrier(0);
            if (w > P) panel_update16<0, 8>(g, x, LP + (P - 1) * (TB * 8), 8 * w);
            if (w + 4 > P) {
                if constexpr (P >= 2) {
                    if (w == P - 1) {
                        double x2[8];
                        panel_row<P - 2>(LP, x2);
                        __builtin_amdgcn_sched_barrier(0);
                        panel_update16<8, 8>(g, x2, LP + (P - 2) * (TB * 8), 8 * w + 24);
                    }
                }
                panel_update16<8, 8>(g, x, LP + (P - 1) * (TB * 8), 8 * w + 24);
            }
        }
        if constexpr (P == 2) {
            if (w == 0) diag_block_inverse(LP, ivs, X, S, 0);
        } else if constexpr (P == 3) {
            if (w == 1) x_sum(LP, X, S, 1, 0);
        } else if constexpr (P == 4) {
            if (w == 2) diag_block_inverse(LP, ivs, X, S, 1);
        } else if constexpr (P == 5) {
            if (w == 0) {
                x_final(X, S + SW, 1, 0);
                store_xrows(X, Linv, 0);
            } else if (w == 2) {
                x_sum(LP, X, S, 2, 1);
                s3_part(LP, X, S3 + SW, 1, 1, 2, false);
            }
        } else if constexpr (P == 6) {
            if (w == 0) {
                diag_block_inverse(LP, ivs, X, S, 2);
            } else if (w == 1) {
                x_sum(LP, X, S, 2, 0);
                s3_part(LP, X, S3, 0, 0, 2, false);
            }
        } else if constexpr (P == 7) {
            if (w == 0) {
                x_final(X, S + SW, 2, 0);
                store_xrows(X, Linv, 16);
            } else if (w == 1) {
                x_final(X, S + SW, 2, 1);
            } else {
                s3_part(LP, X, S3 + 2 * SW, 2, 2, 3, false);
            }
        }
    }
    __syncthreads();
}

// Called by all four wavefronts.  T: the tile (LDS, pitch LDP) at (d0, d0),
// free for scratch once the rows are in registers (ivs and the per-wavefront
// scratch of the inverse); LP: LDS, 64*64; X: LDS, 64*LDP.  L goes to A
// (lower, zeros above) and inv(L), row-major 64x64, to Linv.
__device__ __forceinline__ void diag_factor_inv(double *T, double *LP, double *X, double *A, int ld, int d0,
                                                double *Linv, int *neg) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double g[16];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        g[c] = T[lane * LDP + 8 * w + c];
        g[8 + c] = T[lane * LDP + 32 + 8 * w + c];
    }
    __syncthreads();
    double *ivs = T, *S = T + TB + w * (16 * 17), *S3 = T + TB + 4 * (16 * 17);
    double *lrow = A + (size_t)(d0 + lane) * ld + d0;
    int nneg = 0;
    double myinv = 0.0;
    factor_step<0>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    factor_step<1>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    factor_step<2>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    factor_step<3>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    factor_step<4>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    factor_step<5>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    factor_step<6>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    factor_step<7>(g, LP, ivs, X, S, S3, myinv, nneg, w, lrow, Linv);
    if (nneg && lane == 0) atomicAdd(neg, nneg);
    // X_33, the k = 2 terms of S_30 and S_31, and rows 32-47 of inv(L)
    if (w == 3) {
        diag_block_inverse(LP, ivs, X, S, 3);
    } else if (w == 0) {
        s3_part(LP, X, S3, 0, 2, 3, true);
    } else if (w == 1) {
        s3_part(LP, X, S3 + 16 * 17, 1, 2, 3, true);
    } else {
        store_xrows(X, Linv, 32);
    }
    __syncthreads();
    // X_3j = -X_33 S_3j straight from the matrix-core registers to rows
    // 48-63, columns 16j..16j+15 of inv(L) (register v of lane l: element
    // (l/16 + 4v, l%16))
    if (w < 3) {
        const dbl4 z = {0.0, 0.0, 0.0, 0.0};
        const dbl4 x3 = mm_xs(z, X, S3 + w * (16 * 17), 3);
        const int i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
        for (int v = 0; v < 4; v++) Linv[(48 + k4 + 4 * v) * TB + 16 * w + i16] = -x3[v];
    } else {
        // X_33 to rows 48-63, columns 48-63 of inv(L)
#pragma unroll
        for (int e = lane; e < 16 * 8; e += 64) {
            const int r = 48 + e / 8, c = 48 + 2 * (e % 8);
            dbl2 v;
            v.x = X[r * LDP + c];
            v.y = X[r * LDP + c + 1];
            *reinterpret_cast<dbl2 *>(Linv + r * TB + c) = v;
        }
    }
}


// Trailing tiles of block column kb restricted to the first jcols trailing
// block columns, row-major: rows ii = 0..M-1, columns jj <= min(ii, jcols-1)
// (jcols = M: the whole trailing triangle, in tri_index order).
__device__ __forceinline__ void trap_index(int idx, int jcols, int &i, int &j) {
    const int tri = jcols * (jcols + 1) / 2;
    if (idx < tri) {
        tri_index(idx, i, j);
        return;
    }
    idx -= tri;
    i = jcols + idx / jcols;
    j = idx % jcols;
}

// One block column kb of the factorisation in one launch, one workgroup per
// trailing lower tile (ib, jb), kb < jb <= ib, jb - kb <= jcols:
//   1. the panel tiles L_ib,kb = A_ib,kb inv(L_kb,kb)^T and L_jb,kb on the
//      FP64 matrix cores, from the inverse the previous launch left in
//      Linv; the workgroup of the first column (jb = kb + 1) stores L_ib,kb,
//      so every panel tile is written exactly once, into the upper block
//      (kb, ib), because other workgroups of this launch still read
//      A_ib,kb; the tiles stay there (k_noise, k_trail and the host copy
//      of the factor read them from there);
//   2. A_ib,jb -= L_ib,kb L_jb,kb^T on the matrix cores (FP64, or FP32 /
//      three-slice BF16 with FP32 accumulation for MODE 1 / 2);
//   3. the workgroup of tile (kb+1, kb+1) factors and inverts it, which the
//      next launch needs.
// The panel products are repeated by the workgroups of a block row, in
// exchange for one launch (and one kernel boundary) per block column.
// jcols < M restricts the update to the columns of a super-panel.
template <int MODE>
__global__ void __launch_bounds__(256) k_step(double *A, int ld, int kb, int jcols, double *Linv, int *neg) {
    // three tiles of LDS: inv(L_kb,kb), the row's panel block, and a third
    // holding the target tile in the diagonal tiles' workgroups (the factor's
    // T) and the column's panel block in the others, whose target tile is
    // updated in registers
    __shared__ double W[TB * LDP], As[TB * LDP], BC[TB * LDP];
    // workgroups go round-robin over the eight XCDs (b % 8); XCD x takes a
    // contiguous run of tiles, so the tiles of a block row, which read the
    // same panel tile, share an L2 (workgroup 0, the diagonal tile, stays 0)
    int lb = blockIdx.x;
    const int G8 = gridDim.x / 8;
    if (lb < 8 * G8) lb = (lb % 8) * G8 + lb / 8;
    int ii, jj;
    trap_index(lb, jcols, ii, jj);
    const int ib = kb + 1 + ii, jb = kb + 1 + jj;
    const bool diag = ib == jb;
    {
        // inv(L_kb,kb), the panel blocks (or the diagonal target tile), all in flight
        const double *const src[4] = {Linv + (size_t)kb * TB * TB, A + (size_t)ib * TB * ld + kb * TB,
                                      diag ? A + (size_t)ib * TB * ld + jb * TB : A + (size_t)jb * TB * ld + kb * TB,
                                      nullptr};
        const int lds[4] = {TB, ld, ld, 0};
        double *const dst[4] = {W, As, BC, nullptr};
        load_tiles(3, src, lds, dst);
    }
    __syncthreads();
    {
        dbl4 pi[4], pj[4];
        for (int J = 0; J < 4; J++) pi[J] = pj[J] = dbl4{0.0, 0.0, 0.0, 0.0};
        mma_panel(As, W, pi);
        if (!diag) mma_panel(BC, W, pj);
        __syncthreads();
        panel_to_lds(As, pi);
        if (!diag) panel_to_lds(BC, pj);
    }
    __syncthreads();
    if (jj == 0) store_tile(A, ld, kb * TB, ib * TB, As, false);
    if (diag) {
        if (MODE == 0) {
            syrk_lower(As, BC);
        } else {
            typename AccOf<MODE>::T acc[2][2];
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) acc[a][b] = {};
            tile_mma<MODE>(As, As, acc);
            acc_to_lds(BC, acc, -1.0, false);
        }
        __syncthreads();
        if (ii == 0)
            diag_factor_inv(BC, W, As, A, ld, ib * TB, Linv + (size_t)ib * TB * TB, neg);
        else
            store_tile(A, ld, ib * TB, jb * TB, BC, true);
        return;
    }
    typename AccOf<MODE>::T acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = {};
    tile_mma<MODE>(As, BC, acc);
    // C -= acc in place from registers (the element mapping of acc_to_lds)
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, i16 = lane & 15, k4 = lane >> 4;
    constexpr bool f64 = MODE == 0;
    double *C = A + (size_t)ib * TB * ld + jb * TB;
    double cv[2][2][4];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = wr + 16 * mi + (f64 ? k4 + 4 * v : 4 * k4 + v), c = wc + 16 * ni + i16;
                cv[mi][ni][v] = C[(size_t)r * ld + c];
            }
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = wr + 16 * mi + (f64 ? k4 + 4 * v : 4 * k4 + v), c = wc + 16 * ni + i16;
                C[(size_t)r * ld + c] = cv[mi][ni][v] + -1.0 * (double)acc[mi][ni][v];
            }
}

// The same block column in two launches, for wide trailing matrices (many
// tiles per launch): k_panel forms every panel tile L_ib,kb once and parks
// it in the upper block (kb, ib); k_update then updates each trailing tile
// from two parked panel tiles, and its workgroup of tile (kb+1, kb+1)
// factors and inverts it.  This trades one launch for the panel products
// that k_step repeats in every workgroup of a block row (and for a smaller
// LDS footprint: two panel workgroups fit on a CU).
__global__ void __launch_bounds__(256) k_panel(double *A, int ld, int kb, const double *Linv) {
    __shared__ double W[TB * LDP], As[TB * LDP];
    const int ib = kb + 1 + blockIdx.x;
    {
        const double *const src[4] = {Linv + (size_t)kb * TB * TB, A + (size_t)ib * TB * ld + kb * TB, nullptr,
                                      nullptr};
        const int lds[4] = {TB, ld, 0, 0};
        double *const dst[4] = {W, As, nullptr, nullptr};
        load_tiles(2, src, lds, dst);
    }
    __syncthreads();
    dbl4 pi[4];
    for (int J = 0; J < 4; J++) pi[J] = dbl4{0.0, 0.0, 0.0, 0.0};
    mma_panel(As, W, pi);
    __syncthreads();
    panel_to_lds(As, pi);
    __syncthreads();
    store_tile(A, ld, kb * TB, ib * TB, As, false);
}

template <int MODE>
__global__ void __launch_bounds__(256) k_update(double *A, int ld, int kb, int jcols, double *Linv, int *neg) {
    __shared__ double As[TB * LDP], Bs[TB * LDP], Cs[TB * LDP];
    int lb = blockIdx.x;  // XCD-contiguous runs of tiles, as in k_step
    const int G8 = gridDim.x / 8;
    if (lb < 8 * G8) lb = (lb % 8) * G8 + lb / 8;
    int ii, jj;
    trap_index(lb, jcols, ii, jj);
    const int ib = kb + 1 + ii, jb = kb + 1 + jj;
    const bool diag = ib == jb;
    {
        // the parked panel tiles L_ib,kb and L_jb,kb and the target tile
        const double *const src[4] = {A + (size_t)kb * TB * ld + ib * TB, A + (size_t)ib * TB * ld + jb * TB,
                                      A + (size_t)kb * TB * ld + jb * TB, nullptr};
        const int lds[4] = {ld, ld, ld, 0};
        double *const dst[4] = {As, Cs, Bs, nullptr};
        load_tiles(diag ? 2 : 3, src, lds, dst);
    }
    __syncthreads();
    if (MODE == 0 && diag) {
        syrk_lower(As, Cs);
    } else {
        typename AccOf<MODE>::T acc[2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) acc[a][b] = {};
        tile_mma<MODE>(As, diag ? As : Bs, acc);
        acc_to_lds(Cs, acc, -1.0, false);
    }
    __syncthreads();
    if (ii == 0 && jj == 0)
        diag_factor_inv(Cs, As, Bs, A, ld, ib * TB, Linv + (size_t)ib * TB * TB, neg);
    else
        store_tile(A, ld, ib * TB, jb * TB, Cs, diag);
}

// Trailing update of a super-panel: block columns k0 .. k0+nb-1 are
// factored and their panel tiles parked; every lower tile (ib, jb) of the
// trailing matrix (rows and columns from j0 = k0 + nb) takes
//   A_ib,jb -= sum_k L_ib,k L_jb,k^T   (K = 64 nb)
// with the sum held in registers across the nb panel tiles and the tile
// read and written once.  The next panel tiles are loaded into registers
// while the current ones are multiplied (the target tile during the last).
// Workgroups are grouped in 8x8 blocks of tiles, 64 consecutive workgroups
// per block, so the panel tiles of a block's 8 rows and 8 columns are
// shared through the L2; workgroups past the triangle or the edge exit.
constexpr int NTHR_T = 256;
__device__ __forceinline__ void tile_to_regs(const double *src, int ld, dbl2 (&v)[TB * TB / 2 / NTHR_T]) {
#pragma unroll
    for (int i = 0; i < TB * TB / 2 / NTHR_T; i++) {
        const int e = threadIdx.x + i * NTHR_T, r = e / (TB / 2), c = 2 * (e % (TB / 2));
        v[i] = *reinterpret_cast<const dbl2 *>(src + (size_t)r * ld + c);
    }
}
__device__ __forceinline__ void regs_to_lds(const dbl2 (&v)[TB * TB / 2 / NTHR_T], double *S) {
#pragma unroll
    for (int i = 0; i < TB * TB / 2 / NTHR_T; i++) {
        const int e = threadIdx.x + i * NTHR_T, r = e / (TB / 2), c = 2 * (e % (TB / 2));
        S[r * LDP + c] = v[i].x;
        S[r * LDP + c + 1] = v[i].y;
    }
}
template <int MODE>
__global__ void __launch_bounds__(256, 2) k_trail(double *A, int ld, int k0, int nb, int j0, int nt) {
    __shared__ double As[TB * LDP], Bs[TB * LDP];
    int SI, SJ;
    tri_index(blockIdx.x >> 6, SI, SJ);
    const int ib = j0 + 8 * SI + ((blockIdx.x >> 3) & 7), jb = j0 + 8 * SJ + (blockIdx.x & 7);
    if (ib >= nt || jb > ib) return;
    const bool diag = ib == jb;
    constexpr int PER = TB * TB / 2 / NTHR_T;
    dbl2 va[PER], vb[PER];
    typename AccOf<MODE>::T acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = {};
    // L_ib,k is parked at block (k, ib)
    tile_to_regs(A + (size_t)k0 * TB * ld + ib * TB, ld, va);
    if (!diag) tile_to_regs(A + (size_t)k0 * TB * ld + jb * TB, ld, vb);
    for (int t = 0; t < nb; t++) {
        regs_to_lds(va, As);
        if (!diag) regs_to_lds(vb, Bs);
        __syncthreads();
        if (t + 1 < nb) {
            const size_t row = (size_t)(k0 + t + 1) * TB * ld;
            tile_to_regs(A + row + ib * TB, ld, va);
            if (!diag) tile_to_regs(A + row + jb * TB, ld, vb);
        } else {
            tile_to_regs(A + (size_t)ib * TB * ld + jb * TB, ld, va);
        }
        tile_mma<MODE>(As, diag ? As : Bs, acc);
        __syncthreads();
    }
    regs_to_lds(va, As);
    __syncthreads();
    acc_to_lds(As, acc, -1.0, false);
    __syncthreads();
    store_tile(A, ld, ib * TB, jb * TB, As, diag);
}

// Three-slice BF16 trailing update with the panel tiles split once.  Split
// on the fly (k_trail<2>), every workgroup converts every operand fragment it
// multiplies, so each panel tile is split by all the workgroups of its tile
// row and column and the update is bound by that VALU work.  Here k_split3
// writes the three BF16 slices of every parked panel tile of the super-panel
// once, in the operand order of v_mfma_f32_16x16x32_bf16: fragment (rb, ks)
// of a slice (16 rows from 16 rb, 32 K from 32 ks) is 64 lanes x 16 bytes,
// lane l holding row 16 rb + l % 16 and K entries 32 ks + 8 (l / 16) .. +7.
// Slice s of the tile of row block i and panel column t sits at
// sp[((i - j0) nb + t) 3 + s].
constexpr size_t kSplitTile = (size_t)TB * TB;
__device__ __forceinline__ size_t frag_off(int rb, int ks, int lane) { return (size_t)((rb * 2 + ks) * 64 + lane) * 8; }

// one workgroup per (row block, panel column) tile; wavefront w writes the
// fragments of row block w
__global__ void __launch_bounds__(256) k_split3(const double *A, int ld, int k0, int nb, int j0, __bf16 *sp) {
    const int i = j0 + blockIdx.x / nb, t = blockIdx.x % nb;
    const double *src = A + (size_t)(k0 + t) * TB * ld + (size_t)i * TB;
    __bf16 *dst = sp + (size_t)blockIdx.x * 3 * kSplitTile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
        const int r = 16 * w + (lane & 15), c = 32 * ks + 8 * (lane >> 4);
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            const dbl2 x = *reinterpret_cast<const dbl2 *>(src + (size_t)r * ld + c + q);
            v[q] = x.x;
            v[q + 1] = x.y;
        }
        bf8 h, m, l;
        split3(v, h, m, l);
        const size_t o = frag_off(w, ks, lane);
        *reinterpret_cast<bf8 *>(dst + o) = h;
        *reinterpret_cast<bf8 *>(dst + kSplitTile + o) = m;
        *reinterpret_cast<bf8 *>(dst + 2 * kSplitTile + o) = l;
    }
}

// k_trail_bf: a workgroup takes a 2x2 block of trailing tiles (128x128) and
// wavefront (wr, wc) the whole tile (ib0 + wr, jb0 + wc).  A stage is one K
// step (32) of a panel column: the block's two row tiles' and two column
// tiles' fragments, three slices each, 48 kilobyte slots (op, tile, s, rb),
// copied straight from global memory into LDS (global_load_lds_dwordx4:
// lane l's 16 bytes land at the slot + 16 l, so a fragment lies in LDS as in
// global memory and every wavefront read of it is conflict-free).  The
// copies are what bounds this kernel, so the tile is as large as the
// registers allow (a wavefront reads 24 fragments per stage for 96
// products), and three workgroups share a CU (3 x 48 KB of LDS) to keep
// three stages in flight.  The six products of a 16x16 block, their order
// and the split are those of k_trail<2>: the same factor bit for bit.
typedef __attribute__((address_space(3))) void lds_void;
__device__ __forceinline__ int bf_slot(int op, int tile, int s, int rb) { return ((op * 2 + tile) * 3 + s) * 4 + rb; }
__device__ __forceinline__ const __bf16 *bf_src(const __bf16 *sp, int row, int nb, int t, int s, int rb, int ks,
                                                int lane) {
    return sp + ((size_t)(row * nb + t) * 3 + s) * kSplitTile + frag_off(rb, ks, lane);
}
// Stage st (panel column st / 2, K step st % 2) of tile rows ra, ra + 1 and
// (unless DIAG, where the column tiles are the row tiles) rbw, rbw + 1,
// relative to j0; rows at or past nrow (the odd edge) are not copied.
// Wavefront w copies one tile's 12 fragments (DIAG: half of one tile's).
template <bool DIAG>
__device__ __forceinline__ void bf_issue(const __bf16 *sp, int ra, int rbw, int nb, int nrow, int st, __bf16 *S, int w,
                                         int lane) {
    const int t = st >> 1, ks = st & 1;
    if (DIAG) {
        const int tile = w >> 1, half = w & 1;
        if (ra + tile >= nrow) return;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            const int f = 6 * half + j, sl = f >> 2, rb = f & 3;
            __builtin_amdgcn_global_load_lds(bf_src(sp, ra + tile, nb, t, sl, rb, ks, lane),
                                             (lds_void *)(S + (size_t)bf_slot(0, tile, sl, rb) * 512), 16, 0, 0);
        }
    } else {
        const int op = w >> 1, tile = w & 1, row = (op ? rbw : ra) + tile;
        if (row >= nrow) return;
#pragma unroll
        for (int f = 0; f < 12; f++) {
            const int sl = f >> 2, rb = f & 3;
            __builtin_amdgcn_global_load_lds(bf_src(sp, row, nb, t, sl, rb, ks, lane),
                                             (lds_void *)(S + (size_t)bf_slot(op, tile, sl, rb) * 512), 16, 0, 0);
        }
    }
}
// the six products of column block ni for the four row blocks, each
// block's in the fixed order (smallest first)
__device__ __forceinline__ void bf_mma_col(const bf8 (&fa)[3][4], const bf8 (&fb)[3], flt4 (&acc)[4][4], int ni) {
    constexpr int PA[6] = {2, 1, 0, 1, 0, 0}, PB[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
        for (int mi = 0; mi < 4; mi++)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[PA[p]][mi], fb[PB[p]], acc[mi][ni], 0, 0, 0);
}
__device__ __forceinline__ void bf_read_col(const __bf16 *S, int op, int tile, int q, int lane, bf8 (&fb)[3]) {
#pragma unroll
    for (int sl = 0; sl < 3; sl++)
        fb[sl] = *reinterpret_cast<const bf8 *>(S + (size_t)bf_slot(op, tile, sl, q) * 512 + lane * 8);
}
// Per stage: every wavefront waits for its own copies (vmcnt(0); the LDS
// writes of a copy are done when it retires) and a barrier makes the stage
// whole; the row fragments (12) and two column blocks' (3 each) are read,
// the products of column block 0 run while block 2's fragments are read into
// block 0's registers (1 and 3 likewise); after lgkmcnt(0) a second barrier
// frees the stage, the copy of the next stage is issued, and the products
// of column blocks 2 and 3 run beside it.  The compiler does not order the
// copies against these LDS reads, hence the explicit waits.  Wavefronts
// with no tile (above the diagonal, past the edge) still copy and meet
// every barrier.
template <bool DIAG>
__device__ __forceinline__ void bf_body(const __bf16 *sp, int ra, int rbw, int nb, int nrow, __bf16 *S, int w, int lane,
                                        bool active, flt4 (&acc)[4][4]) {
    constexpr int BOP = DIAG ? 0 : 1;
    const int wr = w >> 1, wc = w & 1;
    bf8 fa[3][4], fb0[3], fb1[3];
    bf_issue<DIAG>(sp, ra, rbw, nb, nrow, 0, S, w, lane);
    for (int st = 0; st < 2 * nb; st++) {
        __builtin_amdgcn_s_waitcnt(0xF70);   // vmcnt(0)
        __builtin_amdgcn_s_barrier();
        if (active) {
#pragma unroll
            for (int sl = 0; sl < 3; sl++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    fa[sl][q] = *reinterpret_cast<const bf8 *>(S + (size_t)bf_slot(0, wr, sl, q) * 512 + lane * 8);
            bf_read_col(S, BOP, wc, 0, lane, fb0);
            bf_read_col(S, BOP, wc, 1, lane, fb1);
            bf_mma_col(fa, fb0, acc, 0);
            bf_read_col(S, BOP, wc, 2, lane, fb0);
            bf_mma_col(fa, fb1, acc, 1);
            bf_read_col(S, BOP, wc, 3, lane, fb1);
        }
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        if (st + 1 < 2 * nb) bf_issue<DIAG>(sp, ra, rbw, nb, nrow, st + 1, S, w, lane);
        if (active) {
            bf_mma_col(fa, fb0, acc, 2);
            bf_mma_col(fa, fb1, acc, 3);
        }
    }
}
// Workgroups go round-robin over the eight XCDs (b % 8); XCD x takes groups
// x, x + 8, ... of 8x8 blocks, 64 consecutive workgroups of its own per
// group, so a group's 16 row and 16 column panel tiles are read into one L2.
__global__ void __launch_bounds__(256, 3) k_trail_bf(double *A, int ld, const __bf16 *sp, int nb, int j0, int nt) {
    __shared__ __attribute__((aligned(16))) __bf16 S[48 * 512];
    const int S2 = (nt - j0 + 1) / 2, SG = (S2 + 7) / 8, l = blockIdx.x >> 3, G = (l >> 6) * 8 + (blockIdx.x & 7);
    if (G >= SG * (SG + 1) / 2) return;
    int GI, GJ;
    tri_index(G, GI, GJ);
    const int IB = 8 * GI + ((l >> 3) & 7), JB = 8 * GJ + (l & 7);
    if (IB >= S2 || JB > IB) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    const int nrow = nt - j0, ra = 2 * IB, rbw = 2 * JB;  // tile rows relative to j0
    const int ib = j0 + ra + wr, jb = j0 + rbw + wc;
    const bool active = ib < nt && jb <= ib;
    flt4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = flt4{0.f, 0.f, 0.f, 0.f};
    if (IB == JB) bf_body<true>(sp, ra, rbw, nb, nrow, S, w, lane, active, acc);
    else bf_body<false>(sp, ra, rbw, nb, nrow, S, w, lane, active, acc);
    if (!active) return;
    const int i16 = lane & 15, k4 = lane >> 4;
    // C -= acc in place, a row of 16x16 blocks at a time: loads, then stores
    // (a diagonal tile keeps its upper triangle)
#pragma unroll
    for (int mi = 0; mi < 4; mi++) {
        double cv[4][4];
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * mi + 4 * k4 + v, c = 16 * ni + i16;
                cv[ni][v] = A[(size_t)(ib * TB + r) * ld + jb * TB + c];
            }
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * mi + 4 * k4 + v, c = 16 * ni + i16;
                if (ib == jb && c > r) continue;
                A[(size_t)(ib * TB + r) * ld + jb * TB + c] = cv[ni][v] + -1.0 * (double)acc[mi][ni][v];
            }
    }
}

// FP64 / FP32 trailing update of a super-panel from panel tiles packed once.
// k_pack64 copies each parked panel tile of the super-panel into the operand
// order of v_mfma_f64_16x16x4f64 (v_mfma_f32_16x16x4f32 has the same lane
// layout): unit (rb, kp) of a tile is 64 lanes x 16 bytes, lane l holding
// row 16 rb + l % 16 at K = 8 kp + l / 16 and 8 kp + 4 + l / 16, the
// operands of K steps 2 kp and 2 kp + 1.  The tile of row block i and panel
// column t sits at pk[((i - j0) nb + t) TB TB].
constexpr size_t kPackTile = (size_t)TB * TB;  // doubles
__device__ __forceinline__ size_t pk_off(int rb, int kp, int lane) { return (size_t)((rb * 8 + kp) * 64 + lane) * 2; }

// one workgroup per (row block, panel column) tile; wavefront w writes the
// units of row block w
__global__ void __launch_bounds__(256) k_pack64(const double *A, int ld, int k0, int nb, int j0, double *pk) {
    const int i = j0 + blockIdx.x / nb, t = blockIdx.x % nb;
    const double *src = A + (size_t)(k0 + t) * TB * ld + (size_t)i * TB;
    double *dst = pk + (size_t)blockIdx.x * kPackTile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = 16 * w + (lane & 15), k4 = lane >> 4;
#pragma unroll
    for (int kp = 0; kp < 8; kp++) {
        dbl2 v;
        v.x = src[(size_t)r * ld + 8 * kp + k4];
        v.y = src[(size_t)r * ld + 8 * kp + 4 + k4];
        *reinterpret_cast<dbl2 *>(dst + pk_off(w, kp, lane)) = v;
    }
}

// k_trail_pk: a workgroup takes a 2x2 block of trailing tiles (128x128),
// wavefront (wr, wc) the whole tile (ib0 + wr, jb0 + wc), as k_trail_bf
// does.  A stage is 16 of a panel column's K: units (rb, kp) of the
// block's two row and two column tiles, 32 KB in slots (op, tile, rb, kp),
// copied straight from global memory into LDS.  Two stages per workgroup:
// the copy of stage s + 1 is issued before the products of stage s, so it
// has a whole stage to land; two workgroups per CU (2 x 64 KB of LDS).
// Every accumulator takes the MFMAs of k_trail<MODE> in the same order (panel
// column, then K step), so the update is the same bit for bit.
__device__ __forceinline__ int pk_slot(int op, int tile, int rb, int kpl) { return ((op * 2 + tile) * 4 + rb) * 2 + kpl; }
template <bool DIAG>
__device__ __forceinline__ void pk_issue(const double *pk, int ra, int rbw, int nb, int nrow, int st, double *S, int w,
                                         int lane) {
    const int t = st >> 2, h = st & 3;
    if (DIAG) {
        const int tile = w >> 1, half = w & 1;
        if (ra + tile >= nrow) return;
        const double *src = pk + (size_t)((ra + tile) * nb + t) * kPackTile;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int rb = 2 * half + (j >> 1), kpl = j & 1;
            __builtin_amdgcn_global_load_lds(src + pk_off(rb, 2 * h + kpl, lane),
                                             (lds_void *)(S + (size_t)pk_slot(0, tile, rb, kpl) * 128), 16, 0, 0);
        }
    } else {
        const int op = w >> 1, tile = w & 1, row = (op ? rbw : ra) + tile;
        if (row >= nrow) return;
        const double *src = pk + (size_t)(row * nb + t) * kPackTile;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int rb = j >> 1, kpl = j & 1;
            __builtin_amdgcn_global_load_lds(src + pk_off(rb, 2 * h + kpl, lane),
                                             (lds_void *)(S + (size_t)pk_slot(op, tile, rb, kpl) * 128), 16, 0, 0);
        }
    }
}
template <int MODE, typename V>
__device__ __forceinline__ V pk_mfma(double a, double b, V c) {
    if constexpr (MODE == 0) return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, c, 0, 0, 0);
}
template <int MODE, bool DIAG, typename V>
__device__ __forceinline__ void pk_body(const double *pk, int ra, int rbw, int nb, int nrow, double *S, int w, int lane,
                                        bool active, V (&acc)[4][4]) {
    constexpr int BOP = DIAG ? 0 : 1;
    const int wr = w >> 1, wc = w & 1, ns = 4 * nb;
    pk_issue<DIAG>(pk, ra, rbw, nb, nrow, 0, S, w, lane);
    for (int st = 0; st < ns; st++) {
        const double *B = S + (size_t)(st & 1) * 32 * 128;
        if (st + 1 < ns) {
            pk_issue<DIAG>(pk, ra, rbw, nb, nrow, st + 1, S + (size_t)((st + 1) & 1) * 32 * 128, w, lane);
            // this wavefront's copies of stage st are done, those of st + 1
            // may be in flight (vector memory counts retire in order)
            if (DIAG) __builtin_amdgcn_s_waitcnt(0xF74);  // vmcnt(4)
            else __builtin_amdgcn_s_waitcnt(0xF78);       // vmcnt(8)
        } else {
            __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
        }
        __builtin_amdgcn_s_barrier();
        if (active) {
            dbl2 a[2][4], b[2][4];
#pragma unroll
            for (int kpl = 0; kpl < 2; kpl++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    a[kpl][q] = *reinterpret_cast<const dbl2 *>(B + (size_t)pk_slot(0, wr, q, kpl) * 128 + lane * 2);
                    b[kpl][q] = *reinterpret_cast<const dbl2 *>(B + (size_t)pk_slot(BOP, wc, q, kpl) * 128 + lane * 2);
                }
#pragma unroll
            for (int kpl = 0; kpl < 2; kpl++) {
#pragma unroll
                for (int mi = 0; mi < 4; mi++)
#pragma unroll
                    for (int ni = 0; ni < 4; ni++) acc[mi][ni] = pk_mfma<MODE>(a[kpl][mi].x, b[kpl][ni].x, acc[mi][ni]);
#pragma unroll
                for (int mi = 0; mi < 4; mi++)
#pragma unroll
                    for (int ni = 0; ni < 4; ni++) acc[mi][ni] = pk_mfma<MODE>(a[kpl][mi].y, b[kpl][ni].y, acc[mi][ni]);
            }
        }
        // every wavefront has read stage st before its slots take stage st + 2
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
    }
}
template <int MODE>
__global__ void __launch_bounds__(256, 2) k_trail_pk(double *A, int ld, const double *pk, int nb, int j0, int nt) {
    __shared__ __attribute__((aligned(16))) double S[2 * 32 * 128];
    const int S2 = (nt - j0 + 1) / 2, SG = (S2 + 7) / 8, l = blockIdx.x >> 3, G = (l >> 6) * 8 + (blockIdx.x & 7);
    if (G >= SG * (SG + 1) / 2) return;
    int GI, GJ;
    tri_index(G, GI, GJ);
    const int IB = 8 * GI + ((l >> 3) & 7), JB = 8 * GJ + (l & 7);
    if (IB >= S2 || JB > IB) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1;
    const int nrow = nt - j0, ra = 2 * IB, rbw = 2 * JB;  // tile rows relative to j0
    const int ib = j0 + ra + wr, jb = j0 + rbw + wc;
    const bool active = ib < nt && jb <= ib;
    typedef typename AccOf<MODE>::T V;
    V acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = V{};
    if (IB == JB) pk_body<MODE, true>(pk, ra, rbw, nb, nrow, S, w, lane, active, acc);
    else pk_body<MODE, false>(pk, ra, rbw, nb, nrow, S, w, lane, active, acc);
    if (!active) return;
    const int i16 = lane & 15, k4 = lane >> 4;
    // C -= acc in place (row of register v: l/16 + 4v for the FP64 MFMA,
    // 4(l/16) + v for the FP32 one), a row of 16x16 blocks at a time
#pragma unroll
    for (int mi = 0; mi < 4; mi++) {
        double cv[4][4];
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * mi + (MODE == 0 ? k4 + 4 * v : 4 * k4 + v), c = 16 * ni + i16;
                cv[ni][v] = A[(size_t)(ib * TB + r) * ld + jb * TB + c];
            }
#pragma unroll
        for (int ni = 0; ni < 4; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = 16 * mi + (MODE == 0 ? k4 + 4 * v : 4 * k4 + v), c = 16 * ni + i16;
                if (ib == jb && c > r) continue;
                A[(size_t)(ib * TB + r) * ld + jb * TB + c] = cv[ni][v] + -1.0 * (double)acc[mi][ni][v];
            }
    }
}

// Factor and inverse of the diagonal tile (d, d) by one workgroup (the
// tile after a super-panel's trailing update, or the first tile of a
// caller's matrix); clear != 0 also restarts the negative-pivot count.
__global__ void __launch_bounds__(256) k_diag(double *A, int ld, int d, double *Linv, int *neg, int clear) {
    __shared__ double T[TB * LDP], LP[TB * TB], X[TB * LDP];
    if (clear && threadIdx.x == 0) *neg = 0;
    load_tile(A, ld, d * TB, d * TB, T);
    __syncthreads();
    diag_factor_inv(T, LP, X, A, ld, d * TB, Linv + (size_t)d * TB * TB, neg);
}

// rd = s2dt * L * p (rd_cal, func_cuda.cu:634-647): one wavefront per row,
// lane k % 64 over the row; rows outside the hydrodynamic set use their
// self-mobility: s2dt * sqrt(a1 / rad) * p.  With du = D inv(L~)^T p (the
// FP64 refinement of a reduced-precision factor) the HI rows take
// s2dt (L~ p + du) / 2 instead.  L is read where the factor
// leaves it: the diagonal tile of the row in place, the off-diagonal tile
// (ib, t) parked in the upper block (t, ib) (a row of the parked tile is
// contiguous, so every read is one 512-byte segment).
__global__ void __launch_bounds__(256) k_noise(const double *A, int ld, int m, int nD3, const double *p,
                                               const double *du, const double *rad_all, double s2dt, double a1,
                                               double *rd) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nD3) return;
    if (row >= m) {
        if (lane == 0) rd[row] = s2dt * sqrt(a1 / rad_all[row / 3]) * p[row];
        return;
    }
    const int ib = row / TB, r = row % TB;
    const double *a = A + (size_t)r * ld + ib * TB + lane;  // parked tile (t, ib): a + t*TB*ld
    const double *q = p + lane;
    double v = 0.0;
    int t = 0;
    // eight tiles' loads in flight at a time, summed in tile order
    for (; t + 8 <= ib; t += 8) {
        double x[8], y[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            x[u] = a[(size_t)(t + u) * TB * ld];
            y[u] = q[(t + u) * TB];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) v += x[u] * y[u];
    }
    for (; t < ib; t++) v += a[(size_t)t * TB * ld] * q[t * TB];
    if (lane <= r) v += A[(size_t)row * ld + ib * TB + lane] * p[ib * TB + lane];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) rd[row] = s2dt * (du ? 0.5 * (v + du[row]) : v);
}

// ---------------------------------------------------------------------------
// FP64 refinement of B xi for a factor whose trailing updates ran in FP32 or
// BF16.  With L~ L~^T = D + E and M = inv(L~) D inv(L~)^T = I + X (X small
// and symmetric), B = L~ M^(1/2) satisfies B B^T = D exactly, and to first
// order B xi = L~ (xi + X xi / 2).  Since L~ X xi = D u - L~ xi for
// u = inv(L~)^T xi (L~^T u = xi), that is
//   B xi = (L~ xi + D u) / 2,   B B^T - D = L~ X^2 L~^T / 4 + O(X^3):
// one FP64 triangular solve with the factor (k_bsolve), one matrix-free
// product D u from the bead positions D was assembled from (k_dmul, the RPY
// blocks of k_rpy), and the mean taken in k_noise beside L~ xi.
// ---------------------------------------------------------------------------
__global__ void k_hi_snapshot(Stat s, Geo g, int nhi, double *xs) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nhi) return;
    const double3 r = hi_pos(s, g, b);
    xs[3 * b] = r.x;
    xs[3 * b + 1] = r.y;
    xs[3 * b + 2] = r.z;
}

__global__ void k_refine_init(const double *p, int m, int mpad, double *pw) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < mpad) pw[r] = r < m ? p[r] : 0.0;
}

// Column part of a 64x64 tile T (global, leading dimension ldt) for this
// thread: column e = t % 64, rows 16q .. 16q+15, q = t / 64 (a wavefront
// reads one row per load), loaded into registers ahead of use.
__device__ __forceinline__ void col_load(const double *T, size_t ldt, double (&v)[16]) {
    const int e = threadIdx.x & 63, q = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 16; j++) v[j] = T[(16 * q + j) * ldt + e];
}
// part[q][e] = sum_j v[j] x[16q + j], then y[e] = sum_q part[q][e] in a
// fixed order for e < 64 (threads 0..63), with barriers around.
__device__ __forceinline__ void col_reduce(const double (&v)[16], const double *x, double *part, double *y,
                                           const double *sub) {
    const int t = threadIdx.x, e = t & 63, q = t >> 6;
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 16; j++) a += v[j] * x[16 * q + j];
    part[q * TB + e] = a;
    __syncthreads();
    if (t < TB) {
        const double sum = ((part[t] + part[TB + t]) + part[2 * TB + t]) + part[3 * TB + t];
        y[t] = sub ? sub[t] - sum : sum;
    }
    __syncthreads();
}

// Blocks b0 = 2g and b1 = 2g + 1 (if present) of u = inv(L~)^T xi; the
// blocks after them were solved by earlier launches, so pw there is final:
//   u_b1 = inv(L~_b1b1)^T pw_b1,  u_b0 = inv(L~_b0b0)^T (pw_b0 - L~_b1b0^T u_b1)
// in every workgroup (all tiles loaded at the start); workgroup 0 stores
// u_b0 and u_b1, workgroup k >= 1 subtracts L~_b0,i^T u_b0 + L~_b1,i^T u_b1
// from pw_i for tile row i = k - 1 < b0 (L~_b,i parked in block (i, b)).
__global__ void __launch_bounds__(256) k_bsolve(const double *A, int ld, const double *Linv, int nt, int g,
                                                double *pw, double *u) {
    __shared__ double x[2 * TB], u0[TB], u1[TB], part[4 * TB];
    const int t = threadIdx.x, b0 = 2 * g, two = b0 + 1 < nt, i = blockIdx.x - 1;
    double la[16], lc[16], lb[16], r0[16], r1[16];
    if (two) {
        col_load(Linv + (size_t)(b0 + 1) * TB * TB, TB, la);
        col_load(A + (size_t)b0 * TB * ld + (b0 + 1) * TB, ld, lc);
    }
    col_load(Linv + (size_t)b0 * TB * TB, TB, lb);
    if (i >= 0) {
        col_load(A + (size_t)i * TB * ld + b0 * TB, ld, r0);
        if (two) col_load(A + (size_t)i * TB * ld + (b0 + 1) * TB, ld, r1);
    }
    if (t < (two ? 2 : 1) * TB) x[t] = pw[b0 * TB + t];
    __syncthreads();
    if (two) {
        col_reduce(la, x + TB, part, u1, nullptr);
        col_reduce(lc, u1, part, x, x);  // pw_b0 - L~_b1b0^T u_b1, in place
    }
    col_reduce(lb, x, part, u0, nullptr);
    if (i < 0) {
        if (t < TB) u[b0 * TB + t] = u0[t];
        else if (two && t < 2 * TB) u[b0 * TB + t] = u1[t - TB];
        return;
    }
    const int e = t & 63, q = t >> 6;
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 16; j++) a += r0[j] * u0[16 * q + j];
    if (two)
#pragma unroll
        for (int j = 0; j < 16; j++) a += r1[j] * u1[16 * q + j];
    part[q * TB + e] = a;
    __syncthreads();
    if (t < TB) pw[i * TB + t] -= ((part[t] + part[TB + t]) + part[2 * TB + t]) + part[3 * TB + t];
}

// w = D u for the hydrodynamic set (m = 3 nhi rows): one wavefront per bead
// i, lanes over the beads j, the 3x3 blocks of k_rpy (a1/rad_i on the
// diagonal), lane sums combined in a fixed order.
__global__ void __launch_bounds__(256) k_dmul(const double *xs, const double *rad_all, int nhi, double a1, double a2,
                                              const double *u, double *w) {
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nhi) return;
    const double3 ri = make_double3(xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]);
    const double radi = rad_all[i];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int j = lane; j < nhi; j += 64) {
        const double uj[3] = {u[3 * j], u[3 * j + 1], u[3 * j + 2]};
        if (j == i) {
            const double d = a1 / radi;
#pragma unroll
            for (int k = 0; k < 3; k++) acc[k] += d * uj[k];
        } else {
            double M[3][3];
            rpy_block(ri, make_double3(xs[3 * j], xs[3 * j + 1], xs[3 * j + 2]), radi, rad_all[j], a1, a2, M);
#pragma unroll
            for (int k = 0; k < 3; k++) acc[k] += (M[k][0] * uj[0] + M[k][1] * uj[1]) + M[k][2] * uj[2];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_down(acc[k], o, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 3; k++) w[3 * i + k] = acc[k];
}

// ---------------------------------------------------------------------------
void launch_rpy(const Stat &s, int n_hi, const Geo &g, const double *rad_all, const CholWork &c, hipStream_t q,
                bool factor_first) {
    const int nt = c.mpad / TB;
    // the positions D is built from, kept only where a refinement can read them
    if (refine_active(c)) k_hi_snapshot<<<nblk(c.m / 3, 256), 256, 0, q>>>(s, g, c.m / 3, c.xs);
    // the first tile's factor inside the assembly launch saves a launch on
    // the factor's chain, but gives every workgroup the factor's LDS (one
    // per CU): from 64 block columns the assembly takes its own launch
    // (four workgroups per CU) and k_diag the first tile
    if (factor_first && nt < 64) {
        k_rpy<true><<<nt * (nt + 1) / 2, 256, 0, q>>>(s, g, rad_all, c.A, c.m, c.ld, s.a1, s.a2, c.neg, c.Linv);
    } else {
        k_rpy<false><<<nt * (nt + 1) / 2, 256, 0, q>>>(s, g, rad_all, c.A, c.m, c.ld, s.a1, s.a2, c.neg, c.Linv);
        if (factor_first) launch_diag(c, 0, false, q);
    }
    (void)n_hi;
}

namespace {
template <int MODE>
void chol_column(const CholWork &c, int kb, int jcols, hipStream_t q) {
    const int nt = c.mpad / TB, M = nt - kb - 1;
    const int grid = jcols * (jcols + 1) / 2 + (M - jcols) * jcols;
    if (jcols == 0) {
        k_panel<<<M, 256, 0, q>>>(c.A, c.ld, kb, c.Linv);
    } else if (grid >= c.split_tiles) {
        k_panel<<<M, 256, 0, q>>>(c.A, c.ld, kb, c.Linv);
        k_update<MODE><<<grid, 256, 0, q>>>(c.A, c.ld, kb, jcols, c.Linv, c.neg);
    } else {
        k_step<MODE><<<grid, 256, 0, q>>>(c.A, c.ld, kb, jcols, c.Linv, c.neg);
    }
}
template <int MODE>
void chol_trail(const CholWork &c, int k0, int nb, hipStream_t q) {
    const int nt = c.mpad / TB, j0 = k0 + nb;
    // the 128x128 blocks pay in FP64 from 192 trailing block rows, in FP32
    // from about 96 (fewer blocks than workgroup slots below that); in BF16
    // always (the direct form splits every operand fragment in every
    // workgroup)
    const int min_rows = c.pack_min_rows >= 0 ? c.pack_min_rows : MODE == 0 ? 192 : MODE == 1 ? 96 : 0;
    if (c.sp && nb <= c.sp_nb && nt - j0 >= min_rows) {
        // panel tiles packed (split) once, 128x128 blocks in 8x8 groups
        const int S2 = (nt - j0 + 1) / 2, SG = (S2 + 7) / 8, grid = (SG * (SG + 1) / 2 + 7) / 8 * 512;
        if constexpr (MODE == 2) {
            __bf16 *sp = static_cast<__bf16 *>(c.sp);
            k_split3<<<(nt - j0) * nb, 256, 0, q>>>(c.A, c.ld, k0, nb, j0, sp);
            k_trail_bf<<<grid, 256, 0, q>>>(c.A, c.ld, sp, nb, j0, nt);
        } else {
            double *pk = static_cast<double *>(c.sp);
            k_pack64<<<(nt - j0) * nb, 256, 0, q>>>(c.A, c.ld, k0, nb, j0, pk);
            k_trail_pk<MODE><<<grid, 256, 0, q>>>(c.A, c.ld, pk, nb, j0, nt);
        }
        return;
    }
    const int SM = (nt - j0 + 7) / 8;
    k_trail<MODE><<<SM * (SM + 1) / 2 * 64, 256, 0, q>>>(c.A, c.ld, k0, nb, j0, nt);
}
}  // namespace

size_t split_bytes(int ld, int nb) {
    return std::max((size_t)(ld / TB) * nb * 3 * kSplitTile * sizeof(__bf16),
                    (size_t)(ld / TB) * nb * kPackTile * sizeof(double));
}

int chol_panel_tiles(const CholWork &c) {
    const int nt = c.mpad / TB;
    if (c.panel_tiles > 0) return c.panel_tiles >= 2 && nt > c.panel_tiles ? c.panel_tiles : 0;
    if (c.panel_tiles < 0) return 0;
    return nt >= 64 ? 8 : 0;
}

void launch_cholesky(const CholWork &c, hipStream_t q, MarkFn mark, void *mark_arg, hipEvent_t gate, int gate_kb) {
    const int nt = c.mpad / TB;
    auto M_ = [&](int id, int ph) { if (mark) mark(mark_arg, id, ph); };
    auto column = [&](int kb, int jcols, int mode) {
        M_(K_CHOL_STEP, 0);
        if (mode == 0) chol_column<0>(c, kb, jcols, q);
        else if (mode == 1) chol_column<1>(c, kb, jcols, q);
        else chol_column<2>(c, kb, jcols, q);
        M_(K_CHOL_STEP, 1);
        if (gate && kb == gate_kb) (void)hipEventRecord(gate, q);
    };
    const int NB = chol_panel_tiles(c);
    int k0 = 0;
    if (NB > 0) {
        // super-panels while a whole one and a trailing matrix remain; the
        // panel's own columns are factored in FP64 in every mode (the
        // reduced-precision modes apply to the trailing update k_trail)
        for (; k0 + NB < nt; k0 += NB) {
            const int kend = k0 + NB;
            for (int kb = k0; kb < kend; kb++) column(kb, kend - 1 - kb, 0);
            M_(K_CHOL_TRAIL, 0);
            if (c.mode == 0) chol_trail<0>(c, k0, NB, q);
            else if (c.mode == 1) chol_trail<1>(c, k0, NB, q);
            else chol_trail<2>(c, k0, NB, q);
            M_(K_CHOL_TRAIL, 1);
            M_(K_CHOL_DIAG, 0);
            k_diag<<<1, 256, 0, q>>>(c.A, c.ld, kend, c.Linv, c.neg, 0);
            M_(K_CHOL_DIAG, 1);
        }
    }
    // one launch (or two) per block column, each updating the whole trailing matrix
    for (int kb = k0; kb + 1 < nt; kb++) column(kb, nt - kb - 1, c.mode);
    if (gate && gate_kb >= nt - 1) (void)hipEventRecord(gate, q);
}

void launch_diag(const CholWork &c, int d, bool clear, hipStream_t q) {
    k_diag<<<1, 256, 0, q>>>(c.A, c.ld, d, c.Linv, c.neg, clear ? 1 : 0);
}

// A caller's matrix into the factor's storage: the lower triangle of D
// (row-major, leading dimension ldd), zeros above the diagonal and the
// identity on the padding rows; and back: the factor, unparked (the
// off-diagonal tile (ib, kb) lives in the upper block (kb, ib)), into L
// with zeros above the diagonal.
__global__ void k_pack(const double *__restrict__ D, int ldd, int m, double *__restrict__ A, int ld, int mpad) {
    const long long total = (long long)mpad * mpad;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(e / mpad), c = (int)(e % mpad);
        double v = 0.0;
        if (c <= r) v = (r < m) ? D[(size_t)r * ldd + c] : (r == c ? 1.0 : 0.0);
        A[(size_t)r * ld + c] = v;
    }
}
__global__ void k_unpark(const double *__restrict__ A, int ld, int m, double *__restrict__ L, int ldl) {
    const long long total = (long long)m * m;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e / m), j = (int)(e % m), ib = i / TB, kb = j / TB;
        double v = 0.0;
        if (kb < ib) v = A[(size_t)(kb * TB + i % TB) * ld + ib * TB + j % TB];
        else if (j <= i) v = A[(size_t)i * ld + j];
        L[(size_t)i * ldl + j] = v;
    }
}

void launch_pack(const double *D, int ldd, const CholWork &c, hipStream_t q) {
    const long long total = (long long)c.mpad * c.mpad;
    const int grid = (int)std::min<long long>((total + 255) / 256, 4096);
    k_pack<<<grid, 256, 0, q>>>(D, ldd, c.m, c.A, c.ld, c.mpad);
}
void launch_unpark(const CholWork &c, double *L, int ldl, hipStream_t q) {
    const long long total = (long long)c.m * c.m;
    const int grid = (int)std::max<long long>(1, std::min<long long>((total + 255) / 256, 4096));
    k_unpark<<<grid, 256, 0, q>>>(c.A, c.ld, c.m, L, ldl);
}

bool refine_active(const CholWork &c) { return c.mode != 0 && c.refine && c.m > 0 && c.xs; }

void launch_refine(const Stat &s, const CholWork &c, const double *p, const double *rad_all, hipStream_t q) {
    const int nt = c.mpad / TB, nhi = c.m / 3, ng = nblk(nt, 2);
    double *pw = c.rv, *u = c.rv + c.ld;
    k_refine_init<<<nblk(c.mpad, 256), 256, 0, q>>>(p, c.m, c.mpad, pw);
    for (int g = ng - 1; g >= 0; g--) k_bsolve<<<1 + 2 * g, 256, 0, q>>>(c.A, c.ld, c.Linv, nt, g, pw, u);
    k_dmul<<<nblk(nhi, 4), 256, 0, q>>>(c.xs, rad_all, nhi, s.a1, s.a2, u, c.pr);
}

void launch_noise(const Stat &s, int n_hi, int n_D3, const CholWork &c, const double *p, const double *rad_all,
                  double *rd, hipStream_t q, const double *du) {
    k_noise<<<nblk(n_D3, 4), 256, 0, q>>>(c.A, c.ld, c.m, n_D3, p, du, rad_all, s.s2dt, s.a1, rd);
    (void)n_hi;
}

}  // namespace bd
