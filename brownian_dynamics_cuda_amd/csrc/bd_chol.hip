// bd_chol.hip -- RPY diffusion tensor, blocked Cholesky and the noise
// displacement rd = s2dt * L * p on gfx950.
//
// Replaces Diffusion_Tensor_CUDA + cusolverDnDpotrf + Cholesky_Decomposition_mod
// + rd_cal (func_cuda.cu:500-647, 3504-3552).  The factor follows the CPU
// Cholesky of func.cpp:806-836, including |f| for a negative pivot.
//
// Matrix: row-major, lower triangle, padded to a multiple of 64 with an
// identity block (so the padded factor is [[L,0],[0,I]] and no kernel needs
// a bounds test).  Right-looking blocked algorithm with 64x64 tiles:
//   panel kb:  trsm   L_ib,kb = A_ib,kb * inv(L_kb,kb)^T   (FP64 MFMA)
//              syrk   A_ib,jb -= L_ib,kb * L_jb,kb^T       (FP64 or FP32 MFMA)
//              and the workgroup owning tile (kb+1,kb+1) factors it in LDS
//              right after its update (potrf + triangular inverse, FP64).
#include "bd_dev.h"
#include "bd_kernels.h"

namespace bd {

constexpr int TB = 64;        // tile edge
constexpr int LDP = TB + 1;   // LDS row pitch (doubles)
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

static inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// RPY tensor assembly (func.cpp:737-804): one element per thread, lower
// triangle only, identity padding.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double3 hi_pos(const Stat &s, const Geo &g, int b) {
    if (b < s.n) return ld3(g.r + 3 * b);
    b -= s.n;
    if (b < s.n_tail) return ld3(g.r_t + 3 * b);
    return ld3(g.r_lh + 3 * (b - s.n_tail));
}

__device__ __forceinline__ double rpy_entry(double3 ri, double3 rj, double radi, double radj, int k, int l,
                                            double a1, double a2) {
    double sij[3] = {ri.x - rj.x, ri.y - rj.y, ri.z - rj.z};
    double ssq = sij[0] * sij[0] + sij[1] * sij[1] + sij[2] * sij[2];
    double s = sqrt(ssq);
    if (s >= radi + radj) {
        double f = (radi * radi + radj * radj) / ssq;
        double f1 = 1.0 + f / 3, f2 = 1.0 - f;
        return (k == l) ? (a2 / s) * (f1 + f2 * sij[k] * sij[k] / ssq) : (a2 / s) * (f2 * sij[k] * sij[l] / ssq);
    }
    double se = pow((radi * radi * radi + radj * radj * radj) / 2.0, 1.0 / 3);
    double sq = sqrt(ssq);
    return (k == l) ? (a1 / se) * (1.0 - 9.0 * sq / (32 * se) + 3.0 * sij[k] * sij[k] / (32.0 * se * sq))
                    : (a1 / se) * (3.0 * sij[k] * sij[l] / (32.0 * se * sq));
}

__device__ __forceinline__ void tri_index(int idx, int &i, int &j) {
    // idx -> (i, j), 0 <= j <= i, row-major over the lower triangle
    int r = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= idx) r++;
    while (r * (r + 1) / 2 > idx) r--;
    i = r;
    j = idx - r * (r + 1) / 2;
}

__global__ void __launch_bounds__(256) k_rpy(Stat s, Geo g, const double *rad_all, double *A, int m, int ld,
                                             double a1, double a2) {
    int ti, tj;
    tri_index(blockIdx.x, ti, tj);
    for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
        const int row = ti * TB + e / TB, col = tj * TB + e % TB;
        double v;
        if (col > row) v = 0.0;
        else if (row >= m || col >= m) v = (row == col) ? 1.0 : 0.0;
        else {
            const int i = row / 3, j = col / 3, k = row % 3, l = col % 3;
            if (i == j) v = (k == l) ? a1 / rad_all[i] : 0.0;
            else v = rpy_entry(hi_pos(s, g, i), hi_pos(s, g, j), rad_all[i], rad_all[j], k, l, a1, a2);
        }
        A[(size_t)row * ld + col] = v;
    }
}

// ---------------------------------------------------------------------------
// tile helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_tile(const double *A, int ld, int r0, int c0, double *S) {
    for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
        int r = e / TB, c = e % TB;
        S[r * LDP + c] = A[(size_t)(r0 + r) * ld + c0 + c];
    }
}
__device__ __forceinline__ void store_tile(double *A, int ld, int r0, int c0, const double *S, bool lower_only) {
    for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
        int r = e / TB, c = e % TB;
        if (lower_only && c > r) continue;
        A[(size_t)(r0 + r) * ld + c0 + c] = S[r * LDP + c];
    }
}

// acc += As * Bs^T for a 64x64 tile; 4 waves, each a 32x32 quarter as 2x2
// tiles of v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][l/16] and
// B^T[l/16][l%16]; result register v of lane l is C[l/16 + 4v][l%16]
// (measured on gfx950; the FP32 16x16x4 form places it at C[4(l/16)+v]).
__device__ __forceinline__ void mma_f64(const double *As, const double *Bs, dbl4 acc[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll 4
    for (int kk = 0; kk < TB; kk += 4) {
        double a0 = As[(wr + i16) * LDP + kk + k4];
        double a1 = As[(wr + 16 + i16) * LDP + kk + k4];
        double b0 = Bs[(wc + i16) * LDP + kk + k4];
        double b1 = Bs[(wc + 16 + i16) * LDP + kk + k4];
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
#pragma unroll 4
    for (int kk = 0; kk < TB; kk += 4) {
        float a0 = (float)As[(wr + i16) * LDP + kk + k4];
        float a1 = (float)As[(wr + 16 + i16) * LDP + kk + k4];
        float b0 = (float)Bs[(wc + i16) * LDP + kk + k4];
        float b1 = (float)Bs[(wc + 16 + i16) * LDP + kk + k4];
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
}

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

// 1/sqrt(f) to full FP64 accuracy: v_rsq_f64 refined by two Newton steps.
__device__ __forceinline__ double rsqrt_f64(double f) {
    double y = __builtin_amdgcn_rsq(f);
    y = y * (1.5 - 0.5 * f * y * y);
    y = y * (1.5 - 0.5 * f * y * y);
    return y;
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Cholesky of the 64x64 LDS tile L (lower, in place), the CPU algorithm of
// func.cpp:806-836 reordered into 16-column sub-panels.  A pivot f < 0 is
// replaced by -f and counted (func.cpp:827-833); the pivot's inverse is
// kept in invd[] for the triangular solves.  Each sub-panel is factored by
// wavefront 0 alone -- lane r holds row r, the 16 columns in registers,
// pivots and column entries are broadcast with readlane, so a column step
// has no barrier -- then all waves apply its rank-16 update to the rest of
// the tile.  Requires blockDim.x == 256.
__device__ void potrf_tile_lds(double *L, double *invd, int *neg) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int sp = 0; sp < 4; sp++) {
        const int c0 = 16 * sp;
        if (w == 0) {
            double l[16];
            const bool act = lane >= c0;
#pragma unroll
            for (int k = 0; k < 16; k++) l[k] = act ? L[lane * LDP + c0 + k] : 0.0;
#pragma unroll
            for (int jj = 0; jj < 16; jj++) {
                const int j = c0 + jj;
                double f = readlane_f64(l[jj], j);
                if (f < 0) {
                    f = -f;
                    if (lane == 0) atomicAdd(neg, 1);
                }
                const double inv = rsqrt_f64(f);
                if (lane == j) { l[jj] = f * inv; invd[j] = inv; }
                else if (lane > j) l[jj] = l[jj] * inv;
#pragma unroll
                for (int kk = jj + 1; kk < 16; kk++) {
                    const double ckj = readlane_f64(l[jj], c0 + kk);  // L[c0+kk][j]
                    if (lane >= c0 + kk) l[kk] -= l[jj] * ckj;
                }
            }
            if (act)
#pragma unroll
                for (int k = 0; k < 16; k++) L[lane * LDP + c0 + k] = l[k];
        }
        __syncthreads();
        // rank-16 update of the tile below and right of the sub-panel
        const int base = c0 + 16;
        const int nr = TB - base;
        for (int e = tid; e < nr * nr; e += blockDim.x) {
            const int r = base + e / nr, k = base + e % nr;
            if (k > r) continue;
            double acc = L[r * LDP + k];
#pragma unroll
            for (int c = 0; c < 16; c++) acc -= L[r * LDP + c0 + c] * L[k * LDP + c0 + c];
            L[r * LDP + k] = acc;
        }
        __syncthreads();
    }
}

// first diagonal tile
__global__ void __launch_bounds__(256) k_potrf0(double *A, int ld, double *invd, int *neg) {
    __shared__ double L[TB * LDP];
    load_tile(A, ld, 0, 0, L);
    __syncthreads();
    potrf_tile_lds(L, invd, neg);
    store_tile(A, ld, 0, 0, L, true);
}

// Panel below the diagonal: X * L_kk^T = A_ib,kb for every row block
// ib > kb, by forward substitution.  Lane = row (64 rows per wavefront,
// four wavefronts per workgroup); the row lives in registers; L_kk and
// the pivot inverses are broadcast from LDS.
__global__ void __launch_bounds__(256) k_trsm(double *A, int ld, int kb, const double *invd, int nrow_blocks) {
    __shared__ double Lk[TB * TB], iv[TB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int e = tid; e < TB * TB; e += blockDim.x) {
        const int r = e / TB, c = e % TB;
        Lk[e] = (c <= r) ? A[(size_t)(kb * TB + r) * ld + kb * TB + c] : 0.0;
    }
    if (tid < TB) iv[tid] = invd[kb * TB + tid];
    __syncthreads();
    const int ib = kb + 1 + blockIdx.x * 4 + w;
    if (blockIdx.x * 4 + w >= nrow_blocks) return;
    double *row = A + (size_t)(ib * TB + lane) * ld + kb * TB;
    double b[TB];
#pragma unroll
    for (int k = 0; k < TB; k++) b[k] = row[k];
#pragma unroll
    for (int j = 0; j < TB; j++) {
        b[j] = b[j] * iv[j];
#pragma unroll
        for (int k = j + 1; k < TB; k++) b[k] -= Lk[k * TB + j] * b[j];
    }
#pragma unroll
    for (int k = 0; k < TB; k++) row[k] = b[k];
}

// trailing update of tile (ib, jb), kb < jb <= ib; the tile (kb+1, kb+1) is
// then factored in place.  FP32 MFMA variant when MODE != 0.
template <int MODE>
__global__ void __launch_bounds__(256) k_syrk(double *A, int ld, int kb, double *invd, int *neg) {
    __shared__ double As[TB * LDP], Bs[TB * LDP];
    int ii, jj;
    tri_index(blockIdx.x, ii, jj);
    const int ib = kb + 1 + ii, jb = kb + 1 + jj;
    load_tile(A, ld, ib * TB, kb * TB, As);
    load_tile(A, ld, jb * TB, kb * TB, Bs);
    __syncthreads();
    if (MODE == 0) {
        dbl4 acc[2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
        mma_f64(As, Bs, acc);
        __syncthreads();
        load_tile(A, ld, ib * TB, jb * TB, As);
        __syncthreads();
        acc_to_lds(As, acc, -1.0, false);
    } else {
        flt4 acc[2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) acc[a][b] = flt4{0.f, 0.f, 0.f, 0.f};
        mma_f32(As, Bs, acc);
        __syncthreads();
        load_tile(A, ld, ib * TB, jb * TB, As);
        __syncthreads();
        acc_to_lds(As, acc, -1.0, false);
    }
    __syncthreads();
    if (ii == 0 && jj == 0) {
        potrf_tile_lds(As, invd + ib * TB, neg);
        store_tile(A, ld, ib * TB, jb * TB, As, true);
    } else {
        store_tile(A, ld, ib * TB, jb * TB, As, ib == jb);
    }
}

// rd = s2dt * L * p (rd_cal, func_cuda.cu:634-647): one wavefront per row,
// coalesced over the row; rows outside the hydrodynamic set use their
// self-mobility: s2dt * sqrt(a1 / rad) * p.
__global__ void __launch_bounds__(256) k_noise(const double *A, int ld, int m, int nD3, const double *p,
                                               const double *rad_all, double s2dt, double a1, double *rd) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nD3) return;
    if (row >= m) {
        if (lane == 0) rd[row] = s2dt * sqrt(a1 / rad_all[row / 3]) * p[row];
        return;
    }
    const double *Lr = A + (size_t)row * ld;
    double v = 0.0;
    for (int k = lane; k <= row; k += 64) v += Lr[k] * p[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) rd[row] = s2dt * v;
}

// ---------------------------------------------------------------------------
void launch_rpy(const Stat &s, int n_hi, const Geo &g, const double *rad_all, const CholWork &c, hipStream_t q) {
    const int nt = c.mpad / TB;
    k_rpy<<<nt * (nt + 1) / 2, 256, 0, q>>>(s, g, rad_all, c.A, c.m, c.ld, s.a1, s.a2);
    (void)n_hi;
}

void launch_cholesky(const CholWork &c, hipStream_t q, MarkFn mark, void *mark_arg) {
    const int nt = c.mpad / TB;
    auto M_ = [&](int id, int ph) { if (mark) mark(mark_arg, id, ph); };
    launch_fill_int(c.neg, 1, 0, q);
    M_(K_POTRF, 0);
    k_potrf0<<<1, 256, 0, q>>>(c.A, c.ld, c.Linv, c.neg);
    M_(K_POTRF, 1);
    for (int kb = 0; kb + 1 < nt; kb++) {
        const int M = nt - kb - 1;
        M_(K_TRSM, 0);
        k_trsm<<<(M + 3) / 4, 256, 0, q>>>(c.A, c.ld, kb, c.Linv, M);
        M_(K_TRSM, 1);
        M_(K_SYRK, 0);
        if (c.mode == 0)
            k_syrk<0><<<M * (M + 1) / 2, 256, 0, q>>>(c.A, c.ld, kb, c.Linv, c.neg);
        else
            k_syrk<1><<<M * (M + 1) / 2, 256, 0, q>>>(c.A, c.ld, kb, c.Linv, c.neg);
        M_(K_SYRK, 1);
    }
}

void launch_noise(const Stat &s, int n_hi, int n_D3, const CholWork &c, const double *p, const double *rad_all,
                  double *rd, hipStream_t q) {
    k_noise<<<nblk(n_D3, 4), 256, 0, q>>>(c.A, c.ld, c.m, n_D3, p, rad_all, s.s2dt, s.a1, rd);
    (void)n_hi;
}

}  // namespace bd
