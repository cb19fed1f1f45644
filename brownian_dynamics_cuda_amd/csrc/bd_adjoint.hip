// bd_adjoint.hip -- the core-core charge term of the force pass and its
// adjoint, for the differentiable trajectory (torch.autograd.Function in
// brownian_dynamics_cuda_amd/diff.py).
//
// Term (func.cpp:1489-1566): for two cores i, j within the 25 nm bead
// rule and every pair of their point charges k (on i) and l (on j), with
// z = Q_k - Q_l, d = |z|,
//   E  = ke q_k q_l exp(-kappa d) / d  +  [d <= 8] kex ((s/d)^12 - (s/d)^6)
//   f_k = g(d) z/d = -f_l,   g = -dE/dd
//       = ke q_k q_l exp(-kappa d) (kappa d + 1) / d^2
//         + [d <= 8] kex (12 (s/d)^12 - 6 (s/d)^6) / d.
// Here the output is the force on every charge site (the caller turns it
// into core forces and body-frame torques, whose derivatives torch takes)
// and each core's half of its pairs' energy.
//
// Adjoint: with lambda_k = dLoss/df_k and eps_i = dLoss/de_i, a pair adds
//   phi = (lambda_k - lambda_l) . g z/d + (eps_i + eps_j)/2 E
// to the loss; with u = lambda_k - lambda_l and uz = u . z/d,
//   dphi/dQ_k = g'(d) uz z/d + g(d) (u - uz z/d) / d - eps g z/d = -dphi/dQ_l,
//   g'_DH = -ke q_k q_l exp(-kappa d) (kappa^2 d^2 + 2 kappa d + 2) / d^3,
//   g'_EV = kex (-156 (s/d)^12 + 42 (s/d)^6) / d^2,
// and dphi/dq_k = q_l (uz gh + eps eh) with gh = g_DH / (q_k q_l),
// eh = E_DH / (q_k q_l) (the excluded volume does not depend on charge).
//
// Layout: one workgroup per core i, its charge sites dealt over the
// threads; the partner cores of i (CSR list, both directions) staged one
// at a time in LDS.  Every pair is evaluated from both sides, so each
// output has one owner and no atomics: the results are deterministic.
// Templated on the arithmetic (float for the FP32 force passes of the
// differentiable trajectory, double for its FP64 checks).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/bd_hip.h"
#include "bd_kernels.h"

namespace bd {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxSites = 640;  // charges per core held in LDS

struct CCParams {
    double ke, kappa, kex, sigma, cut_ev;
};

template <typename T>
struct PairTerms {
    T g, gp, e, gh, eh;  // g, g', energy, g_DH/(q_k q_l), E_DH/(q_k q_l)
};

template <typename T>
__device__ __forceinline__ PairTerms<T> pair_terms(T d, T qkql, const CCParams &p) {
    const T id = T(1) / d;
    const T kap = (T)p.kappa, ke = (T)p.ke;
    const T ex = exp(-kap * d);
    PairTerms<T> r;
    r.eh = ke * ex * id;
    r.gh = r.eh * (kap * d + T(1)) * id;
    r.e = qkql * r.eh;
    r.g = qkql * r.gh;
    r.gp = -qkql * ke * ex * (kap * kap * d * d + T(2) * kap * d + T(2)) * id * id * id;
    if (d <= (T)p.cut_ev) {
        const T sd = (T)p.sigma * id, s2 = sd * sd, s6 = s2 * s2 * s2, s12 = s6 * s6, kex = (T)p.kex;
        r.e += kex * (s12 - s6);
        r.g += kex * (T(12) * s12 - T(6) * s6) * id;
        r.gp += kex * (T(-156) * s12 + T(42) * s6) * id * id;
    }
    return r;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *buf) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) buf[w] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0)
        for (int k = 0; k < kThreads / 64; k++) s += buf[k];
    return s;
}

// f[i][k][3] (written) and e[i] = half the energy of core i's pairs
template <typename T>
__global__ void __launch_bounds__(kThreads) k_cc_fwd(int Nq, const T *__restrict__ Q, const T *__restrict__ q,
                                                     const int *__restrict__ off, const int *__restrict__ nbr,
                                                     CCParams p, T *__restrict__ f, T *__restrict__ e) {
    __shared__ T sQ[3 * kMaxSites], sq[kMaxSites], red[kThreads / 64];
    const int i = blockIdx.x;
    constexpr int KP = (kMaxSites + kThreads - 1) / kThreads;
    T fx[KP], fy[KP], fz[KP], qk[KP], px[KP], py[KP], pz[KP];
    T en = T(0);
    for (int t = 0; t < KP; t++) {
        const int k = threadIdx.x + t * kThreads;
        fx[t] = fy[t] = fz[t] = T(0);
        if (k < Nq) {
            const T *src = Q + ((size_t)i * Nq + k) * 3;
            px[t] = src[0]; py[t] = src[1]; pz[t] = src[2];
            qk[t] = q[k];
        } else {
            px[t] = py[t] = pz[t] = qk[t] = T(0);
        }
    }
    for (int s = off[i]; s < off[i + 1]; s++) {
        const int j = nbr[s];
        __syncthreads();
        for (int x = threadIdx.x; x < 3 * Nq; x += kThreads) sQ[x] = Q[(size_t)j * Nq * 3 + x];
        for (int x = threadIdx.x; x < Nq; x += kThreads) sq[x] = q[x];
        __syncthreads();
        for (int t = 0; t < KP; t++) {
            const int k = threadIdx.x + t * kThreads;
            if (k >= Nq) continue;
            for (int l = 0; l < Nq; l++) {
                const T zx = px[t] - sQ[3 * l], zy = py[t] - sQ[3 * l + 1], zz = pz[t] - sQ[3 * l + 2];
                const T d = sqrt(zx * zx + zy * zy + zz * zz);
                const PairTerms<T> r = pair_terms<T>(d, qk[t] * sq[l], p);
                const T gi = r.g / d;
                fx[t] += gi * zx;
                fy[t] += gi * zy;
                fz[t] += gi * zz;
                en += T(0.5) * r.e;
            }
        }
    }
    for (int t = 0; t < KP; t++) {
        const int k = threadIdx.x + t * kThreads;
        if (k < Nq) {
            T *dst = f + ((size_t)i * Nq + k) * 3;
            dst[0] = fx[t]; dst[1] = fy[t]; dst[2] = fz[t];
        }
    }
    const T tot = block_sum(en, red);
    if (threadIdx.x == 0) e[i] = tot;
}

// gQ[i][k][3] and gq[i][k] (per-site charge gradient; the template
// gradient is its sum over cores), both written
template <typename T>
__global__ void __launch_bounds__(kThreads) k_cc_bwd(int Nq, const T *__restrict__ Q, const T *__restrict__ q,
                                                     const int *__restrict__ off, const int *__restrict__ nbr,
                                                     CCParams p, const T *__restrict__ gf, const T *__restrict__ ge,
                                                     T *__restrict__ gQ, T *__restrict__ gqs) {
    __shared__ T sQ[3 * kMaxSites], sL[3 * kMaxSites], sq[kMaxSites];
    const int i = blockIdx.x;
    constexpr int KP = (kMaxSites + kThreads - 1) / kThreads;
    T gx[KP], gy[KP], gz[KP], gqk[KP], qk[KP], px[KP], py[KP], pz[KP], lx[KP], ly[KP], lz[KP];
    const T ei = ge[i];
    for (int t = 0; t < KP; t++) {
        const int k = threadIdx.x + t * kThreads;
        gx[t] = gy[t] = gz[t] = gqk[t] = T(0);
        if (k < Nq) {
            const size_t o = ((size_t)i * Nq + k) * 3;
            px[t] = Q[o]; py[t] = Q[o + 1]; pz[t] = Q[o + 2];
            lx[t] = gf[o]; ly[t] = gf[o + 1]; lz[t] = gf[o + 2];
            qk[t] = q[k];
        } else {
            px[t] = py[t] = pz[t] = lx[t] = ly[t] = lz[t] = qk[t] = T(0);
        }
    }
    for (int s = off[i]; s < off[i + 1]; s++) {
        const int j = nbr[s];
        const T eps = T(0.5) * (ei + ge[j]);
        __syncthreads();
        for (int x = threadIdx.x; x < 3 * Nq; x += kThreads) {
            sQ[x] = Q[(size_t)j * Nq * 3 + x];
            sL[x] = gf[(size_t)j * Nq * 3 + x];
        }
        for (int x = threadIdx.x; x < Nq; x += kThreads) sq[x] = q[x];
        __syncthreads();
        for (int t = 0; t < KP; t++) {
            const int k = threadIdx.x + t * kThreads;
            if (k >= Nq) continue;
            for (int l = 0; l < Nq; l++) {
                const T zx = px[t] - sQ[3 * l], zy = py[t] - sQ[3 * l + 1], zz = pz[t] - sQ[3 * l + 2];
                const T d = sqrt(zx * zx + zy * zy + zz * zz), id = T(1) / d;
                const T ux = lx[t] - sL[3 * l], uy = ly[t] - sL[3 * l + 1], uz_ = lz[t] - sL[3 * l + 2];
                const T hx = zx * id, hy = zy * id, hz = zz * id;
                const T uz = ux * hx + uy * hy + uz_ * hz;
                const PairTerms<T> r = pair_terms<T>(d, qk[t] * sq[l], p);
                const T a = r.gp * uz - eps * r.g - r.g * uz * id, b = r.g * id;
                gx[t] += a * hx + b * ux;
                gy[t] += a * hy + b * uy;
                gz[t] += a * hz + b * uz_;
                gqk[t] += sq[l] * (uz * r.gh + eps * r.eh);
            }
        }
    }
    for (int t = 0; t < KP; t++) {
        const int k = threadIdx.x + t * kThreads;
        if (k < Nq) {
            const size_t o = ((size_t)i * Nq + k) * 3;
            gQ[o] = gx[t]; gQ[o + 1] = gy[t]; gQ[o + 2] = gz[t];
            gqs[(size_t)i * Nq + k] = gqk[t];
        }
    }
}

int cc_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_last_error(buf);
    return code;
}

int cc_check(int n_c, int Nq, const void *Q, const void *q, const int *off, const int *nbr) {
    if (n_c < 1 || Nq < 1 || Nq > kMaxSites)
        return cc_fail(BD_E_ARG, "need n_c >= 1 and 1 <= Nq <= %d (n_c=%d, Nq=%d)", kMaxSites, n_c, Nq);
    if (!Q || !q || !off || !nbr) return cc_fail(BD_E_ARG, "null argument");
    return BD_OK;
}

}  // namespace
}  // namespace bd

using namespace bd;

extern "C" {

int bd_cc_forward(int fp64, int n_c, int Nq, const void *Q, const void *q, const int *nbr_off, const int *nbr,
                  double ke, double kappa, double kex, double sigma, void *f, void *e, void *stream) {
    int r = cc_check(n_c, Nq, Q, q, nbr_off, nbr);
    if (r) return r;
    if (!f || !e) return cc_fail(BD_E_ARG, "null output");
    const CCParams p{ke, kappa, kex, sigma, 8.0};
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (fp64)
        k_cc_fwd<double><<<n_c, kThreads, 0, s>>>(Nq, (const double *)Q, (const double *)q, nbr_off, nbr, p,
                                                  (double *)f, (double *)e);
    else
        k_cc_fwd<float><<<n_c, kThreads, 0, s>>>(Nq, (const float *)Q, (const float *)q, nbr_off, nbr, p,
                                                 (float *)f, (float *)e);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return cc_fail(BD_E_HIP, "k_cc_fwd: %s", hipGetErrorString(err));
    return BD_OK;
}

int bd_cc_backward(int fp64, int n_c, int Nq, const void *Q, const void *q, const int *nbr_off, const int *nbr,
                   double ke, double kappa, double kex, double sigma, const void *gf, const void *ge, void *gQ,
                   void *gq_sites, void *stream) {
    int r = cc_check(n_c, Nq, Q, q, nbr_off, nbr);
    if (r) return r;
    if (!gf || !ge || !gQ || !gq_sites) return cc_fail(BD_E_ARG, "null argument");
    const CCParams p{ke, kappa, kex, sigma, 8.0};
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (fp64)
        k_cc_bwd<double><<<n_c, kThreads, 0, s>>>(Nq, (const double *)Q, (const double *)q, nbr_off, nbr, p,
                                                  (const double *)gf, (const double *)ge, (double *)gQ,
                                                  (double *)gq_sites);
    else
        k_cc_bwd<float><<<n_c, kThreads, 0, s>>>(Nq, (const float *)Q, (const float *)q, nbr_off, nbr, p,
                                                 (const float *)gf, (const float *)ge, (float *)gQ,
                                                 (float *)gq_sites);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return cc_fail(BD_E_HIP, "k_cc_bwd: %s", hipGetErrorString(err));
    return BD_OK;
}

}  // extern "C"
