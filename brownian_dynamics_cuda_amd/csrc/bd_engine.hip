// bd_engine.hip -- host side of libbdhip.so: the context that owns every
// device buffer of one fiber, the step orchestration in the order of
// main_cuda (func_cuda.cu:3434-3680) and the extern "C" API of bd_hip.h.
//
// Differences from the reference's host code, all by design:
//   * every buffer is allocated once at bd_create (the reference frees and
//     re-allocates LH arrays, D, its factor and the noise buffers each step,
//     func_cuda.cu:3436-3548) -- a step allocates nothing;
//   * two HIP streams: D, its Cholesky factor and rd are formed on a side
//     stream while force pass 1 runs on the main one (both only read the
//     state at x0); the whole step can be captured once and replayed as a
//     hipGraph (no hipMemsetAsync anywhere in it, counters are cleared by
//     fill kernels);
//   * D is assembled and factored in place (lower triangle); the drift
//     D*F is formed without a stored D (closed-form RPY blocks).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bd_hip.h"
#include "bd_dev.h"
#include "bd_kernels.h"

using namespace bd;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(BD_E_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                       \
    } while (0)

const char *kKernelNames[K_COUNT] = {
    "sites", "mech", "linker", "tail", "lh", "core", "cc_list", "cc_work", "rpy",
    "chol_step", "noise", "frames", "betas", "translate",
    "commit", "energy_reduce", "cc_gather", "chol_trail", "chol_diag", "refine"};

}  // namespace

namespace bd {
void set_last_error(const char *msg) { g_err = msg; }
}  // namespace bd

struct bd_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of the step: D, its factor and rd run there, beside
    // force pass 1 (both read only the state at x0); joined before the
    // predictor translation
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_gate = nullptr;
    // force pass 1 waits for the factor to pass this block column (-1: the
    // default fraction of the block columns; BD_GATE_KB overrides it)
    int gate_kb = -1;
    // branches of the two force passes of a step (streams shared, events per pass)
    ForceStreams fstr[2]{};
    std::vector<void *> allocs;
    // sizes
    int n = 0, n_c = 0, n_tail = 0, n_LH = 0, lh_cap = 0, n_D = 0, n_D3 = 0, n_hi = 0, hi_all = 0;
    int nlh = 0, groups_per_core = 10;
    Stat st{};
    Geo cur{}, pred{};
    Frc f0{}, f1{};
    Sites sites{};
    ForceScratch fw{};
    CholWork chol{};
    double *rd = nullptr, *p = nullptr, *rr = nullptr, *energy = nullptr, *Epart = nullptr;
    double *rad_all = nullptr;
    int ne = 0;
    bool state_set = false;
    bool use_graph = true;
    // captured steps: [0] with D and its factor, [1] reusing the factor
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // frequency_RP (main.cpp:45): D and L are rebuilt on the first step after
    // an invalidation (new state, new hydrodynamic size, precision change)
    // and then every rp_freq-th step; the other steps reuse L for rd
    int rp_freq = 1;
    long long rp_count = 0;
    bool factor_fresh = false;
    // chol.xs holds the positions of the last assembly of D (it is written
    // only in the modes whose noise takes the refinement)
    bool xs_valid = false;
    // noise and energies of bd_run_steps, grown on demand
    double *batch = nullptr;
    size_t batch_cap = 0;
    // timing
    bool timing = false;
    std::vector<std::pair<int, hipEvent_t>> pending;  // (kernel id, event) pairs, start/stop
    double total_ms[K_COUNT] = {0};
    long long launches[K_COUNT] = {0};
    std::vector<int> nc_lh_flag, core_bead_host;

    template <typename T>
    int alloc(T **p, size_t count) {
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(BD_E_HIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        allocs.push_back(q);
        *p = static_cast<T *>(q);
        return BD_OK;
    }
    template <typename T>
    int upload(T **p, const T *src, size_t count) {
        int r = alloc(p, count);
        if (r) return r;
        if (count && src) {
            hipError_t e = hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(BD_E_HIP, "hipMemcpy H2D: %s", hipGetErrorString(e));
        }
        return BD_OK;
    }
};

namespace {

__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, long long nw) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nw;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void mark_cb(void *arg, int id, int phase) {
    bd_context *c = static_cast<bd_context *>(arg);
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return;
    (void)hipEventRecord(ev, c->stream);
    c->pending.push_back({id * 2 + phase, ev});
}

void collect_timing(bd_context *c) {
    (void)hipStreamSynchronize(c->stream);
    hipEvent_t open[K_COUNT] = {nullptr};
    for (auto &pe : c->pending) {
        int id = pe.first / 2, ph = pe.first % 2;
        if (ph == 0) {
            open[id] = pe.second;
        } else if (open[id]) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, open[id], pe.second) == hipSuccess) {
                c->total_ms[id] += ms;
                c->launches[id] += 1;
            }
        }
    }
    for (auto &pe : c->pending) (void)hipEventDestroy(pe.second);
    c->pending.clear();
}

struct Marker {
    bd_context *c;
    MarkFn fn() const { return c->timing ? mark_cb : nullptr; }
    void operator()(int id, int ph) const {
        if (c->timing) mark_cb(c, id, ph);
    }
};

// The force pass at configuration g into f (energy only when Epart given).
// Pass 2 (pass == 1) runs its branches on their own streams unless the
// timing markers are on; pass 1 stays on one stream, because it already
// shares the GPU with the factorisation on the side stream.
void enqueue_forces(bd_context *c, const Geo &g, const Frc &f, double *Epart, int terms, int pass,
                    bool sites_done = false) {
    launch_forces(c->st, g, c->sites, f, c->fw, Epart, terms, c->stream, c->timing ? mark_cb : nullptr, c,
                  (c->timing || pass == 0) ? nullptr : &c->fstr[pass], sites_done);
}

void enqueue_factor(bd_context *c, const Geo &g, hipStream_t q, hipEvent_t gate = nullptr, int gate_kb = -1) {
    Marker M{c};
    M(K_RPY, 0);
    launch_rpy(c->st, c->n_hi, g, c->rad_all, c->chol, q, true);
    c->xs_valid = refine_active(c->chol);
    M(K_RPY, 1);
    launch_cholesky(c->chol, q, M.fn(), c, gate, gate_kb);
}

void drop_graphs(bd_context *c) {
    for (auto &g : c->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// One predictor-corrector step (func_cuda.cu:3494-3660) reading c->rr, c->p.
// refactor=false keeps the L of an earlier step (the frequency_RP gate of
// func_cuda.cu:3504-3535) and only recomputes rd = s2dt L p.
void enqueue_step(bd_context *c, bool refactor) {
    Marker M{c};
    hipStream_t q = c->stream;
    const Stat &s = c->st;
    // D, its factor and the noise displacement (3504-3552) on the side
    // stream (serial when the per-kernel timing markers are on)
    hipStream_t qf = q;
    if (!c->timing) {
        (void)hipEventRecord(c->ev_fork, q);
        (void)hipStreamWaitEvent(c->side, c->ev_fork, 0);
        qf = c->side;
    }
    // force pass 1 starts once the factor is past block column gate_kb:
    // the early block columns fill the GPU, and force kernels beside them
    // hold the CUs that the diagonal-tile workgroups wait for
    const int nt = c->chol.mpad / 64;
    const int gate_kb = c->gate_kb >= 0 ? c->gate_kb : (nt - 1) * 55 / 100;
    const bool gated = refactor && qf != q;
    if (refactor) enqueue_factor(c, c->cur, qf, gated ? c->ev_gate : nullptr, gate_kb);
    const double *du = nullptr;
    if (refine_active(c->chol)) {
        M(K_REFINE, 0); launch_refine(s, c->chol, c->p, c->rad_all, qf); M(K_REFINE, 1);
        du = c->chol.pr;
    }
    M(K_NOISE, 0); launch_noise(s, c->n_hi, c->n_D3, c->chol, c->p, c->rad_all, c->rd, qf, du); M(K_NOISE, 1);
    if (gated) (void)hipStreamWaitEvent(q, c->ev_gate, 0);
    // LH bending angles at the current state (3497)
    M(K_BETAS, 0); launch_betas(s, c->cur, 0, 1, q); M(K_BETAS, 1);
    // force pass 1 at x0 and its energy (3556-3615)
    enqueue_forces(c, c->cur, c->f0, c->Epart, BD_TERM_ALL, 0);
    M(K_EREDUCE, 0); launch_energy_reduce(c->Epart, c->ne, c->energy, q); M(K_EREDUCE, 1);
    if (qf != q) {
        (void)hipEventRecord(c->ev_join, qf);
        (void)hipStreamWaitEvent(q, c->ev_join, 0);
    }
    // predictor (3617-3619): translation and rotation in one launch
    M(K_TRANSLATE, 0);
    launch_translate(s, c->hi_all, c->cur, c->pred, c->f0, Frc{}, c->rd, c->rad_all, c->rr, q);
    M(K_TRANSLATE, 1);
    // rebuild at the predicted state (3621-3625), with force pass 2's sites
    // in the same launch (outside the per-kernel timing mode)
    M(K_FRAMES, 0);
    const bool sites_done = !c->timing && launch_frames_sites(s, c->pred, c->sites, c->fw.cc_count, q);
    if (!sites_done && !launch_frames_betas(s, c->pred, 1, 1, q)) {
        launch_betas(s, c->pred, 1, 1, q);
        launch_frames(s, c->pred, q);
    }
    M(K_FRAMES, 1);
    // force pass 2 (3628-3646)
    enqueue_forces(c, c->pred, c->f1, nullptr, BD_TERM_ALL, 1, sites_done);
    // corrector from x0 with the averaged forces, same rd and rr (3648-3652)
    // (the Heun means of forces and torques formed where they are read)
    M(K_TRANSLATE, 0);
    launch_translate(s, c->hi_all, c->cur, c->pred, c->f0, c->f1, c->rd, c->rad_all, c->rr, q);
    M(K_TRANSLATE, 1);
    // commit and rebuild (3654-3660)
    M(K_COMMIT, 0); launch_commit(s, c->pred, c->cur, q); M(K_COMMIT, 1);
    M(K_FRAMES, 0);
    if (!launch_frames_betas(s, c->cur, 1, 1, q)) {
        launch_frames(s, c->cur, q);
        launch_betas(s, c->cur, 1, 1, q);
    }
    M(K_FRAMES, 1);
}

int alloc_geo(bd_context *c, Geo &g) {
    int n = c->n, nc = c->n_c, nt = c->n_tail, nl = c->lh_cap;
    int r = 0;
    r |= c->alloc(&g.r, 3 * n); r |= c->alloc(&g.a, 3 * n); r |= c->alloc(&g.b, 3 * n);
    r |= c->alloc(&g.c, 3 * n);
    r |= c->alloc(&g.alpha, n); r |= c->alloc(&g.beta, n); r |= c->alloc(&g.gamma, n);
    r |= c->alloc(&g.length, n);
    r |= c->alloc(&g.a_dna, 3 * nc); r |= c->alloc(&g.b_dna, 3 * nc); r |= c->alloc(&g.c_dna, 3 * nc);
    r |= c->alloc(&g.alpha_p, nc); r |= c->alloc(&g.beta_p, nc); r |= c->alloc(&g.gamma_p, nc);
    r |= c->alloc(&g.r_t, 3 * nt); r |= c->alloc(&g.beta_t, nt);
    r |= c->alloc(&g.r_lh, 3 * nl); r |= c->alloc(&g.beta_lh, nl);
    return r ? BD_E_HIP : BD_OK;
}

int alloc_frc(bd_context *c, Frc &f) {
    int r = 0;
    r |= c->alloc(&f.F, 3 * c->n); r |= c->alloc(&f.T, 3 * c->n);
    r |= c->alloc(&f.tF, 3 * c->n_tail); r |= c->alloc(&f.lF, 3 * c->lh_cap);
    return r ? BD_E_HIP : BD_OK;
}

int check(const bd_context *c) {
    if (!c) return fail(BD_E_ARG, "null context");
    return BD_OK;
}

int sync_check(bd_context *c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipGetLastError());
    return BD_OK;
}

template <typename T>
int d2h(T *dst, const T *src, size_t count) {
    if (!dst || !count) return BD_OK;
    HIPCHK(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDeviceToHost));
    return BD_OK;
}
template <typename T>
int h2d(T *dst, const T *src, size_t count) {
    if (!src || !count) return BD_OK;
    HIPCHK(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return BD_OK;
}

}  // namespace

extern "C" {

const char *bd_last_error(void) { return g_err.c_str(); }
int bd_version(void) { return 1; }

int bd_create(const bd_system *sys, const bd_options *opt, bd_context **out) {
    if (!sys || !out) return fail(BD_E_ARG, "null argument");
    *out = nullptr;
    const int n = sys->n, n_c = sys->n_c, nt = sys->n_tail, Nq = sys->Nq;
    if (n < 3) return fail(BD_E_ARG, "need at least 3 main beads (n=%d)", n);
    if (!sys->type || sys->type[0] != 1)
        return fail(BD_E_ARG, "the first bead must be a core (func.cpp:1062 reads type[i-1])");
    int ncount = 0;
    for (int i = 0; i < n; i++) {
        if (sys->type[i] != 0 && sys->type[i] != 1) return fail(BD_E_ARG, "type[%d]=%d not 0/1", i, sys->type[i]);
        ncount += sys->type[i];
    }
    if (ncount != n_c) return fail(BD_E_ARG, "n_c=%d but type[] has %d cores", n_c, ncount);
    if (Nq < 0 || Nq > 640) return fail(BD_E_ARG, "Nq=%d out of range (at most 640 charges per core)", Nq);
    if (sys->n_lh_g > kMaxGSites) return fail(BD_E_ARG, "n_lh_g=%d exceeds %d", sys->n_lh_g, kMaxGSites);
    const int nlh = sys->n_lh_n + sys->n_lh_c;
    if (sys->n_LH > 0 && sys->n_lh_n != 0)
        return fail(BD_E_ARG, "N-terminal LH beads are not supported (func.cpp:2067 uses i %% n_lh_c)");
    if (sys->n_LH > 0 && sys->n_lh_c < 3) return fail(BD_E_ARG, "n_lh_c must be >= 3");
    if (nt > 0 && (sys->n_t <= 0 || nt != sys->n_t * n_c))
        return fail(BD_E_ARG, "n_tail=%d must equal n_t*n_c (every core carries tails)", nt);
    if (sys->groups_per_core <= 0) return fail(BD_E_ARG, "groups_per_core must be > 0");
    int nlhc = 0;
    for (int k = 0; k < n_c; k++) nlhc += (sys->nc_lh_flag && sys->nc_lh_flag[k] == 1) ? 1 : 0;
    if (sys->n_LH != nlhc * nlh) return fail(BD_E_ARG, "n_LH=%d but %d bound cores x %d beads", sys->n_LH, nlhc, nlh);

    bd_options o{0, BD_CHOL_FP64, 1, 0};
    if (opt) o = *opt;
    HIPCHK(hipSetDevice(o.device));
    bd_context *c = new bd_context();
    c->device = o.device;
    c->use_graph = o.use_graph != 0;
    if (const char *gk = std::getenv("BD_GATE_KB")) c->gate_kb = std::atoi(gk);
    c->n = n; c->n_c = n_c; c->n_tail = nt; c->n_LH = sys->n_LH; c->nlh = nlh;
    c->lh_cap = std::max(sys->n_LH, o.lh_capacity);
    c->hi_all = sys->hi_all ? 1 : 0;
    c->groups_per_core = sys->groups_per_core;
    c->n_D = n + nt + sys->n_LH;
    c->n_D3 = 3 * c->n_D;
    c->n_hi = c->hi_all ? c->n_D : n;
    const int nD_cap = n + nt + c->lh_cap;   // buffers sized for the LH capacity
    const int nhi_cap = c->hi_all ? nD_cap : n;
    auto bail = [&](int r) { bd_destroy(c); return r; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(BD_E_HIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_gate, hipEventDisableTiming) != hipSuccess)
        return bail(fail(BD_E_HIP, "hipEventCreate failed"));
    for (int i = 0; i < 3; i++) {
        hipStream_t b = nullptr;
        if (hipStreamCreateWithFlags(&b, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(BD_E_HIP, "hipStreamCreate failed"));
        c->fstr[0].b[i] = c->fstr[1].b[i] = b;
    }
    for (auto &fs : c->fstr) {
        bool ok = hipEventCreateWithFlags(&fs.fork, hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&fs.mech, hipEventDisableTiming) == hipSuccess;
        for (auto &j : fs.join) ok = ok && hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess;
        if (!ok) return bail(fail(BD_E_HIP, "hipEventCreate failed"));
    }

    // host-side index tables
    std::vector<int> ic_of(n), core_bead, elig(n), lnk, lhc_bead, lh_of_core(n_c, -1);
    int ic = -1;
    for (int i = 0; i < n; i++) {
        if (sys->type[i] == 1) { ic++; core_bead.push_back(i); }
        else lnk.push_back(i);
        ic_of[i] = ic;
        bool e = (i == 0 || sys->type[i - 1] == 0) && (i == n - 1 || sys->type[i + 1] == 0);
        elig[i] = e ? 1 : 0;
    }
    for (int k = 0; k < n_c; k++)
        if (sys->nc_lh_flag && sys->nc_lh_flag[k] == 1) { lh_of_core[k] = (int)lhc_bead.size(); lhc_bead.push_back(core_bead[k]); }
    // frame-pass segments: cut only in front of cores (pass 1 writes linker
    // frames and reads across beads only core frames; pass 2 of bead i reads
    // pass 1 of beads i and i + 1), at most 64 beads each
    std::vector<int> fseg;
    if (n >= 2) {
        int cur = 0;
        fseg.push_back(0);
        while (n - 1 - cur > 64) {
            int cut = -1;
            for (int b = cur + 64; b > cur; b--)
                if (sys->type[b] != 0) { cut = b; break; }
            if (cut < 0) break;
            fseg.push_back(cut);
            cur = cut;
        }
        if (n - 1 - cur <= 64) fseg.push_back(n - 1);
        else fseg.clear();
    }
    std::vector<int> t_rcore(std::max(nt, 1), -1), rt_off(n_c + 1, 0), rt_list;
    {
        std::vector<std::vector<int>> per(n_c);
        for (int j = 0; j < nt; j++) {
            int target = (sys->t_grp[j] - 1) / sys->groups_per_core;
            if (sys->t_fix[j] == 1) {
                if (target < 0 || target >= n_c)
                    return bail(fail(BD_E_ARG, "tail %d maps to core %d ((t_grp-1)/%d) outside [0,%d)", j, target,
                                     sys->groups_per_core, n_c));
                t_rcore[j] = core_bead[target];
                per[target].push_back(j);
            }
        }
        for (int k = 0; k < n_c; k++) {
            rt_off[k + 1] = rt_off[k] + (int)per[k].size();
            rt_list.insert(rt_list.end(), per[k].begin(), per[k].end());
        }
    }
    const int ntpl = sys->n_lh_n + sys->n_lh_g + sys->n_lh_c;
    double rmax_q = 0.0, rmax_g = 0.0;
    for (int q = 0; q < Nq; q++) {
        const double *p = sys->core_pos + 3 * q;
        rmax_q = std::max(rmax_q, std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
    }
    for (int m = 0; m < sys->n_lh_g; m++) {
        const double *p = sys->LH_g_pos + 3 * m;
        rmax_g = std::max(rmax_g, std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
    }

    Stat &s = c->st;
    s.n = n; s.n_c = n_c; s.n_tail = nt; s.n_t = sys->n_t; s.n_lh_n = sys->n_lh_n; s.n_lh_g = sys->n_lh_g;
    s.n_lh_c = sys->n_lh_c; s.nlh = nlh; s.n_LH = sys->n_LH; s.n_lhc = (int)lhc_bead.size(); s.Nq = Nq;
    s.n_lnk = (int)lnk.size();
    s.h = sys->h; s.g = sys->g; s.s = sys->s; s.debyell = sys->debyell; s.debye = sys->debye;
    s.q_l = sys->q_l; s.k_e = sys->k_e; s.k_ex = sys->k_ex;
    s.sig_dd = sys->sig_dd; s.sig_dc = sys->sig_dc; s.sig_cc = sys->sig_cc; s.sig_tt = sys->sig_tt;
    s.sig_tl = sys->sig_tl; s.sig_tc = sys->sig_tc;
    s.a1 = sys->a1; s.a2 = sys->a2; s.s2dt = sys->s2dt; s.del = sys->del; s.time_step = sys->time_step;
    {
        double eps = 1 / (4 * kPI * sys->k_e);  // func.cpp:1874 passes 1/(4*PI*k_e)
        s.c4pe = 4 * kPI * eps;
        for (int i = 0; i < 11; i++) s.ec.c[i] = kExpCoef.c[i];
    }
    s.rmax_q = rmax_q; s.rmax_g = rmax_g;

    int r = 0;
    int *p_i;
    double *p_d;
    const double *pd;
#define UPI(field, src, cnt) do { r |= c->upload(&p_i, (const int *)(src), (cnt)); s.field = p_i; } while (0)
#define UPD(field, src, cnt) do { r |= c->upload(&p_d, (const double *)(src), (cnt)); s.field = p_d; } while (0)
    UPI(type, sys->type, n);
    UPI(ic_of, ic_of.data(), n);
    UPI(core_bead, core_bead.data(), core_bead.size());
    UPI(elig, elig.data(), n);
    UPI(lnk, lnk.data(), lnk.size());
    UPI(fseg, fseg.data(), fseg.size());
    s.nfseg = fseg.empty() ? 0 : (int)fseg.size() - 1;
    UPD(phi_o, sys->phi_o, n);
    UPD(rad, sys->rad, n);
    UPD(cpos, sys->core_pos, 3 * Nq);
    UPD(cq, sys->core_q, Nq);
    UPD(tail_pos, sys->tail_pos, 3 * std::max(sys->n_t, 0));
    UPD(h_t, sys->h_t, nt); UPD(g_t, sys->g_t, nt); UPD(lo_t, sys->lo_t, nt);
    UPD(beta_o_t, sys->beta_o_t, nt); UPD(t_q, sys->t_q, nt); UPD(t_rad, sys->t_rad, nt);
    UPI(t_grp, sys->t_grp, nt); UPI(t_fix, sys->t_fix, nt);
    UPI(t_rcore, t_rcore.data(), nt);
    UPI(rt_off, rt_off.data(), rt_off.size());
    UPI(rt_list, rt_list.data(), rt_list.size());
    UPD(LH_g_pos, sys->LH_g_pos, 3 * sys->n_lh_g);
    UPI(LH_conn, sys->LH_conn, ntpl);
    UPD(LH_q, sys->LH_q, ntpl); UPD(LH_hh, sys->LH_vdw_hh, ntpl); UPD(LH_hc, sys->LH_vdw_hc, ntpl);
    UPD(LH_hl, sys->LH_vdw_hl, ntpl); UPD(LH_ht, sys->LH_vdw_ht, ntpl); UPD(LH_kstr, sys->LH_kstr, ntpl);
    UPD(LH_kben, sys->LH_kben, ntpl); UPD(LH_streq, sys->LH_streq, ntpl); UPD(LH_betaeq, sys->LH_betaeq, ntpl);
    {
        std::vector<int> lb(n_c, 0);
        std::copy(lhc_bead.begin(), lhc_bead.end(), lb.begin());
        UPI(lhc_bead, lb.data(), n_c);
    }
    UPI(lh_of_core, lh_of_core.data(), n_c);
#undef UPI
#undef UPD
    (void)pd;
    // rad_all = [rad; t_rad; lh_rad] (main.cpp:301-322)
    {
        std::vector<double> ra(nD_cap, 0.5);
        for (int i = 0; i < n; i++) ra[i] = sys->rad[i];
        for (int i = 0; i < nt; i++) ra[n + i] = sys->t_rad[i];
        for (int i = 0; i < sys->n_LH; i++) ra[n + nt + i] = sys->lh_rad[i];
        r |= c->upload(&c->rad_all, ra.data(), ra.size());
    }
    if (r) return bail(BD_E_HIP);

    if (alloc_geo(c, c->cur) || alloc_geo(c, c->pred)) return bail(BD_E_HIP);
    if (alloc_frc(c, c->f0) || alloc_frc(c, c->f1)) return bail(BD_E_HIP);
    r |= c->alloc(&c->sites.Q, 3 * (size_t)n_c * Nq);
    r |= c->alloc(&c->sites.Qp, 4 * (size_t)n_c * Nq);
    r |= c->alloc(&c->sites.G, 3 * (size_t)n_c * std::max(sys->n_lh_g, 1));
    r |= c->alloc(&c->sites.ch_t, nchunks(nt));
    r |= c->alloc(&c->sites.ch_lh, nchunks(c->lh_cap));
    r |= c->alloc(&c->sites.ch_g, nchunks(n_c * std::max(sys->n_lh_g, 1)));
    r |= c->alloc(&c->sites.ch_l, nchunks(n));
    {
        const size_t cap[NSETS] = {(size_t)nt, (size_t)c->lh_cap, (size_t)n_c * std::max(sys->n_lh_g, 1), (size_t)n};
        for (int k = 0; k < NSETS; k++) {
            CellGrid &cg = c->sites.cg[k];
            r |= c->alloc(&cg.hdr, 4);
            r |= c->alloc(&cg.dim, 3);
            r |= c->alloc(&cg.start, (size_t)kCellMax + 1);
            r |= c->alloc(&cg.cursor, (size_t)kCellMax);
            r |= c->alloc(&cg.members, cap[k]);
            r |= c->alloc(&cg.cid, cap[k]);
        }
        // the grid build costs more than it saves on small fibers (measured:
        // 100 nucleosomes 987 steps/s with chunk spheres, 923 with cells);
        // cells from 20000 tail and LH beads up (1000 nucleosomes: the tail
        // and LH kernels 1.56 -> 0.66 ms per step)
        c->sites.use_cells = (nt + c->lh_cap) >= 20000 ? 1 : 0;
        if (const char *ev = getenv("BD_CELLS")) c->sites.use_cells = atoi(ev) != 0;
    }
    // core-core scratch
    // one wavefront per (pair, 64 charges of the first core, about 80
    // charges of the second)
    c->fw.cc_nkb = (Nq + 63) / 64;
    c->fw.cc_nlc = std::max(1, (Nq + 79) / 80);
    {
        long long npairs = (long long)n_c * (n_c - 1) / 2;
        long long units = npairs * c->fw.cc_nkb * c->fw.cc_nlc;
        c->fw.cc_grid = (int)std::max(1LL, std::min((units + 3) / 4, 2048LL));
        r |= c->alloc(&c->fw.cc_valid, (size_t)n_c * n_c);
        r |= c->alloc(&c->fw.cc_list, 2 * (size_t)std::max(npairs, 1LL));
        r |= c->alloc(&c->fw.cc_count, 1);
        r |= c->alloc(&c->fw.cc_slots, (size_t)n_c * n_c * c->fw.cc_nkb * c->fw.cc_nlc * 10);
        c->fw.core_split[0] = 1;
        c->fw.core_split[1] = 2;
        c->fw.core_split[2] = 1;
        r |= c->alloc(&c->fw.core_part,
                      (size_t)n_c * (c->fw.core_split[0] + c->fw.core_split[1] + c->fw.core_split[2]) * 8);
        r |= c->alloc(&c->fw.core_cnt, n_c);
    }
    // Cholesky workspace
    c->chol.m = 3 * c->n_hi;
    c->chol.mpad = ((c->chol.m + 63) / 64) * 64;
    c->chol.ld = ((3 * nhi_cap + 63) / 64) * 64;
    if (o.chol_mode != BD_CHOL_FP64 && o.chol_mode != BD_CHOL_FP32 && o.chol_mode != BD_CHOL_BF16)
        return bail(fail(BD_E_ARG, "unsupported chol_mode %d", o.chol_mode));
    c->chol.mode = o.chol_mode;
    c->chol.panel_tiles = 0;
    if (const char *ev = getenv("BD_PANEL_TILES")) c->chol.panel_tiles = atoi(ev);
    {
        int ncu = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, o.device) == hipSuccess && prop.multiProcessorCount > 0)
            ncu = prop.multiProcessorCount;
        c->chol.split_tiles = ncu * 19 / 10;
    }
    r |= c->alloc(&c->chol.A, (size_t)c->chol.ld * c->chol.ld);
    r |= c->alloc(&c->chol.Linv, (size_t)(c->chol.ld / 64) * 64 * 64);
    r |= c->alloc(&c->chol.neg, 1);
    c->chol.refine = 1;
    if (const char *ev = getenv("BD_REFINE")) c->chol.refine = atoi(ev) != 0;
    r |= c->alloc(&c->chol.xs, (size_t)c->chol.ld);
    r |= c->alloc(&c->chol.rv, 2 * (size_t)c->chol.ld);
    r |= c->alloc(&c->chol.pr, (size_t)c->chol.ld);
    // super-panel panel tiles packed (FP64 / FP32) or split (BF16) once for
    // the trailing update; BD_PANEL_PACK=0 reads the parked tiles directly
    c->chol.pack_min_rows = getenv("BD_PANEL_PACK_MIN_ROWS") ? atoi(getenv("BD_PANEL_PACK_MIN_ROWS")) : -1;
    if (!getenv("BD_PANEL_PACK") || atoi(getenv("BD_PANEL_PACK")) != 0) {
        double *spd = nullptr;
        c->chol.sp_nb = 8;
        r |= c->alloc(&spd, (split_bytes(c->chol.ld, c->chol.sp_nb) + 7) / 8);
        c->chol.sp = spd;
    }
    r |= c->alloc(&c->rd, 3 * nD_cap);
    r |= c->alloc(&c->p, 3 * nD_cap);
    r |= c->alloc(&c->rr, 3 * n);
    r |= c->alloc(&c->energy, 1);
    c->ne = energy_slots(s);
    {
        Stat cap = s;
        cap.n_LH = c->lh_cap;
        r |= c->alloc(&c->Epart, energy_slots(cap));
    }
    if (r) return bail(BD_E_HIP);
    launch_fill(c->rr, 3 * n, 0.0, c->stream);
    launch_fill(c->p, 3 * nD_cap, 0.0, c->stream);
    launch_fill_int(c->fw.core_cnt, n_c, 0, c->stream);
    // entries the step never writes (e.g. the exit frame of a final core)
    // read as 0, like the host arrays of the reference driver
    for (Geo *g : {&c->cur, &c->pred}) {
        launch_fill(g->alpha, n, 0.0, c->stream); launch_fill(g->beta, n, 0.0, c->stream);
        launch_fill(g->gamma, n, 0.0, c->stream); launch_fill(g->length, n, 0.0, c->stream);
        launch_fill(g->a_dna, 3 * n_c, 0.0, c->stream); launch_fill(g->b_dna, 3 * n_c, 0.0, c->stream);
        launch_fill(g->c_dna, 3 * n_c, 0.0, c->stream); launch_fill(g->alpha_p, n_c, 0.0, c->stream);
        launch_fill(g->beta_p, n_c, 0.0, c->stream); launch_fill(g->gamma_p, n_c, 0.0, c->stream);
        launch_fill(g->beta_t, nt, 0.0, c->stream); launch_fill(g->beta_lh, c->lh_cap, 0.0, c->stream);
    }
    c->nc_lh_flag.assign(n_c, 0);
    for (int k = 0; k < n_c; k++) c->nc_lh_flag[k] = sys->nc_lh_flag ? sys->nc_lh_flag[k] : 0;
    c->core_bead_host = core_bead;
    int e = sync_check(c);
    if (e) return bail(e);
    *out = c;
    return BD_OK;
}

int bd_destroy(bd_context *c) {
    if (!c) return BD_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (auto &pe : c->pending) (void)hipEventDestroy(pe.second);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->batch) (void)hipFree(c->batch);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_gate) (void)hipEventDestroy(c->ev_gate);
    for (int i = 0; i < 3; i++)
        if (c->fstr[0].b[i]) {
            (void)hipStreamSynchronize(c->fstr[0].b[i]);
            (void)hipStreamDestroy(c->fstr[0].b[i]);
        }
    for (auto &fs : c->fstr) {
        if (fs.fork) (void)hipEventDestroy(fs.fork);
        if (fs.mech) (void)hipEventDestroy(fs.mech);
        for (auto &j : fs.join)
            if (j) (void)hipEventDestroy(j);
    }
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BD_OK;
}

int bd_get_sizes(bd_context *c, int *n, int *n_tail, int *n_LH, int *n_hi) {
    if (check(c)) return BD_E_ARG;
    if (n) *n = c->n;
    if (n_tail) *n_tail = c->n_tail;
    if (n_LH) *n_LH = c->n_LH;
    if (n_hi) *n_hi = c->n_hi;
    return BD_OK;
}

int bd_set_lh(bd_context *c, const int *nc_lh_flag, int n_LH, const double *r_lh, const double *lh_rad) {
    if (check(c) || !nc_lh_flag) return fail(BD_E_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    const int n_c = c->n_c, nlh = c->nlh;
    std::vector<int> lhc, loc(n_c, -1);
    for (int k = 0; k < n_c; k++)
        if (nc_lh_flag[k] == 1) { loc[k] = (int)lhc.size(); lhc.push_back(c->core_bead_host[k]); }
    if (n_LH != (int)lhc.size() * nlh)
        return fail(BD_E_ARG, "n_LH=%d but %zu bound cores x %d beads", n_LH, lhc.size(), nlh);
    if (n_LH > c->lh_cap) return fail(BD_E_CAPACITY, "n_LH=%d exceeds the LH capacity %d", n_LH, c->lh_cap);
    if (n_LH && (!r_lh || !lh_rad)) return fail(BD_E_ARG, "r_lh and lh_rad are required");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n_LH == c->n_LH && std::equal(nc_lh_flag, nc_lh_flag + n_c, c->nc_lh_flag.begin())) {
        // same binding: positions and radii only (the captured step stays valid)
        int e = h2d(c->cur.r_lh, r_lh, 3 * n_LH);
        e |= h2d(c->rad_all + c->n + c->n_tail, lh_rad, n_LH);
        return e ? BD_E_HIP : BD_OK;
    }
    const int n_hi_old = c->n_hi;
    lhc.resize(n_c, 0);
    HIPCHK(hipMemcpy((void *)c->st.lhc_bead, lhc.data(), sizeof(int) * n_c, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void *)c->st.lh_of_core, loc.data(), sizeof(int) * n_c, hipMemcpyHostToDevice));
    int e = h2d(c->cur.r_lh, r_lh, 3 * n_LH);
    e |= h2d(c->rad_all + c->n + c->n_tail, lh_rad, n_LH);
    if (e) return BD_E_HIP;
    int nb = 0;
    for (int k = 0; k < n_c; k++) nb += nc_lh_flag[k] == 1;
    c->nc_lh_flag.assign(nc_lh_flag, nc_lh_flag + n_c);
    c->n_LH = n_LH;
    c->st.n_LH = n_LH;
    c->st.n_lhc = nb;
    c->n_D = c->n + c->n_tail + n_LH;
    c->n_D3 = 3 * c->n_D;
    c->n_hi = c->hi_all ? c->n_D : c->n;
    c->chol.m = 3 * c->n_hi;
    c->chol.mpad = ((c->chol.m + 63) / 64) * 64;
    c->ne = energy_slots(c->st);
    drop_graphs(c);
    if (c->n_hi != n_hi_old) {
        c->factor_fresh = false;
        c->rp_count = 0;
    }
    if (n_LH) launch_betas(c->st, c->cur, 0, 1, c->stream);
    return sync_check(c);
}

int bd_set_integrator(bd_context *c, double a1, double a2, double s2dt, double del, double time_step) {
    if (check(c)) return BD_E_ARG;
    c->st.a1 = a1; c->st.a2 = a2; c->st.s2dt = s2dt; c->st.del = del; c->st.time_step = time_step;
    if (c->graph[0] || c->graph[1]) {
        HIPCHK(hipStreamSynchronize(c->stream));
        drop_graphs(c);
    }
    return BD_OK;
}

int bd_set_state(bd_context *c, const double *r, const double *a, const double *b, const double *cc,
                 const double *r_t, const double *r_lh, int rebuild) {
    if (check(c)) return BD_E_ARG;
    if (!r || !a || !b || !cc) return fail(BD_E_ARG, "r, a, b, c are required");
    if (c->n_tail && !r_t) return fail(BD_E_ARG, "r_t is required (n_tail > 0)");
    if (c->n_LH && !r_lh) return fail(BD_E_ARG, "r_lh is required (n_LH > 0)");
    HIPCHK(hipSetDevice(c->device));
    int e = 0;
    e |= h2d(c->cur.r, r, 3 * c->n); e |= h2d(c->cur.a, a, 3 * c->n);
    e |= h2d(c->cur.b, b, 3 * c->n); e |= h2d(c->cur.c, cc, 3 * c->n);
    e |= h2d(c->cur.r_t, r_t, 3 * c->n_tail); e |= h2d(c->cur.r_lh, r_lh, 3 * c->n_LH);
    if (e) return BD_E_HIP;
    if (rebuild) {
        launch_frames(c->st, c->cur, c->stream);
        launch_betas(c->st, c->cur, 1, 1, c->stream);
    }
    c->state_set = true;
    c->factor_fresh = false;
    c->rp_count = 0;
    return sync_check(c);
}

int bd_get_state(bd_context *c, double *r, double *a, double *b, double *cc, double *r_t, double *r_lh) {
    if (check(c)) return BD_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    int e = 0;
    e |= d2h(r, c->cur.r, 3 * c->n); e |= d2h(a, c->cur.a, 3 * c->n);
    e |= d2h(b, c->cur.b, 3 * c->n); e |= d2h(cc, c->cur.c, 3 * c->n);
    e |= d2h(r_t, c->cur.r_t, 3 * c->n_tail); e |= d2h(r_lh, c->cur.r_lh, 3 * c->n_LH);
    return e ? BD_E_HIP : BD_OK;
}

int bd_set_derived(bd_context *c, const bd_derived *d) {
    if (check(c) || !d) return fail(BD_E_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    int n = c->n, nc = c->n_c, e = 0;
    e |= h2d(c->cur.alpha, d->alpha, n); e |= h2d(c->cur.beta, d->beta, n);
    e |= h2d(c->cur.gamma, d->gamma, n); e |= h2d(c->cur.length, d->length, n);
    e |= h2d(c->cur.a_dna, d->a_dna, 3 * nc); e |= h2d(c->cur.b_dna, d->b_dna, 3 * nc);
    e |= h2d(c->cur.c_dna, d->c_dna, 3 * nc); e |= h2d(c->cur.alpha_p, d->alpha_p, nc);
    e |= h2d(c->cur.beta_p, d->beta_p, nc); e |= h2d(c->cur.gamma_p, d->gamma_p, nc);
    e |= h2d(c->cur.beta_t, d->beta_t, c->n_tail); e |= h2d(c->cur.beta_lh, d->beta_lh, c->n_LH);
    return e ? BD_E_HIP : BD_OK;
}

int bd_get_derived(bd_context *c, bd_derived *d) {
    if (check(c) || !d) return fail(BD_E_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    int n = c->n, nc = c->n_c, e = 0;
    e |= d2h(d->alpha, c->cur.alpha, n); e |= d2h(d->beta, c->cur.beta, n);
    e |= d2h(d->gamma, c->cur.gamma, n); e |= d2h(d->length, c->cur.length, n);
    e |= d2h(d->a_dna, c->cur.a_dna, 3 * nc); e |= d2h(d->b_dna, c->cur.b_dna, 3 * nc);
    e |= d2h(d->c_dna, c->cur.c_dna, 3 * nc); e |= d2h(d->alpha_p, c->cur.alpha_p, nc);
    e |= d2h(d->beta_p, c->cur.beta_p, nc); e |= d2h(d->gamma_p, c->cur.gamma_p, nc);
    e |= d2h(d->beta_t, c->cur.beta_t, c->n_tail); e |= d2h(d->beta_lh, c->cur.beta_lh, c->n_LH);
    return e ? BD_E_HIP : BD_OK;
}

static int run_step(bd_context *c) {
    // a refinement switched on since the last assembly needs a new one
    const bool refactor = !c->factor_fresh || c->rp_count % c->rp_freq == 0 ||
                          (refine_active(c->chol) && !c->xs_valid);
    c->rp_count++;
    c->factor_fresh = true;
    if (c->timing || !c->use_graph) {
        enqueue_step(c, refactor);
        HIPCHK(hipGetLastError());
        if (c->timing) collect_timing(c);
        return BD_OK;
    }
    hipGraphExec_t &ge = c->graph[refactor ? 0 : 1];
    if (!ge) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_step(c, refactor);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess) return fail(BD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(BD_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    HIPCHK(hipGraphLaunch(ge, c->stream));
    return BD_OK;
}

int bd_step(bd_context *c, const double *rr, const double *p, double *energy) {
    if (check(c)) return BD_E_ARG;
    if (!c->state_set) return fail(BD_E_STATE, "bd_set_state must be called before bd_step");
    if (!rr || !p) return fail(BD_E_ARG, "rr and p are required");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->rr, rr, sizeof(double) * 3 * c->n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->p, p, sizeof(double) * c->n_D3, hipMemcpyHostToDevice, c->stream));
    int e = run_step(c);
    if (e) return e;
    if (energy) HIPCHK(hipMemcpyAsync(energy, c->energy, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return sync_check(c);
}

int bd_step_device(bd_context *c, const double *d_rr, const double *d_p, double *d_energy) {
    if (check(c)) return BD_E_ARG;
    if (!c->state_set) return fail(BD_E_STATE, "bd_set_state must be called before bd_step_device");
    if (!d_rr || !d_p) return fail(BD_E_ARG, "d_rr and d_p are required");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->rr, d_rr, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->p, d_p, sizeof(double) * c->n_D3, hipMemcpyDeviceToDevice, c->stream));
    int e = run_step(c);
    if (e) return e;
    if (d_energy)
        HIPCHK(hipMemcpyAsync(d_energy, c->energy, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return BD_OK;
}

int bd_run_steps(bd_context *c, int k, const double *rr, const double *p, double *energies) {
    if (check(c)) return BD_E_ARG;
    if (!c->state_set) return fail(BD_E_STATE, "bd_set_state must be called before bd_run_steps");
    if (k < 0 || (k && (!rr || !p))) return fail(BD_E_ARG, "k >= 0 and rr, p are required");
    if (!k) return BD_OK;
    HIPCHK(hipSetDevice(c->device));
    const size_t n3 = 3 * (size_t)c->n, nd3 = (size_t)c->n_D3;
    const size_t need = (size_t)k * (n3 + nd3 + 1);
    if (need > c->batch_cap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->batch) (void)hipFree(c->batch);
        c->batch = nullptr;
        c->batch_cap = 0;
        hipError_t e = hipMalloc((void **)&c->batch, need * sizeof(double));
        if (e != hipSuccess) return fail(BD_E_HIP, "hipMalloc(%zu): %s", need * sizeof(double), hipGetErrorString(e));
        c->batch_cap = need;
    }
    double *d_rr = c->batch, *d_p = d_rr + (size_t)k * n3, *d_e = d_p + (size_t)k * nd3;
    HIPCHK(hipMemcpyAsync(d_rr, rr, sizeof(double) * k * n3, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_p, p, sizeof(double) * k * nd3, hipMemcpyHostToDevice, c->stream));
    for (int s = 0; s < k; s++) {
        HIPCHK(hipMemcpyAsync(c->rr, d_rr + s * n3, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->p, d_p + s * nd3, sizeof(double) * nd3, hipMemcpyDeviceToDevice, c->stream));
        int e = run_step(c);
        if (e) return e;
        HIPCHK(hipMemcpyAsync(d_e + s, c->energy, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    if (energies) HIPCHK(hipMemcpyAsync(energies, d_e, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
    return sync_check(c);
}

int bd_forces(bd_context *c, int terms, double *F, double *T, double *tF, double *lhF, double *energy) {
    if (check(c)) return BD_E_ARG;
    if (!c->state_set) return fail(BD_E_STATE, "bd_set_state must be called first");
    HIPCHK(hipSetDevice(c->device));
    enqueue_forces(c, c->cur, c->f0, c->Epart, terms, 0);
    launch_energy_reduce(c->Epart, c->ne, c->energy, c->stream);
    HIPCHK(hipGetLastError());
    if (c->timing) collect_timing(c);
    int e = sync_check(c);
    if (e) return e;
    e |= d2h(F, c->f0.F, 3 * c->n); e |= d2h(T, c->f0.T, 3 * c->n);
    e |= d2h(tF, c->f0.tF, 3 * c->n_tail); e |= d2h(lhF, c->f0.lF, 3 * c->n_LH);
    e |= d2h(energy, c->energy, 1);
    return e ? BD_E_HIP : BD_OK;
}

int bd_factor(bd_context *c) {
    if (check(c)) return BD_E_ARG;
    if (!c->state_set) return fail(BD_E_STATE, "bd_set_state must be called first");
    HIPCHK(hipSetDevice(c->device));
    launch_betas(c->st, c->cur, 0, 1, c->stream);
    enqueue_factor(c, c->cur, c->stream);
    HIPCHK(hipGetLastError());
    if (c->timing) collect_timing(c);
    return sync_check(c);
}

int bd_get_factor(bd_context *c, double *L, int *neg) {
    if (check(c)) return BD_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int m = c->chol.m, mp = c->chol.mpad;
    if (L) {
        // the factor's storage: diagonal tiles in place, the off-diagonal
        // tile (ib, kb) parked in the upper block (kb, ib)
        std::vector<double> raw((size_t)mp * mp);
        HIPCHK(hipMemcpy2D(raw.data(), sizeof(double) * mp, c->chol.A, sizeof(double) * c->chol.ld,
                           sizeof(double) * mp, mp, hipMemcpyDeviceToHost));
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                const int ib = i / 64, kb = j / 64;
                double v = 0.0;
                if (kb < ib) v = raw[(size_t)(kb * 64 + i % 64) * mp + ib * 64 + j % 64];
                else if (j <= i) v = raw[(size_t)i * mp + j];
                L[(size_t)i * m + j] = v;
            }
    }
    if (neg) HIPCHK(hipMemcpy(neg, c->chol.neg, sizeof(int), hipMemcpyDeviceToHost));
    return BD_OK;
}

int bd_get_tensor(bd_context *c, double *D) {
    if (check(c) || !D) return fail(BD_E_ARG, "null argument");
    if (!c->state_set) return fail(BD_E_STATE, "bd_set_state must be called first");
    HIPCHK(hipSetDevice(c->device));
    launch_rpy(c->st, c->n_hi, c->cur, c->rad_all, c->chol, c->stream);
    c->xs_valid = refine_active(c->chol);
    c->factor_fresh = false;  // D now overwrites the factor's storage
    int e = sync_check(c);
    if (e) return e;
    const int m = c->chol.m;
    HIPCHK(hipMemcpy2D(D, sizeof(double) * m, c->chol.A, sizeof(double) * c->chol.ld, sizeof(double) * m, m,
                       hipMemcpyDeviceToHost));
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++) D[(size_t)i * m + j] = D[(size_t)j * m + i];
    return BD_OK;
}

int bd_noise_displacement(bd_context *c, const double *p, double *rd) {
    if (check(c) || !p || !rd) return fail(BD_E_ARG, "null argument");
    const bool refine = refine_active(c->chol);
    if (refine && !c->xs_valid)
        return fail(BD_E_STATE, "the factor was formed without the refinement (FP64 mode or refinement off): "
                                "call bd_factor again");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->p, p, sizeof(double) * c->n_D3, hipMemcpyHostToDevice, c->stream));
    const double *du = nullptr;
    if (refine) {
        launch_refine(c->st, c->chol, c->p, c->rad_all, c->stream);
        du = c->chol.pr;
    }
    launch_noise(c->st, c->n_hi, c->n_D3, c->chol, c->p, c->rad_all, c->rd, c->stream, du);
    int e = sync_check(c);
    if (e) return e;
    return d2h(rd, c->rd, c->n_D3);
}

int bd_set_chol_mode(bd_context *c, int mode) {
    if (check(c)) return BD_E_ARG;
    if (mode != BD_CHOL_FP64 && mode != BD_CHOL_FP32 && mode != BD_CHOL_BF16)
        return fail(BD_E_ARG, "unsupported chol_mode %d", mode);
    if (mode != c->chol.mode) {
        HIPCHK(hipStreamSynchronize(c->stream));
        drop_graphs(c);
        c->factor_fresh = false;
        c->rp_count = 0;
    }
    c->chol.mode = mode;
    return BD_OK;
}

int bd_set_refine(bd_context *c, int enable) {
    if (check(c)) return BD_E_ARG;
    const int v = enable != 0;
    if (v != c->chol.refine) {
        HIPCHK(hipStreamSynchronize(c->stream));
        drop_graphs(c);
    }
    c->chol.refine = v;
    return BD_OK;
}

int bd_set_chol_schedule(bd_context *c, int panel_tiles) {
    if (check(c)) return BD_E_ARG;
    if (panel_tiles < -1 || panel_tiles == 1) return fail(BD_E_ARG, "panel_tiles must be -1, 0 or >= 2 (got %d)", panel_tiles);
    if (panel_tiles != c->chol.panel_tiles) {
        HIPCHK(hipStreamSynchronize(c->stream));
        drop_graphs(c);
        c->factor_fresh = false;
        c->rp_count = 0;
    }
    c->chol.panel_tiles = panel_tiles;
    return BD_OK;
}

int bd_set_cell_list(bd_context *c, int enable) {
    if (check(c)) return BD_E_ARG;
    const int v = enable != 0;
    if (v != c->sites.use_cells) {
        HIPCHK(hipStreamSynchronize(c->stream));
        drop_graphs(c);
    }
    c->sites.use_cells = v;
    return BD_OK;
}

int bd_get_chol_schedule(bd_context *c, int *panel_tiles_used) {
    if (check(c) || !panel_tiles_used) return fail(BD_E_ARG, "null argument");
    *panel_tiles_used = chol_panel_tiles(c->chol);
    return BD_OK;
}

int bd_set_frequency_rp(bd_context *c, int frequency_rp) {
    if (check(c)) return BD_E_ARG;
    if (frequency_rp < 1) return fail(BD_E_ARG, "frequency_RP must be >= 1 (got %d)", frequency_rp);
    c->rp_freq = frequency_rp;
    c->rp_count = 0;
    c->factor_fresh = false;
    return BD_OK;
}

int bd_set_timing(bd_context *c, int enable) {
    if (check(c)) return BD_E_ARG;
    c->timing = enable != 0;
    if (c->timing)
        for (int k = 0; k < K_COUNT; k++) { c->total_ms[k] = 0; c->launches[k] = 0; }
    return BD_OK;
}

int bd_get_timing(bd_context *c, bd_kernel_time *out, int max_entries, int *n_entries) {
    if (check(c)) return BD_E_ARG;
    int k = 0;
    for (int i = 0; i < K_COUNT && k < max_entries; i++) {
        if (!c->launches[i]) continue;
        std::snprintf(out[k].name, sizeof out[k].name, "%s", kKernelNames[i]);
        out[k].total_ms = c->total_ms[i];
        out[k].launches = c->launches[i];
        k++;
    }
    if (n_entries) *n_entries = k;
    return BD_OK;
}

int bd_get_work_counts(bd_context *c, bd_work_counts *w) {
    if (check(c) || !w) return fail(BD_E_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    int cnt = 0;
    HIPCHK(hipMemcpy(&cnt, c->fw.cc_count, sizeof(int), hipMemcpyDeviceToHost));
    w->core_pairs = cnt;
    w->core_charge_pairs = (long long)cnt * c->st.Nq * c->st.Nq;
    w->core_linker_pairs = -1;
    return BD_OK;
}

int bd_device_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(BD_E_ARG, "null argument");
    HIPCHK(hipMalloc(ptr, bytes));
    return BD_OK;
}
int bd_device_free(void *ptr) {
    HIPCHK(hipFree(ptr));
    return BD_OK;
}
int bd_memcpy_htod(void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BD_OK;
}
int bd_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BD_OK;
}
int bd_device_copy(void *dst, const void *src, size_t bytes) {
    if (!dst || !src) return fail(BD_E_ARG, "null argument");
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return fail(BD_E_ARG, "bd_device_copy needs 16-byte aligned pointers and sizes");
    const long long nw = (long long)(bytes / 16);
    if (nw == 0) return BD_OK;
    const long long blocks = std::min<long long>((nw + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, 0, (uint4 *)dst,
                       (const uint4 *)src, nw);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return BD_OK;
}
int bd_synchronize(bd_context *c) {
    if (check(c)) return BD_E_ARG;
    return sync_check(c);
}

}  // extern "C"
