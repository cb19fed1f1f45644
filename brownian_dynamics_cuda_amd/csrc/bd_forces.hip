// bd_forces.hip -- one force pass of the BD step on gfx950.
//
// Restates the CPU force routines of the reference (func.cpp:951-2781) as
// receiver-side ("gather") kernels: each bead's force and torque is summed
// by exactly one thread or one workgroup, so the result is a fixed-order
// FP64 sum with no atomics.  Pairs that contribute to two receivers are
// evaluated from both sides (cheap short-range terms) except the
// dominant core-core charge sum, which is evaluated once per core pair
// into a slot matrix and gathered in index order by the core kernel.
//
// Energies: every pair/bond energy is counted once, by a fixed owner
// (see the comment at each term), into per-thread slots of Epart.
#include "bd_dev.h"
#include "bd_kernels.h"
#include "bd_frames.h"

namespace bd {

// ---------------------------------------------------------------------------
// world positions of core charges and globular-head sites
// ---------------------------------------------------------------------------
// head site t (globular-domain site m of LH-bearing core k) in the lab frame
__device__ __forceinline__ double3 head_site(const Stat &s, const Geo &g, int t) {
    const int k = t / s.n_lh_g, m = t % s.n_lh_g, cb = s.lhc_bead[k];
    return body_to_world(ld3(g.r + 3 * cb), ld3(g.a + 3 * cb), ld3(g.b + 3 * cb), ld3(g.c + 3 * cb),
                         ld3(s.LH_g_pos + 3 * m));
}

// ---------------------------------------------------------------------------
// linker-DNA mechanics, mechanical_force_and_torque (func.cpp:951-1289)
// ---------------------------------------------------------------------------
struct SegOut {
    double3 rs, rf;          // exit point of bead i, entry point of bead i+1
    double3 s1, s2, b1, b2, t1, t2;  // stretch / bend / twist force pairs
    double E;
};

__device__ __forceinline__ double3 entry_pt(int t, double3 r, double3 a, double3 b, double3 c) {
    if (t == 0) return r;
    const double si = sin(kTHETA), co = cos(kTHETA);
    double3 bm = make_double3(-si * a.x + co * b.x, -si * a.y + co * b.y, -si * a.z + co * b.z);
    return make_double3(r.x - kRO * bm.x + kD1 * c.x, r.y - kRO * bm.y + kD1 * c.y,
                        r.z - kRO * bm.z + kD1 * c.z);
}
__device__ __forceinline__ double3 exit_pt(int t, double3 r, double3 b, double3 c) {
    if (t == 0) return r;
    return make_double3(r.x - (kRO * b.x + kD1 * c.x), r.y - (kRO * b.y + kD1 * c.y),
                        r.z - (kRO * b.z + kD1 * c.z));
}

__device__ void segment(const Stat &s, const Geo &g, int i, SegOut &o) {
    const double si = sin(kTHETA), co = cos(kTHETA);
    const int ti = s.type[i], tf = s.type[i + 1];
    double3 ri = ld3(g.r + 3 * i), ai = ld3(g.a + 3 * i), bi = ld3(g.b + 3 * i), ci = ld3(g.c + 3 * i);
    double3 rn = ld3(g.r + 3 * i + 3), an = ld3(g.a + 3 * i + 3), bn = ld3(g.b + 3 * i + 3),
            cn = ld3(g.c + 3 * i + 3);
    o.rf = entry_pt(tf, rn, an, bn, cn);
    o.rs = exit_pt(ti, ri, bi, ci);
    o.E = 0.0;
    // stretching (func.cpp:978-1018)
    o.s1 = stretch_f1(s.h, o.rs, o.rf, kLO, o.E);
    o.s2 = make_double3(-o.s1.x, -o.s1.y, -o.s1.z);
    const int ic = s.ic_of[i];
    // bending (func.cpp:1022-1105, Bending_force_projection 198-232)
    double3 af, ab, ao;
    double beta_b;
    if (ti == 0) {
        af = (tf == 0) ? an : make_double3(co * an.x + si * bn.x, co * an.y + si * bn.y, co * an.z + si * bn.z);
        ab = (s.type[i - 1] == 0) ? ld3(g.a + 3 * i - 3) : ld3(g.a_dna + 3 * ic);
        ao = ai;
        beta_b = g.beta[i - 1];
    } else {
        af = an;
        ab = ai;
        ao = ld3(g.a_dna + 3 * ic);
        beta_b = g.beta_p[ic];
    }
    {
        double bb = g.beta[i], len = g.length[i];
        o.E += s.g * (bb - beta_b) * (bb - beta_b) / 2;
        double s1, c1, s2, c2;
        sincos_small(bb, s1, c1);
        sincos_small(beta_b, s2, c2);
        double g1 = (bb >= 1e-10) ? bb / (s1 * len) : 1.0 / len;
        double g2 = (beta_b >= 1e-10) ? beta_b / (s2 * len) : 1.0 / len;
        double3 A = make_double3(g1 * (af.x - c1 * ao.x), g1 * (af.y - c1 * ao.y), g1 * (af.z - c1 * ao.z));
        double3 B = make_double3(g2 * (ab.x - c2 * ao.x), g2 * (ab.y - c2 * ao.y), g2 * (ab.z - c2 * ao.z));
        o.b1 = make_double3(-s.g * (A.x + B.x), -s.g * (A.y + B.y), -s.g * (A.z + B.z));
        o.b2 = make_double3(s.g * (A.x + B.x), s.g * (A.y + B.y), s.g * (A.z + B.z));
    }
    // twisting (func.cpp:1110-1179, Twisting_force_projection 234-260)
    {
        double al_b, be_b, ga_b, ph_b, ga_n;
        double3 bo, cc;
        if (ti == 0) {
            al_b = g.alpha[i - 1]; be_b = g.beta[i - 1]; ga_b = g.gamma[i - 1];
            ph_b = s.phi_o[i - 1]; ga_n = g.gamma[i - 1];
            bo = bi; cc = ci;
        } else {
            al_b = g.alpha[i]; be_b = g.beta_p[ic]; ga_b = g.gamma[i];
            ph_b = s.phi_o[i]; ga_n = g.gamma_p[ic];
            bo = ld3(g.b_dna + 3 * ic); cc = ld3(g.c_dna + 3 * ic);
        }
        double tw = g.alpha[i] + g.gamma[i] - s.phi_o[i];
        o.E += s.s * (tw * tw) / 2;
        double g1 = tw * tan_small(0.5 * g.beta[i]) / g.length[i];
        double ca, sa, cg, sg;
        sincos_small(g.alpha[i], sa, ca);
        double g2 = (al_b + ga_b - ph_b) * tan_small(0.5 * be_b) / g.length[i];
        sincos_small(ga_n, sg, cg);
        double3 ch = make_double3(g1 * (ca * cc.x - sa * bo.x), g1 * (ca * cc.y - sa * bo.y), g1 * (ca * cc.z - sa * bo.z));
        double3 zh = make_double3(g2 * (cg * cc.x + sg * bo.x), g2 * (cg * cc.y + sg * bo.y), g2 * (cg * cc.z + sg * bo.z));
        o.t1 = make_double3(s.s * (ch.x + zh.x), s.s * (ch.y + zh.y), s.s * (ch.z + zh.z));
        o.t2 = make_double3(-s.s * (ch.x + zh.x), -s.s * (ch.y + zh.y), -s.s * (ch.z + zh.z));
    }
}

// torque_due_to_force (func.cpp:374-394) about bead centre pc
__device__ __forceinline__ double3 torque_abs(double3 f, double3 pf, double3 pc, double3 a, double3 b, double3 c) {
    double3 d = pf - pc;
    return torque_rel(f, make_double3(dot(d, a), dot(d, b), dot(d, c)), a, b, c);
}

// Thread per main bead j: assigns F[j], T[j] (this kernel opens the pass).
// Bounding sphere of chunk c of kChunk consecutive partners (centroid and
// the largest distance from it), by one wavefront, a member per lane; idx
// maps a partner to its bead (NULL: the partner is the bead).  Returns the
// sphere in every lane.
// sphere of the cnt points p held by lanes 0 .. cnt-1 of a wavefront
__device__ __forceinline__ double4 sphere_of(double3 p, bool in, int cnt) {
    double3 m = p;
    for (int o = 32; o > 0; o >>= 1) {
        m.x += __shfl_xor(m.x, o, 64);
        m.y += __shfl_xor(m.y, o, 64);
        m.z += __shfl_xor(m.z, o, 64);
    }
    const double w = 1.0 / cnt;
    m = make_double3(m.x * w, m.y * w, m.z * w);
    const double3 d = p - m;
    double r2 = in ? dot(d, d) : 0.0;
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    return make_double4(m.x, m.y, m.z, sqrt(r2));
}
__device__ __forceinline__ double4 chunk_sphere(const double *x, const int *idx, int c, int n) {
    const int lane = threadIdx.x & 63, k = c * kChunk + lane;
    const int cnt = min(n - c * kChunk, kChunk);
    const bool in = lane < cnt;
    const double3 p = in ? ld3(x + 3 * (idx ? idx[k] : k)) : zero3();
    return sphere_of(p, in, cnt);
}

// Chunk spheres of the tail beads, free LH beads, head sites and linker
// beads, one wavefront per chunk.
__global__ void __launch_bounds__(256) k_chunks(Stat s, Geo g, Sites st) {
    int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool w0 = (threadIdx.x & 63) == 0;
    const int a = nchunks(s.n_tail), b = nchunks(s.n_LH), c = nchunks(s.n_lhc * s.n_lh_g);
    if (t < a) { const double4 h = chunk_sphere(g.r_t, nullptr, t, s.n_tail); if (w0) st.ch_t[t] = h; return; }
    t -= a;
    if (t < b) { const double4 h = chunk_sphere(g.r_lh, nullptr, t, s.n_LH); if (w0) st.ch_lh[t] = h; return; }
    t -= b;
    if (t < c) { const double4 h = chunk_sphere(st.G, nullptr, t, s.n_lhc * s.n_lh_g); if (w0) st.ch_g[t] = h; return; }
    t -= c;
    if (t < nchunks(s.n_lnk)) { const double4 h = chunk_sphere(g.r, s.lnk, t, s.n_lnk); if (w0) st.ch_l[t] = h; }
}

// Charge sites Q (with charges, Qp) of every core and head sites G of every
// LH-bearing core in the lab frame, a thread per site; with chunks, the
// workgroups from nsb on form the chunk spheres as k_chunks does, a head
// chunk from its sites formed here again (the same values k_sites writes).
__device__ __forceinline__ void sites_block(const Stat &s, const Geo &g, const Sites &st, int *cc_count, int bid,
                                            int nsb) {
    if (bid >= nsb) {
        int t = (bid - nsb) * 4 + (threadIdx.x >> 6);
        const bool w0 = (threadIdx.x & 63) == 0;
        const int a = nchunks(s.n_tail), b = nchunks(s.n_LH), c = nchunks(s.n_lhc * s.n_lh_g);
        if (t < a) { const double4 h = chunk_sphere(g.r_t, nullptr, t, s.n_tail); if (w0) st.ch_t[t] = h; return; }
        t -= a;
        if (t < b) { const double4 h = chunk_sphere(g.r_lh, nullptr, t, s.n_LH); if (w0) st.ch_lh[t] = h; return; }
        t -= b;
        if (t < c) {
            const int nh = s.n_lhc * s.n_lh_g, lane = threadIdx.x & 63, k = t * kChunk + lane;
            const int cnt = min(nh - t * kChunk, kChunk);
            const bool in = lane < cnt;
            const double4 h = sphere_of(in ? head_site(s, g, k) : zero3(), in, cnt);
            if (w0) st.ch_g[t] = h;
            return;
        }
        t -= c;
        if (t < nchunks(s.n_lnk)) { const double4 h = chunk_sphere(g.r, s.lnk, t, s.n_lnk); if (w0) st.ch_l[t] = h; }
        return;
    }
    int t = bid * 256 + threadIdx.x;
    if (t == 0) *cc_count = 0;  // the core-core pair list of this pass starts empty
    int nq = s.n_c * s.Nq;
    if (t < nq) {
        int k = t / s.Nq, q = t % s.Nq, cb = s.core_bead[k];
        double3 w = body_to_world(ld3(g.r + 3 * cb), ld3(g.a + 3 * cb), ld3(g.b + 3 * cb),
                                  ld3(g.c + 3 * cb), ld3(s.cpos + 3 * q));
        st3(st.Q + 3 * t, w);
        double *p = st.Qp + 4 * (size_t)t;
        p[0] = w.x; p[1] = w.y; p[2] = w.z; p[3] = s.cq[q];
        return;
    }
    t -= nq;
    if (t < s.n_lhc * s.n_lh_g) st3(st.G + 3 * t, head_site(s, g, t));
}
__global__ void __launch_bounds__(256) k_sites(Stat s, Geo g, Sites st, int *cc_count, int nsb) {
    sites_block(s, g, st, cc_count, blockIdx.x, nsb);
}
// The predicted state's frame passes and tail / LH bending angles (the
// chain segments of k_frames_seg, on the first threads of a workgroup) and,
// in the same launch, force pass 2's charge and head sites and chunk
// spheres: the sites read positions and core frames only, which the frame
// passes do not write.
__global__ void __launch_bounds__(256) k_frames_sites(Stat s, Geo g, Sites st, int *cc_count, int nbeta, int nsb) {
    const int b = blockIdx.x;
    if (b < s.nfseg) {
        const int i = s.fseg[b] + threadIdx.x;
        const bool mine = i < s.fseg[b + 1];
        if (mine) euler1_bead(s, g, i);
        __syncthreads();
        if (mine) euler2_bead(s, g, i);
        return;
    }
    if (b < s.nfseg + nbeta) {
        const int i = (b - s.nfseg) * 256 + threadIdx.x;
        if (i < s.n_tail + s.n_LH) beta_site(s, g, i, 1, 1);
        return;
    }
    sites_block(s, g, st, cc_count, b - s.nfseg - nbeta, nsb);
}

// Ordered stream compaction over an NT-thread workgroup: threads with keep
// write value to list in thread order; returns the count.  Contains
// barriers: every thread of the workgroup calls it.
template <int NT>
__device__ __forceinline__ int compact_ordered(bool keep, int value, int *list, int *wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long m = __ballot(keep);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
        const int c = wsum[i];
        off += (i < w) ? c : 0;
        tot += c;
    }
    if (keep) list[off + before] = value;
    __syncthreads();
    return tot;
}

// The partners k < n within reach of the sphere (p, rp): the chunks whose
// sphere comes within reach are compacted in index order, and their members
// are dealt out in that order, member f of the list going to the flat thread
// index f (mod stride) -- start is this thread's flat index -- so every
// thread visits a fixed sequence: fn(k).  Partners left out are beyond the
// cut-off of every term that uses this (7 nm), where the terms are exactly 0.
// Contains barriers: every thread of the workgroup calls it.
constexpr double kReach = 7.0 + 1e-6;
template <int NT, typename Fn>
__device__ __forceinline__ void for_near(const double4 *ch, int n, double3 p, double rp, int start, int stride,
                                         int *list, int *wsum, Fn &&fn) {
    const int nch = nchunks(n);
    for (int base = 0; base < nch; base += NT) {
        const int c = base + threadIdx.x;
        bool keep = false;
        if (c < nch) {
            const double4 h = ch[c];
            const double3 d = make_double3(p.x - h.x, p.y - h.y, p.z - h.z);
            keep = !(sqrt(dot(d, d)) - h.w - rp > kReach);
        }
        const int cnt = compact_ordered<NT>(keep, c, list, wsum);
        for (int f = start; f < cnt * kChunk; f += stride) {
            const int k = list[f / kChunk] * kChunk + f % kChunk;
            if (k < n) fn(k);
        }
        __syncthreads();
    }
}

// --- cell grids of the partner sets (the same pair set, fewer candidates) ---
__device__ __forceinline__ int set_size(const Stat &s, int set) {
    return set == SET_TAIL ? s.n_tail : set == SET_LH ? s.n_LH : set == SET_HEAD ? s.n_lhc * s.n_lh_g : s.n_lnk;
}
__device__ __forceinline__ double3 grid_point(const CellGrid &cg, int k) {
    return ld3(cg.pos + 3 * (cg.idx ? cg.idx[k] : k));
}
__device__ __forceinline__ int cell_of(double3 p, const double *o, double inv, int nx, int ny, int nz) {
    const int ix = min(max((int)floor((p.x - o[0]) * inv), 0), nx - 1);
    const int iy = min(max((int)floor((p.y - o[1]) * inv), 0), ny - 1);
    const int iz = min(max((int)floor((p.z - o[2]) * inv), 0), nz - 1);
    return (iz * ny + iy) * nx + ix;
}

// Exclusive scan of v over an NT-thread workgroup; returns this thread's
// offset and the total in *tot.  Contains barriers.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int *tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int off = 0, t = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
        off += (i < w) ? wsum[i] : 0;
        t += wsum[i];
    }
    *tot = t;
    __syncthreads();
    return off + x - v;
}

// One workgroup per partner set: bounding box, grid, counts, offsets and
// the fill (atomic cursors; k_cell_sort orders each cell afterwards).
constexpr int kCellThreads = 1024;
__global__ void __launch_bounds__(kCellThreads) k_cells(Sites st) {
    __shared__ double red[6 * (kCellThreads / 64)];
    __shared__ double s_hdr[4];
    __shared__ int s_dim[3], s_wsum[kCellThreads / 64];
    const int set = blockIdx.x, tid = threadIdx.x;
    const CellGrid cg = st.cg[set];
    const int n = cg.n;
    if (n == 0) {
        if (tid < 3) cg.dim[tid] = 0;
        return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = tid; k < n; k += kCellThreads) {
        const double3 p = grid_point(cg, k);
        lo[0] = fmin(lo[0], p.x); lo[1] = fmin(lo[1], p.y); lo[2] = fmin(lo[2], p.z);
        hi[0] = fmax(hi[0], p.x); hi[1] = fmax(hi[1], p.y); hi[2] = fmax(hi[2], p.z);
    }
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    const int lane = tid & 63, w = tid >> 6;
    if (lane == 0)
        for (int d = 0; d < 3; d++) { red[6 * w + d] = lo[d]; red[6 * w + 3 + d] = hi[d]; }
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < kCellThreads / 64; i++)
            for (int d = 0; d < 3; d++) {
                red[d] = fmin(red[d], red[6 * i + d]);
                red[3 + d] = fmax(red[3 + d], red[6 * i + 3 + d]);
            }
        // a box that is not finite (a trajectory that blew up) takes one cell
        double ext[3];
        for (int d = 0; d < 3; d++) {
            ext[d] = red[3 + d] - red[d];
            if (!(ext[d] >= 0.0 && ext[d] < 1e12)) { ext[d] = 0.0; red[d] = 0.0; }
        }
        double edge = kReach;
        int nd[3] = {1, 1, 1};
        for (int it = 0; it < 200; it++) {
            long long tot = 1;
            for (int d = 0; d < 3; d++) {
                nd[d] = (int)floor(ext[d] / edge) + 1;
                tot *= nd[d];
            }
            if (tot <= kCellMax) break;
            edge *= 1.25;
        }
        for (int d = 0; d < 3; d++) { s_hdr[d] = red[d]; s_dim[d] = nd[d]; cg.hdr[d] = red[d]; cg.dim[d] = nd[d]; }
        s_hdr[3] = 1.0 / edge;
        cg.hdr[3] = s_hdr[3];
    }
    __syncthreads();
    const int nx = s_dim[0], ny = s_dim[1], nz = s_dim[2], nc = nx * ny * nz;
    const double inv = s_hdr[3];
    // counts, offsets and cursors: plain stores complete (the barrier waits
    // for them) before the atomics of the next phase act on the same words,
    // and the counts are read back through atomics, where they were formed
    for (int c = tid; c <= nc; c += kCellThreads) cg.start[c] = 0;
    __syncthreads();
    for (int k = tid; k < n; k += kCellThreads) {
        const int c = cell_of(grid_point(cg, k), s_hdr, inv, nx, ny, nz);
        cg.cid[k] = c;
        atomicAdd(cg.start + c, 1);
    }
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < nc; base += kCellThreads) {
        const int c = base + tid;
        const int v = c < nc ? atomicAdd(cg.start + c, 0) : 0;
        int tot;
        const int ex = block_excl_scan<kCellThreads>(v, s_wsum, &tot);
        if (c < nc) {
            cg.start[c] = carry + ex;
            cg.cursor[c] = carry + ex;
        }
        carry += tot;
    }
    if (tid == 0) cg.start[nc] = carry;
    __syncthreads();
    for (int k = tid; k < n; k += kCellThreads) {
        const int slot = atomicAdd(cg.cursor + cg.cid[k], 1);
        cg.members[slot] = k;
    }
}

// Members of every cell in increasing index order, one wavefront per cell
// (rank of each member = number of smaller members); a cell of more than
// 256 members is ordered by one lane (insertion sort).
__global__ void __launch_bounds__(256) k_cell_sort(Stat s, Sites st) {
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + (threadIdx.x >> 6), tw = gridDim.x * 4;
    for (int set = 0; set < NSETS; set++) {
        if (set_size(s, set) == 0) continue;
        const CellGrid cg = st.cg[set];
        const int nc = cg.dim[0] * cg.dim[1] * cg.dim[2];
        for (int c = gw; c < nc; c += tw) {
            const int b = cg.start[c], m = cg.start[c + 1] - b;
            if (m <= 1) continue;
            int *mem = cg.members + b;
            if (m <= 256) {
                int x[4], r[4] = {0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i < 4; i++) x[i] = (lane + 64 * i < m) ? mem[lane + 64 * i] : 0;
                for (int j = 0; j < m; j++) {
                    const int y = mem[j];
#pragma unroll
                    for (int i = 0; i < 4; i++) r[i] += (y < x[i]) ? 1 : 0;
                }
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (lane + 64 * i < m) mem[r[i]] = x[i];
            } else if (lane == 0) {
                for (int i = 1; i < m; i++) {
                    const int v = mem[i];
                    int j = i - 1;
                    while (j >= 0 && mem[j] > v) {
                        mem[j + 1] = mem[j];
                        j--;
                    }
                    mem[j + 1] = v;
                }
            }
        }
    }
}

// The partners within reach of the sphere (p, rp) through a cell grid: the
// runs of cells along x that meet the box [p - rp - reach, p + rp + reach],
// rows in (z, y) order, each run a contiguous range of members, dealt out
// in that order, member f of the concatenation going to the flat thread
// index f (mod stride) -- start is this thread's flat index.  Partners left
// out are beyond the 7 nm cut-off of every term that uses this, where the
// terms are exactly 0.  Contains barriers: every thread of the workgroup
// calls it with the same sphere.
template <int NT, typename Fn>
__device__ __forceinline__ void for_near_cells(const CellGrid &cg, double3 p, double rp, int start, int stride,
                                               int *off, int *wsum, Fn &&fn) {
    __shared__ int s_beg[NT];
    const int nx = cg.dim[0], ny = cg.dim[1], nz = cg.dim[2];
    if (nx == 0) return;
    const double inv = cg.hdr[3], reach = rp + kReach;
    const double pc[3] = {p.x, p.y, p.z};
    const int nd[3] = {nx, ny, nz};
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lo[d] = max((int)floor((pc[d] - reach - cg.hdr[d]) * inv), 0);
        hi[d] = min((int)floor((pc[d] + reach - cg.hdr[d]) * inv), nd[d] - 1);
    }
    if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;
    const int nyr = hi[1] - lo[1] + 1, nruns = nyr * (hi[2] - lo[2] + 1);
    for (int rb = 0; rb < nruns; rb += NT) {
        const int r = rb + threadIdx.x;
        int beg = 0, len = 0;
        if (r < nruns) {
            const int c0 = ((lo[2] + r / nyr) * ny + lo[1] + r % nyr) * nx + lo[0];
            beg = cg.start[c0];
            len = cg.start[c0 + hi[0] - lo[0] + 1] - beg;
        }
        int tot;
        const int ex = block_excl_scan<NT>(len, wsum, &tot);
        off[threadIdx.x] = ex;
        s_beg[threadIdx.x] = beg;
        __syncthreads();
        const int nr = min(NT, nruns - rb);
        for (int f = start; f < tot; f += stride) {
            int a = 0, z = nr - 1;  // last run with off <= f
            while (a < z) {
                const int mid = (a + z + 1) >> 1;
                if (off[mid] <= f) a = mid;
                else z = mid - 1;
            }
            fn(cg.members[s_beg[a] + f - off[a]]);
        }
        __syncthreads();
    }
}

// The partners of set `set` within reach of (p, rp): through the cell
// grid, or through the chunk spheres ch.
template <int NT, typename Fn>
__device__ __forceinline__ void for_near_set(const Sites &st, int set, const double4 *ch, int n, double3 p, double rp,
                                             int start, int stride, int *list, int *wsum, Fn &&fn) {
    if (st.use_cells) for_near_cells<NT>(st.cg[set], p, rp, start, stride, list, wsum, fn);
    else for_near<NT>(ch, n, p, rp, start, stride, list, wsum, fn);
}

// Sphere around a set of points (centroid, largest distance from it).
template <int N>
__device__ __forceinline__ void points_sphere(const double3 (&x)[N], int n, double3 &c, double &r) {
    c = zero3();
#pragma unroll
    for (int i = 0; i < N; i++)
        if (i < n) c += x[i];
    const double w = 1.0 / n;
    c = make_double3(c.x * w, c.y * w, c.z * w);
    double r2 = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++)
        if (i < n) {
            const double3 d = x[i] - c;
            r2 = fmax(r2, dot(d, d));
        }
    r = sqrt(r2);
}

__global__ void k_mech(Stat s, Geo g, Frc f, double *Epart, int terms) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= s.n) return;
    double3 F = zero3(), T = zero3();
    double E = 0.0;
    const int n = s.n;
    if (terms & TERM_MECH) {
        double3 rj = ld3(g.r + 3 * j), aj = ld3(g.a + 3 * j), bj = ld3(g.b + 3 * j), cj = ld3(g.c + 3 * j);
        SegOut L, R;  // segment j-1 (bead j on its right) and segment j
        bool hasL = j >= 1, hasR = j <= n - 2;
        if (hasL) segment(s, g, j - 1, L);
        if (hasR) segment(s, g, j, R);
        // reference order: all stretches, then bends, then twists
        if (hasL) { F += L.s2; T += torque_abs(L.s2, L.rf, rj, aj, bj, cj); }
        if (hasR) { F += R.s1; T += torque_abs(R.s1, R.rs, rj, aj, bj, cj); }
        if (hasL) { F += L.b2; T += torque_abs(L.b2, L.rf, rj, aj, bj, cj); }
        if (hasR) { F += R.b1; T += torque_abs(R.b1, R.rs, rj, aj, bj, cj); }
        if (hasL) { F += L.t2; T += torque_abs(L.t2, L.rf, rj, aj, bj, cj); }
        if (hasR) { F += R.t1; T += torque_abs(R.t1, R.rs, rj, aj, bj, cj); E = R.E; }
        const double si = sin(kTHETA), co = cos(kTHETA);
        const double *al = g.alpha, *ga = g.gamma, *ph = s.phi_o;
        if (j == n - 1) {
            // last bead: assigned (func.cpp:1281-1287)
            T = make_double3(-s.s * (al[n - 2] + ga[n - 2] - ph[n - 2]), 0.0, 0.0);
        } else if (s.type[j] == 0) {
            // linker bead: assigned (func.cpp:1206-1209)
            T = make_double3(s.s * (al[j] + ga[j] + ph[j] - al[j - 1] - ga[j - 1] - ph[j - 1]), 0.0, 0.0);
        } else {
            // core: explicit torques (func.cpp:1210-1273)
            int ic = s.ic_of[j];
            double3 ad = ld3(g.a_dna + 3 * ic), bd = ld3(g.b_dna + 3 * ic), cd = ld3(g.c_dna + 3 * ic);
            double ada = dot(ad, aj), adb = dot(ad, bj), adc = dot(ad, cj);
            double mag = s.s * (al[j] + ga[j] - ph[j]);
            T.x += mag * ada; T.y += mag * adb; T.z += mag * adc;
            if (j > 0) {
                mag = -s.s * (al[j - 1] + ga[j - 1] - ph[j - 1]);
                T.x += mag * co; T.y += mag * si;
            }
            double bp = g.beta_p[ic], sbp, cbp;
            sincos_small(bp, sbp, cbp);
            T.y = T.y - s.g * bp * adc / sbp;
            T.z = T.z + s.g * bp * adb / sbp;
            if (j > 0) {
                double3 ab = ld3(g.a + 3 * j - 3);
                double xa = dot(ab, aj), xb = dot(ab, bj), xc = dot(ab, cj);
                double bm = g.beta[j - 1], sbm, cbm;
                sincos_small(bm, sbm, cbm);
                T.x = T.x + s.g * bm * (si * xc) / sbm;
                T.y = T.y - s.g * bm * (co * xc) / sbm;
                T.z = T.z + s.g * bm * (co * xb - si * xa) / sbm;
            }
            double s1, c1;
            sincos_small(g.alpha_p[ic], s1, c1);
            mag = s.s * (al[j] + ga[j] - ph[j]) * tan_small(0.5 * bp);
            T.x = T.x - mag * (s1 * dot(cd, aj) + c1 * dot(bd, aj));
            T.y = T.y - mag * (s1 * dot(cd, bj) + c1 * dot(bd, bj));
            T.z = T.z - mag * (s1 * dot(cd, cj) + c1 * dot(bd, cj));
            if (j > 0) {
                sincos_small(ga[j - 1], s1, c1);
                mag = s.s * (al[j - 1] + ga[j - 1] - ph[j - 1]) * tan_small(0.5 * g.beta[j - 1]);
                T.x = T.x - mag * (s1 * 0.0 - c1 * (-si));
                T.y = T.y - mag * (s1 * 0.0 - c1 * co);
                T.z = T.z - mag * (s1 * 1.0 - c1 * 0.0);
            }
        }
    }
    st3(f.F + 3 * j, F);
    st3(f.T + 3 * j, T);
    if (Epart) Epart[j] = E;  // segment j's energy
}

// ---------------------------------------------------------------------------
// core / linker nonbonded helpers (func.cpp:1291-1605, no clamps, EV <= 8 nm)
// ---------------------------------------------------------------------------
// scalar g1 of the reference for a charge pair at distance d (id = 1/d);
// adds energy
template <typename EC>
__device__ __forceinline__ double dh_core(const EC &ec, double ke_q1q2, double kappa, double d, double id,
                                          double &E) {
    const double qe = ke_q1q2 * exp_neg_c(-kappa * d, ec);
    E += qe * id;
    return qe * (kappa * d + 1.0) * (id * id);
}
__device__ __forceinline__ double ev_core(double kex, double sg, double id, double &E) {
    double x = sg * id, x2 = x * x, x6 = x2 * x2 * x2, x7 = x6 * x, x13 = x6 * x7;
    E += kex * (x6 * x6 - x6);
    return kex * ((12.0 / sg) * x13 - (6.0 / sg) * x7);
}

// ---------------------------------------------------------------------------
// Receiver-side kernels for linker, tail and LH beads: one workgroup per
// receiving bead, its threads split the partner lists, and a fixed-order
// block reduction forms the bead's force (deterministic, no atomics).
// ---------------------------------------------------------------------------
// Workgroup sizes: a tail bead has few partners within reach and many
// receivers run side by side (one wavefront each); linker beads carry the
// most partners (DNA and core charges).
constexpr int kTailThreads = 64, kLHThreads = 128, kLinkerThreads = 256;

// Linker bead (func.cpp:1304-1487 linker side, 1898-1917, 2249-2279,
// 2481-2522): DNA-DNA, DNA-core charges, tail-linker, LH-DNA, LHG-DNA.
// F[j] +=.  Energy owner: DNA-DNA pairs with partner index > j.
__global__ void __launch_bounds__(kLinkerThreads) k_linker(Stat s, Geo g, Sites st, Frc f, double *Epart, int terms) {
    __shared__ double red[4 * (kLinkerThreads / 64)];
    __shared__ int s_list[kLinkerThreads], s_wsum[kLinkerThreads / 64];
    const int li = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const int j = s.lnk[li];
    const double3 rj = ld3(g.r + 3 * j);
    double3 F = zero3();
    double E = 0.0, Ed = 0.0;
    if ((terms & TERM_ELEC) && s.elig[j]) {
        const double kll = s.debyell, qlql = s.q_l * s.q_l;
        for (int pi = tid; pi < s.n_lnk; pi += nth) {
            int i = s.lnk[pi];
            if (!s.elig[i] || abs(i - j) <= 1) continue;
            double3 dv = rj - ld3(g.r + 3 * i);
            const double d2 = dot(dv, dv);
            if (d2 > kD25Sq) continue;  // dist > 25
            double Ep = 0.0;
            const double id = rsqrt_f64(d2), dist = d2 * id, qe = s.k_e * qlql * exp_neg_c(-kll * dist, s.ec);
            double g1 = qe * (kll * dist + 1.0) * (id * id);
            Ep += qe * id;
            if (d2 <= kD8Sq) g1 += ev_core(s.k_ex, s.sig_dd, id, Ep);  // dist <= 8
            F += (g1 * id) * dv;
            if (i > j) E += Ep;
        }
        // DNA-core: every charge of the eligible cores within 25 nm, the
        // cores tested by the whole workgroup and listed in index order;
        // energy on the core side
        for (int base = 0; base < s.n_c; base += nth) {
            const int k = base + tid;
            bool keep = false;
            if (k < s.n_c) {
                const int cb = s.core_bead[k];
                if (s.elig[cb]) {
                    const double3 dc = ld3(g.r + 3 * cb) - rj;
                    keep = !(sqrt(dot(dc, dc)) > 25.0);
                }
            }
            const int cnt = compact_ordered<kLinkerThreads>(keep, k, s_list, s_wsum);
            for (int e = 0; e < cnt; e++) {
                const int kc = s_list[e];
                const bool dh = abs(s.core_bead[kc] - j) > 1;
                const double *Q = st.Q + 3 * (size_t)kc * s.Nq;
                for (int q = tid; q < s.Nq; q += nth) {
                    double3 z = rj - ld3(Q + 3 * q);
                    const double d2 = dot(z, z), id = rsqrt_f64(d2), d = d2 * id;
                    double cq = s.cq[q];
                    double g1 = 0.0;
                    if (dh) g1 += dh_core(s.ec, s.k_e * s.q_l * cq, s.debye, d, id, Ed);
                    if (d2 <= kD8Sq && cq > 0) g1 += ev_core(s.k_ex, s.sig_dc, id, Ed);  // d <= 8
                    F += (g1 * id) * z;
                }
            }
            __syncthreads();
        }
    }
    if (terms & TERM_TAIL)  // tail-linker, energy owned by tails
        for_near_set<kLinkerThreads>(st, SET_TAIL, st.ch_t, s.n_tail, rj, 0.0, tid, nth, s_list, s_wsum, [&](int t) {
            dhev_recv(s.ec, s.t_q[t] * s.q_l, s.c4pe, s.debye, s.k_ex, s.sig_tl, rj, ld3(g.r_t + 3 * t), F, Ed);
        });
    if (terms & TERM_LH) {
        // LH-DNA, energy owned by LH beads
        for_near_set<kLinkerThreads>(st, SET_LH, st.ch_lh, s.n_LH, rj, 0.0, tid, nth, s_list, s_wsum, [&](int i) {
            int idx = i % s.nlh;
            idx = (idx < s.n_lh_n) ? idx : idx + s.n_lh_g;
            dhev_recv(s.ec, s.LH_q[idx] * s.q_l, s.c4pe, s.debye, s.k_ex, s.LH_hl[idx], rj, ld3(g.r_lh + 3 * i), F, Ed);
        });
        // LHG-DNA, energy on the core side
        for_near_set<kLinkerThreads>(st, SET_HEAD, st.ch_g, s.n_lhc * s.n_lh_g, rj, 0.0, tid, nth, s_list, s_wsum, [&](int x) {
            int m = x % s.n_lh_g;
            dhev_recv(s.ec, s.LH_q[m] * s.q_l, s.c4pe, s.debye, s.k_ex, s.LH_hl[m], rj, ld3(st.G + 3 * x), F, Ed);
        });
    }
    double v[4] = {F.x, F.y, F.z, E};
    block_sum_n<kLinkerThreads, 4>(v, red);
    if (tid == 0) {
        double *Fj = f.F + 3 * j;
        Fj[0] += v[0]; Fj[1] += v[1]; Fj[2] += v[2];
        if (Epart) Epart[li] = v[3];
    }
}

// Tail bead (func.cpp:1793-2026 and the tail side of 2343-2476).  Assigns
// tF[j].  Energy owner: bonds / angles starting at j, tail-tail pairs with
// partner > j, tail-linker pairs, the anchor spring.
__global__ void __launch_bounds__(kTailThreads) k_tail(Stat s, Geo g, Sites st, Frc f, double *Epart, int terms) {
    __shared__ double red[4 * (kTailThreads / 64)];
    __shared__ int s_list[kTailThreads], s_wsum[kTailThreads / 64];
    const int j = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const int nt = s.n_tail;
    const int *tg = s.t_grp;
    const double3 rj = ld3(g.r_t + 3 * j);
    double3 F = zero3();
    double E = 0.0, Ed = 0.0;
    if (terms & TERM_TAIL) {
        if (tid == 0) {
            // stretching of bonds (j-1,j) and (j,j+1) (1811-1827)
            if (j >= 1 && tg[j - 1] == tg[j])
                F -= stretch_f1(s.h_t[j - 1], ld3(g.r_t + 3 * j - 3), rj, s.lo_t[j - 1], Ed);
            if (j + 1 < nt && tg[j] == tg[j + 1]) F += stretch_f1(s.h_t[j], rj, ld3(g.r_t + 3 * j + 3), s.lo_t[j], E);
            // bending of triples starting at j-2, j-1, j (1832-1858), subtracted
            for (int c0 = j - 2; c0 <= j; c0++) {
                if (c0 < 0 || c0 + 2 >= nt) continue;
                if (!(tg[c0] == tg[c0 + 1] && tg[c0] == tg[c0 + 2])) continue;
                double3 f1, f2, f3;
                bend3(s.g_t[c0], g.beta_t[c0], s.beta_o_t[c0], ld3(g.r_t + 3 * c0), ld3(g.r_t + 3 * c0 + 3),
                      ld3(g.r_t + 3 * c0 + 6), f1, f2, f3, (c0 == j) ? E : Ed);
                F -= (c0 == j) ? f1 : (c0 == j - 1 ? f2 : f3);
            }
            // anchor spring to the owning core (1971-2022)
            if (s.t_fix[j] == 1) {
                int cb = s.core_bead[j / s.n_t];
                double3 ro = body_to_world(ld3(g.r + 3 * cb), ld3(g.a + 3 * cb), ld3(g.b + 3 * cb),
                                           ld3(g.c + 3 * cb), ld3(s.tail_pos + 3 * (j % s.n_t)));
                double3 S = rj - ro;
                E += s.h * (S.x * S.x + S.y * S.y + S.z * S.z) / 2;
                F -= make_double3(s.h * S.x, s.h * S.y, s.h * S.z);
            }
        }
        // tail-tail DH (1863-1891), skipping same-group neighbours
        const double qj = s.t_q[j];
        for_near_set<kTailThreads>(st, SET_TAIL, st.ch_t, nt, rj, 0.0, tid, nth, s_list, s_wsum, [&](int k) {
            if (k == j) return;
            if (!(tg[k] != tg[j] || abs(k - j) > 2)) return;
            dh_recv(s.ec, qj * s.t_q[k], s.c4pe, s.debye, rj, ld3(g.r_t + 3 * k), F, (k > j) ? E : Ed);
        });
        // tail-linker (1898-1917)
        for_near_set<kTailThreads>(st, SET_LINKER, st.ch_l, s.n_lnk, rj, 0.0, tid, nth, s_list, s_wsum, [&](int pi) {
            dhev_recv(s.ec, qj * s.q_l, s.c4pe, s.debye, s.k_ex, s.sig_tl, rj, ld3(g.r + 3 * s.lnk[pi]), F, E);
        });
    }
    if ((terms & TERM_LH) && s.t_fix[j] == 0) {
        // LH-tail and LHG-tail, tail side; anchored tails pass these to a core (k_core)
        for_near_set<kTailThreads>(st, SET_LH, st.ch_lh, s.n_LH, rj, 0.0, tid, nth, s_list, s_wsum, [&](int i) {
            int idx = i % s.nlh;
            idx = (idx < s.n_lh_n) ? idx : idx + s.n_lh_g;
            dhev_recv(s.ec, s.LH_q[idx] * s.t_q[j], s.c4pe, s.debye, s.k_ex, s.LH_ht[idx], rj, ld3(g.r_lh + 3 * i), F, Ed);
        });
        for_near_set<kTailThreads>(st, SET_HEAD, st.ch_g, s.n_lhc * s.n_lh_g, rj, 0.0, tid, nth, s_list, s_wsum, [&](int x) {
            int m = x % s.n_lh_g;
            dhev_recv(s.ec, s.LH_q[m] * s.t_q[j], s.c4pe, s.debye, s.k_ex, s.LH_ht[m], rj, ld3(st.G + 3 * x), F, Ed);
        });
    }
    double v[4] = {F.x, F.y, F.z, E};
    block_sum_n<kTailThreads, 4>(v, red);
    if (tid == 0) {
        st3(f.tF + 3 * j, make_double3(v[0], v[1], v[2]));
        if (Epart) Epart[j] = v[3];
    }
}

// Free LH bead (func.cpp:2093-2781, LH side).  Assigns lF[j].  Energy
// owner: bonds / angles at j, LH-LH pairs with partner > j, LH-DNA, LH-tail.
__global__ void __launch_bounds__(kLHThreads) k_lh(Stat s, Geo g, Sites st, Frc f, double *Epart, int terms) {
    __shared__ double red[4 * (kLHThreads / 64)];
    __shared__ int s_list[kLHThreads], s_wsum[kLHThreads / 64];
    const int j = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const int nlh = s.nlh;
    auto tpl = [&](int i) { int x = i % nlh; return (x < s.n_lh_n) ? x : x + s.n_lh_g; };
    const int idx = tpl(j);
    const int mine = j / nlh;
    const double3 rj = ld3(g.r_lh + 3 * j);
    double3 F = zero3();
    double E = 0.0, Ed = 0.0;
    if (terms & TERM_LH) {
        if (tid == 0) {
            // stretching (2111-2143)
            if (j >= 1 && s.LH_conn[tpl(j - 1)] == 1) {
                int ip = tpl(j - 1);
                F -= stretch_f1(s.LH_kstr[ip], ld3(g.r_lh + 3 * j - 3), rj, s.LH_streq[ip], Ed);
            }
            if (j + 1 < s.n_LH && s.LH_conn[idx] == 1)
                F += stretch_f1(s.LH_kstr[idx], rj, ld3(g.r_lh + 3 * j + 3), s.LH_streq[idx], E);
            // bending centred at j-1, j, j+1 (2149-2181)
            for (int c0 = j - 1; c0 <= j + 1; c0++) {
                if (c0 < 1 || c0 + 1 >= s.n_LH) continue;
                if (!((c0 % s.n_lh_c != 0) && (c0 % s.n_lh_c != s.n_lh_c - 1))) continue;
                int ic = tpl(c0);
                double3 f1, f2, f3;
                bend3(s.LH_kben[ic], g.beta_lh[c0], s.LH_betaeq[ic], ld3(g.r_lh + 3 * c0 - 3),
                      ld3(g.r_lh + 3 * c0), ld3(g.r_lh + 3 * c0 + 3), f1, f2, f3, (c0 == j) ? E : Ed);
                F -= (c0 == j + 1) ? f1 : (c0 == j ? f2 : f3);
            }
            // globular head - first C-term bead spring: force x10 (2727-2777)
            if (j % nlh == 0) {
                for (int m = 0; m < s.n_lh_g; m++) {
                    int ti = m + s.n_lh_n;
                    if (s.LH_conn[ti] != 1) continue;
                    double3 G = ld3(st.G + 3 * (mine * s.n_lh_g + m));
                    double3 dv = rj - G;
                    double d = sqrt(dot(dv, dv));
                    double ks = s.LH_kstr[ti], l0 = s.LH_streq[ti];
                    double3 S = make_double3((d - l0) * dv.x / d, (d - l0) * dv.y / d, (d - l0) * dv.z / d);
                    F -= make_double3(ks * S.x * 10, ks * S.y * 10, ks * S.z * 10);
                }
            }
        }
        // LH-LH (2190-2244)
        for_near_set<kLHThreads>(st, SET_LH, st.ch_lh, s.n_LH, rj, 0.0, tid, nth, s_list, s_wsum, [&](int k) {
            if (k == j) return;
            int i2 = tpl(k);
            double sg = (s.LH_hh[idx] + s.LH_hh[i2]) / 2;
            double &Ek = (k > j) ? E : Ed;
            if (k / nlh == mine) {
                if (abs(k - j) > 1) ev_recv(s.k_ex, sg, rj, ld3(g.r_lh + 3 * k), F, Ek);
            } else {
                dhev_recv(s.ec, s.LH_q[idx] * s.LH_q[i2], s.c4pe, s.debye, s.k_ex, sg, rj, ld3(g.r_lh + 3 * k), F, Ek);
            }
        });
        // LH-DNA (2249-2279)
        for_near_set<kLHThreads>(st, SET_LINKER, st.ch_l, s.n_lnk, rj, 0.0, tid, nth, s_list, s_wsum, [&](int pi) {
            dhev_recv(s.ec, s.LH_q[idx] * s.q_l, s.c4pe, s.debye, s.k_ex, s.LH_hl[idx], rj, ld3(g.r + 3 * s.lnk[pi]), F, E);
        });
        // LH-tail (2343-2398)
        for_near_set<kLHThreads>(st, SET_TAIL, st.ch_t, s.n_tail, rj, 0.0, tid, nth, s_list, s_wsum, [&](int t) {
            dhev_recv(s.ec, s.LH_q[idx] * s.t_q[t], s.c4pe, s.debye, s.k_ex, s.LH_ht[idx], rj, ld3(g.r_t + 3 * t), F, E);
        });
        // LHG of other cores acting on this bead (2590-2644), energy on the core side
        for_near_set<kLHThreads>(st, SET_HEAD, st.ch_g, s.n_lhc * s.n_lh_g, rj, 0.0, tid, nth, s_list, s_wsum, [&](int x) {
            int k = x / s.n_lh_g, m = x % s.n_lh_g;
            if (k == mine) return;
            dhev_recv(s.ec, s.LH_q[m] * s.LH_q[idx], s.c4pe, s.debye, s.k_ex, (s.LH_hh[m] + s.LH_hh[idx]) / 2, rj,
                      ld3(st.G + 3 * x), F, Ed);
        });
    }
    double v[4] = {F.x, F.y, F.z, E};
    block_sum_n<kLHThreads, 4>(v, red);
    if (tid == 0) {
        st3(f.lF + 3 * j, make_double3(v[0], v[1], v[2]));
        if (Epart) Epart[j] = v[3];
    }
}

// ---------------------------------------------------------------------------
// Workgroup per core (ordinal kc): everything acting on the core body.
// F[cb] +=, T[cb] +=.  It runs after k_cc_work and also gathers this core's
// row (as c1) and column (as c2) of the core-core slot matrix.
// Energy owners: core-linker charge pairs, LHG-* pairs of this core's head
// (LHG-LHG only towards higher LH ordinals), the LHG-C spring (slot
// Ecore[kc]); core-core pairs with this core as the lower index (Ecc[kc]).
//
// Partner lists are culled by the whole workgroup, 320 candidates at a
// time, into LDS in index order (ballot + prefix count), so the per-charge
// loops visit the surviving partners in the reference's order from LDS.
// ---------------------------------------------------------------------------
constexpr int kCoreThreads = 320;
constexpr int kCoreQPer = 2;   // charges per thread: Nq <= 640
constexpr int kMaxRT = 8;      // redirected tails summed per pass over the LH partners

// The gridDim.y = S0 + S1 + S2 workgroups of core kc take three phases:
// y < S0 the charge-owned terms, then S1 workgroups the core's own
// globular-head sites, then S2 the redirected tails and anchor springs.
// Workgroup sub of the S of its phase takes the partners e = sub (mod S)
// of the charge-owned terms, the other cores k = sub (mod S) of the
// head-charge terms, and every S-th element of the flat loops.  The last
// workgroup of the core to finish sums the partial results in order of y
// (atomic arrival count, no waiting).
__global__ void __launch_bounds__(kCoreThreads) k_core(Stat s, Geo g, Sites st, Frc f, double *core_part,
                                                      int *core_cnt, double *Ecore, int terms, int S0, int S1) {
    __shared__ int s_list[kCoreThreads];
    __shared__ double s_pos[3 * kCoreThreads];
    __shared__ int s_wsum[kCoreThreads / 64];
    __shared__ double red[3 * kMaxRT * (kCoreThreads / 64)];
    const int kc = blockIdx.x;
    const int cb = s.core_bead[kc];
    const int tid = threadIdx.x;
    const int NS = gridDim.y, y = blockIdx.y;
    const int phase = y < S0 ? 0 : (y < S0 + S1 ? 1 : 2);
    const int S = phase == 0 ? S0 : (phase == 1 ? S1 : NS - S0 - S1);
    const int sub = phase == 0 ? y : (phase == 1 ? y - S0 : y - S0 - S1);
    const int ft = sub * kCoreThreads + tid, FT = S * kCoreThreads;
    const double3 rc = ld3(g.r + 3 * cb), ac = ld3(g.a + 3 * cb), bc = ld3(g.b + 3 * cb), cc = ld3(g.c + 3 * cb);
    double3 F = zero3(), T = zero3();
    double E = 0.0, Ed = 0.0, Ec = 0.0;  // Ec: kept for the layout of the partial sums
    const double *Qc = st.Q + 3 * (size_t)kc * s.Nq;
    const int lho = s.lh_of_core[kc];

    // --- charge-owned terms: thread per charge q of this core ---
    if (phase == 0) {
        double3 Fq[kCoreQPer];
#pragma unroll
        for (int i = 0; i < kCoreQPer; i++) Fq[i] = zero3();
        if ((terms & TERM_ELEC) && s.elig[cb]) {
            // core-linker (func.cpp:1356-1487): force on the charge
            for (int base = 0; base < s.n_lnk; base += kCoreThreads) {
                const int pi = base + tid;
                bool keep = false;
                int L = 0;
                if (pi < s.n_lnk) {
                    L = s.lnk[pi];
                    if (s.elig[L]) {
                        const double3 dc = rc - ld3(g.r + 3 * L);
                        keep = !(sqrt(dot(dc, dc)) > 25.0);
                    }
                }
                const int cnt = compact_ordered<kCoreThreads>(keep, L, s_list, s_wsum);
                for (int e = tid; e < cnt; e += kCoreThreads) st3(s_pos + 3 * e, ld3(g.r + 3 * s_list[e]));
                __syncthreads();
#pragma unroll
                for (int i = 0; i < kCoreQPer; i++) {
                    const int q = tid + i * kCoreThreads;
                    if (q < s.Nq) {
                        const double3 Qq = ld3(Qc + 3 * q);
                        const double cq = s.cq[q];
                        for (int e = sub; e < cnt; e += S) {
                            const double3 z = Qq - ld3(s_pos + 3 * e);
                            const double d2 = dot(z, z), id = rsqrt_f64(d2), d = d2 * id;
                            double g1 = 0.0;
                            if (abs(cb - s_list[e]) > 1) g1 += dh_core(s.ec, s.k_e * s.q_l * cq, s.debye, d, id, E);
                            if (d2 <= kD8Sq && cq > 0) g1 += ev_core(s.k_ex, s.sig_dc, id, E);  // d <= 8
                            Fq[i] += (g1 * id) * z;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (terms & TERM_LH) {
            // LHG of other cores acting on this charge (func.cpp:2526-2586, charge side)
            for (int base = 0; base < s.n_lhc; base += kCoreThreads) {
                const int k = base + tid;
                bool keep = false;
                if (k < s.n_lhc) {
                    const int ob = s.lhc_bead[k];
                    if (ob != cb) {
                        const double3 dc = ld3(g.r + 3 * ob) - rc;
                        keep = !(sqrt(dot(dc, dc)) - s.rmax_g - s.rmax_q > 7.0 + 1e-6);
                    }
                }
                const int cnt = compact_ordered<kCoreThreads>(keep, k, s_list, s_wsum);
#pragma unroll
                for (int i = 0; i < kCoreQPer; i++) {
                    const int q = tid + i * kCoreThreads;
                    if (q < s.Nq) {
                        const double3 Qq = ld3(Qc + 3 * q);
                        const double cq = s.cq[q];
                        for (int e = sub; e < cnt; e += S) {
                            const int k2 = s_list[e];
                            for (int m = 0; m < s.n_lh_g; m++)
                                dhev_recv(s.ec, s.LH_q[m] * cq, s.c4pe, s.debye, s.k_ex, s.LH_hc[m], Qq,
                                          ld3(st.G + 3 * (k2 * s.n_lh_g + m)), Fq[i], Ed);
                        }
                    }
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int i = 0; i < kCoreQPer; i++) {
            const int q = tid + i * kCoreThreads;
            if (q < s.Nq) {
                F += Fq[i];
                T += torque_rel(Fq[i], ld3(s.cpos + 3 * q), ac, bc, cc);
            }
        }
    }

    // --- own globular-head sites (LH-bound cores) ---
    if (phase == 1 && (terms & TERM_LH) && lho >= 0) {
        const int ng = s.n_lh_g;
        double3 Gm[kMaxGSites], Fm[kMaxGSites];
        #pragma unroll
        for (int m = 0; m < kMaxGSites; m++) if (m < ng) { Gm[m] = ld3(st.G + 3 * (lho * ng + m)); Fm[m] = zero3(); }
        // sphere around the head sites: partners beyond it by the cut-off
        // are left out of the loops below
        double3 hc;
        double hr;
        points_sphere(Gm, ng, hc, hr);
        // LHG-tail (2403-2476)
        for_near_set<kCoreThreads>(st, SET_TAIL, st.ch_t, s.n_tail, hc, hr, ft, FT, s_list, s_wsum, [&](int t) {
            double3 rt = ld3(g.r_t + 3 * t);
#pragma unroll
            for (int m = 0; m < kMaxGSites; m++) if (m < ng)
                dhev_recv(s.ec, s.LH_q[m] * s.t_q[t], s.c4pe, s.debye, s.k_ex, s.LH_ht[m], Gm[m], rt, Fm[m], E);
        });
        // LHG-DNA (2481-2522)
        for_near_set<kCoreThreads>(st, SET_LINKER, st.ch_l, s.n_lnk, hc, hr, ft, FT, s_list, s_wsum, [&](int pi) {
            double3 rl = ld3(g.r + 3 * s.lnk[pi]);
#pragma unroll
            for (int m = 0; m < kMaxGSites; m++) if (m < ng)
                dhev_recv(s.ec, s.LH_q[m] * s.q_l, s.c4pe, s.debye, s.k_ex, s.LH_hl[m], Gm[m], rl, Fm[m], E);
        });
        // LHG-charges of every other core (2526-2586): the cores k = sub
        // (mod S) whose charges (within rmax_q of the core) can come within
        // reach of the head sphere, listed by the workgroup in index order
        for (int base = 0; base < s.n_c; base += kCoreThreads) {
            const int k = base + tid;
            bool keep = false;
            if (k < s.n_c && k % S == sub) {
                const int ob = s.core_bead[k];
                if (ob != cb) {
                    const double3 dc = ld3(g.r + 3 * ob) - hc;
                    keep = !(sqrt(dot(dc, dc)) - hr - s.rmax_q > kReach);
                }
            }
            const int cnt = compact_ordered<kCoreThreads>(keep, k, s_list, s_wsum);
            for (int e = 0; e < cnt; e++) {
                const double *Qk = st.Q + 3 * (size_t)s_list[e] * s.Nq;
                for (int q = tid; q < s.Nq; q += blockDim.x) {
                    const double3 Qq = ld3(Qk + 3 * q);
#pragma unroll
                    for (int m = 0; m < kMaxGSites; m++) if (m < ng)
                        dhev_recv(s.ec, s.LH_q[m] * s.cq[q], s.c4pe, s.debye, s.k_ex, s.LH_hc[m], Gm[m], Qq, Fm[m], E);
                }
            }
            __syncthreads();
        }
        // LHG-free LH beads of other cores (2590-2644)
        for_near_set<kCoreThreads>(st, SET_LH, st.ch_lh, s.n_LH, hc, hr, ft, FT, s_list, s_wsum, [&](int jl) {
            if (jl / s.nlh == lho) return;
            int i2 = jl % s.nlh;
            i2 = (i2 < s.n_lh_n) ? i2 : i2 + s.n_lh_g;
            double3 rl = ld3(g.r_lh + 3 * jl);
#pragma unroll
            for (int m = 0; m < kMaxGSites; m++) if (m < ng)
                dhev_recv(s.ec, s.LH_q[m] * s.LH_q[i2], s.c4pe, s.debye, s.k_ex, (s.LH_hh[m] + s.LH_hh[i2]) / 2, Gm[m], rl, Fm[m], E);
        });
        // LHG-LHG with the other LH-bound cores (2649-2719)
        for_near_set<kCoreThreads>(st, SET_HEAD, st.ch_g, s.n_lhc * ng, hc, hr, ft, FT, s_list, s_wsum, [&](int x) {
            int k = x / ng, m1 = x % ng;
            if (k == lho) return;
            double3 G2 = ld3(st.G + 3 * x);
#pragma unroll
            for (int m = 0; m < kMaxGSites; m++) if (m < ng)
                dhev_recv(s.ec, s.LH_q[m] * s.LH_q[m1], s.c4pe, s.debye, s.k_ex, (s.LH_hh[m] + s.LH_hh[m1]) / 2, Gm[m], G2, Fm[m],
                          (k > lho) ? E : Ed);
        });
        // head - first C-term bead spring (2727-2777): force x10, torque x1
        if (ft == 0 && s.n_LH > 0) {
            double3 p = ld3(g.r_lh + 3 * (size_t)lho * s.nlh);
            #pragma unroll
        for (int m = 0; m < kMaxGSites; m++) if (m < ng) {
                int ti = m + s.n_lh_n;
                if (s.LH_conn[ti] != 1) continue;
                double3 dv = p - Gm[m];
                double d = sqrt(dot(dv, dv));
                double ks = s.LH_kstr[ti], l0 = s.LH_streq[ti];
                E += ks * (d - l0) * (d - l0) / 2;
                double3 S = make_double3((d - l0) * dv.x / d, (d - l0) * dv.y / d, (d - l0) * dv.z / d);
                F += make_double3(ks * S.x * 10, ks * S.y * 10, ks * S.z * 10);
                T += torque_rel(make_double3(ks * S.x, ks * S.y, ks * S.z), ld3(s.LH_g_pos + 3 * m), ac, bc, cc);
            }
        }
        #pragma unroll
        for (int m = 0; m < kMaxGSites; m++) if (m < ng) {
            F += Fm[m];
            T += torque_rel(Fm[m], ld3(s.LH_g_pos + 3 * m), ac, bc, cc);
        }
    }

    // --- anchored tails whose LH forces are redirected to this core, up to
    //     kMaxRT per pass over the LH partners ---
    if (phase == 2 && (terms & TERM_LH)) {
        for (int e0 = s.rt_off[kc]; e0 < s.rt_off[kc + 1]; e0 += kMaxRT) {
            const int ne = min(kMaxRT, s.rt_off[kc + 1] - e0);
            double3 rt[kMaxRT], Ft[kMaxRT];
            double tq[kMaxRT];
#pragma unroll
            for (int e = 0; e < kMaxRT; e++) {
                const int tj = s.rt_list[e0 + min(e, ne - 1)];
                rt[e] = ld3(g.r_t + 3 * tj);
                tq[e] = s.t_q[tj];
                Ft[e] = zero3();
            }
            double3 tc;
            double tr;
            points_sphere(rt, ne, tc, tr);
            for_near_set<kCoreThreads>(st, SET_LH, st.ch_lh, s.n_LH, tc, tr, ft, FT, s_list, s_wsum, [&](int i) {
                int idx = i % s.nlh;
                idx = (idx < s.n_lh_n) ? idx : idx + s.n_lh_g;
                const double3 rl = ld3(g.r_lh + 3 * i);
#pragma unroll
                for (int e = 0; e < kMaxRT; e++)
                    if (e < ne) dhev_recv(s.ec, s.LH_q[idx] * tq[e], s.c4pe, s.debye, s.k_ex, s.LH_ht[idx], rt[e], rl, Ft[e], Ed);
            });
            for_near_set<kCoreThreads>(st, SET_HEAD, st.ch_g, s.n_lhc * s.n_lh_g, tc, tr, ft, FT, s_list, s_wsum, [&](int x) {
                const int m = x % s.n_lh_g;
                const double3 Gx = ld3(st.G + 3 * x);
#pragma unroll
                for (int e = 0; e < kMaxRT; e++)
                    if (e < ne) dhev_recv(s.ec, s.LH_q[m] * tq[e], s.c4pe, s.debye, s.k_ex, s.LH_ht[m], rt[e], Gx, Ft[e], Ed);
            });
            double v[3 * kMaxRT];
#pragma unroll
            for (int e = 0; e < kMaxRT; e++) {
                v[3 * e] = Ft[e].x;
                v[3 * e + 1] = Ft[e].y;
                v[3 * e + 2] = Ft[e].z;
            }
            block_sum_n<kCoreThreads, 3 * kMaxRT>(v, red);
            if (tid == 0)
#pragma unroll
                for (int e = 0; e < kMaxRT; e++)
                    if (e < ne) {
                        const double3 ff = make_double3(v[3 * e], v[3 * e + 1], v[3 * e + 2]);
                        F += ff;
                        T += torque_abs(ff, rt[e], rc, ac, bc, cc);
                    }
        }
    }
    // --- anchor springs of this core's own tails (1971-2022); energy on the tail side ---
    if (phase == 2 && (terms & TERM_TAIL) && ft == 0 && s.n_tail > 0) {
        for (int jj = 0; jj < s.n_t; jj++) {
            int tj = kc * s.n_t + jj;
            if (s.t_fix[tj] != 1) continue;
            double3 ro = body_to_world(rc, ac, bc, cc, ld3(s.tail_pos + 3 * jj));
            double3 S = ld3(g.r_t + 3 * tj) - ro;
            double3 hs = make_double3(s.h * S.x, s.h * S.y, s.h * S.z);
            F += hs;
            T += torque_rel(hs, ld3(s.tail_pos + 3 * jj), ac, bc, cc);
        }
    }
    // block reduction in fixed order, then the core's 3 S partial sums in order
    double v[8] = {F.x, F.y, F.z, T.x, T.y, T.z, E, Ec};
    block_sum_n<kCoreThreads, 8>(v, red);
    if (tid == 0) {
        double *part = core_part + (size_t)kc * NS * 8;
#pragma unroll
        for (int i = 0; i < 8; i++) part[blockIdx.y * 8 + i] = v[i];
        __threadfence();
        if (atomicAdd(core_cnt + kc, 1) == NS - 1) {
            __threadfence();
            double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int j = 0; j < NS; j++)
#pragma unroll
                for (int i = 0; i < 8; i++) t[i] += part[j * 8 + i];
            core_cnt[kc] = 0;
            double *Fc = f.F + 3 * cb, *Tc = f.T + 3 * cb;
            Fc[0] += t[0]; Fc[1] += t[1]; Fc[2] += t[2];
            Tc[0] += t[3]; Tc[1] += t[4]; Tc[2] += t[5];
            if (Ecore) Ecore[kc] = t[6];
        }
    }
}

// The core-core charge sums of core kc: its row (as c1) and column (as c2)
// of the slot matrix that k_cc_work fills (func.cpp:1489-1566), one
// workgroup per core: the partner cores of kc's pairs are listed in index
// order, then threads take their (partner, unit) slots in that order and a
// fixed-order block sum adds them; adds to F, T of the core bead and owns
// the energy of the pairs where kc is c1.
constexpr int kGatherThreads = 256;
__global__ void __launch_bounds__(kGatherThreads) k_cc_gather(Stat s, Geo g, Frc f, const unsigned char *cc_valid,
                                                              int nchunk, const double *cc_slots, double *Ecc,
                                                              int terms) {
    __shared__ double red[10 * (kGatherThreads / 64)];
    __shared__ int s_list[kGatherThreads], s_wsum[kGatherThreads / 64];
    const int kc = blockIdx.x, tid = threadIdx.x;
    const int cb = s.core_bead[kc];
    double3 F = zero3(), T = zero3(), Tw = zero3();
    double Ec = 0.0;
    // without the electrostatic terms only the energy slot is cleared
    const int no = (terms & TERM_ELEC) ? s.n_c : 0;
    for (int base = 0; base < no; base += kGatherThreads) {
        const int o = base + tid;
        bool keep = false;
        if (o < no && o != kc) keep = cc_valid[kc < o ? (size_t)kc * s.n_c + o : (size_t)o * s.n_c + kc] != 0;
        const int cnt = compact_ordered<kGatherThreads>(keep, o, s_list, s_wsum);
        for (int x = tid; x < cnt * nchunk; x += kGatherThreads) {
            const int o2 = s_list[x / nchunk], ch = x % nchunk;
            const bool first = kc < o2;
            const size_t pidx = first ? (size_t)kc * s.n_c + o2 : (size_t)o2 * s.n_c + kc;
            const double *sl = cc_slots + 10 * (pidx * nchunk + ch);
            if (first) {
                F += make_double3(sl[0], sl[1], sl[2]);
                T += make_double3(sl[3], sl[4], sl[5]);
                Ec += sl[9];
            } else {
                F -= make_double3(sl[0], sl[1], sl[2]);
                Tw += make_double3(sl[6], sl[7], sl[8]);
            }
        }
        __syncthreads();
    }
    double v[10] = {F.x, F.y, F.z, T.x, T.y, T.z, Tw.x, Tw.y, Tw.z, Ec};
    block_sum_n<kGatherThreads, 10>(v, red);
    if (tid == 0) {
        const double3 ac = ld3(g.a + 3 * cb), bc = ld3(g.b + 3 * cb), cc = ld3(g.c + 3 * cb);
        const double3 tw = make_double3(v[6], v[7], v[8]);
        double *Fc = f.F + 3 * cb, *Tc = f.T + 3 * cb;
        Fc[0] += v[0]; Fc[1] += v[1]; Fc[2] += v[2];
        Tc[0] += v[3] + dot(ac, tw);
        Tc[1] += v[4] + dot(bc, tw);
        Tc[2] += v[5] + dot(cc, tw);
        if (Ecc) Ecc[kc] = v[9];
    }
}

// ---------------------------------------------------------------------------
// core-core charge sums (func.cpp:1489-1566)
// ---------------------------------------------------------------------------
// Flags core pairs (k1 < k2) inside 25 nm with both cores eligible and
// appends them to a list; the slot matrix is indexed by (k1, k2), so the
// list order does not affect the result.
__global__ void k_cc_list(Stat s, Geo g, unsigned char *valid, int *list, int *count) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long nc = s.n_c;
    if (t >= nc * nc) return;
    int k1 = (int)(t / nc), k2 = (int)(t % nc);
    unsigned char v = 0;
    if (k1 < k2) {
        int b1 = s.core_bead[k1], b2 = s.core_bead[k2];
        if (s.elig[b1] && s.elig[b2]) {
            double3 dv = ld3(g.r + 3 * b2) - ld3(g.r + 3 * b1);
            if (!(sqrt(dot(dv, dv)) > 25.0)) v = 1;
        }
    }
    valid[t] = v;
    if (v) {
        int pos = atomicAdd(count, 1);
        list[2 * pos] = k1;
        list[2 * pos + 1] = k2;
    }
}

constexpr int kCCThreads = 256;   // four wavefronts, a work unit each

// Work unit = (core pair, block of 64 charges of the first core, chunk of
// the second core's charges), one wavefront: lane = charge k of the first
// core; the second core's charge records are the same for every lane and
// come through scalar loads (no LDS, no barriers).  Slot per unit = {F on
// c1 (3), T on c1 body (3), T on c2 world (3), E}.  The torque on c2 is
// summed in world axes about its centre and turned into body components by
// k_cc_gather (equal to the reference's body-frame sum for an orthonormal
// frame).  Lanes past the last charge repeat it and are weighted by 0.
// exp(-kappa d) is the table form exp_neg_t: 2^(j/32) comes in as a kernel
// argument and is copied to LDS once (32 doubles over the 64 banks, so any
// mix of lanes reads it without conflict); the Taylor coefficients of e^r
// come in as a kernel argument too, so they sit in scalar registers and each
// Horner step is one FMA with a scalar addend (compile-time constants end up
// in vector registers, where every step first copies its addend).
template <bool WITH_E>
__global__ void __launch_bounds__(kCCThreads) k_cc_work(Stat s, Geo g, const double *__restrict__ Qp,
                                                       const int *__restrict__ list, const int *__restrict__ count,
                                                       int nkb, int nlc, double *__restrict__ slots, ExpCoef6 ec,
                                                       Exp2Tab et) {
    __shared__ double etab[32];
    if (threadIdx.x < 32) etab[threadIdx.x] = et.t[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nsub = nkb * nlc, nunits = (*count) * nsub, Nq = s.Nq;
    const int chunk = (Nq + nlc - 1) / nlc;
    const double kappa = s.debye, kex = s.k_ex, sg = s.sig_cc, c12 = 12.0 / sg, c6 = 6.0 / sg;
    for (int u = blockIdx.x * (kCCThreads / 64) + wv; u < nunits; u += gridDim.x * (kCCThreads / 64)) {
        const int pr = u / nsub, sub = u % nsub, kb = sub / nlc, lc = sub % nlc;
        const int k1 = list[2 * pr], k2 = list[2 * pr + 1];
        const int b1 = s.core_bead[k1];
        const int k = min(kb * 64 + lane, Nq - 1);
        const double act = (kb * 64 + lane < Nq) ? 1.0 : 0.0;
        const int l0 = lc * chunk, l1 = min(Nq, l0 + chunk);
        const double *pk = Qp + 4 * ((size_t)k1 * Nq + k);
        const double3 Qk = make_double3(pk[0], pk[1], pk[2]);
        const double keq = s.k_e * pk[3];
        // Fk = sum_l f_l, f_l = g_l (Qk - Q_l)
        double3 Fk = zero3();
        double E = 0.0;
        const double *pl = Qp + 4 * (size_t)k2 * Nq;
        for (int l = l0; l < l1; l++) {
            const double *d = pl + 4 * l;
            const double3 z = make_double3(Qk.x - d[0], Qk.y - d[1], Qk.z - d[2]);
            // Debye-Hueckel + excluded volume of the reference (dh_core, ev_core)
            // with 1/d from a refined reciprocal square root
            const double r2s = dot(z, z);
            const double inv = rsqrt_f64(r2s);
            const double dist = r2s * inv;
            const double t = keq * d[3] * exp_neg_t(-kappa * dist, etab, ec);
            if (WITH_E) E += t * inv;
            double g1d = t * (kappa + inv) * (inv * inv);  // (kappa d + 1) / d^3
            if (r2s <= kD8Sq) {  // the reference's d <= 8, exactly
                const double x = sg * inv, x2 = x * x, x6 = x2 * x2 * x2, x7 = x6 * x, x13 = x6 * x7;
                if (WITH_E) E += kex * (x6 * x6 - x6);
                g1d += kex * (c12 * x13 - c6 * x7) * inv;
            }
            Fk.x = __builtin_fma(g1d, z.x, Fk.x);
            Fk.y = __builtin_fma(g1d, z.y, Fk.y);
            Fk.z = __builtin_fma(g1d, z.z, Fk.z);
        }
        Fk = act * Fk;
        const double3 a1 = ld3(g.a + 3 * b1), bb1 = ld3(g.b + 3 * b1), c1 = ld3(g.c + 3 * b1);
        const double3 T1 = torque_rel(Fk, ld3(s.cpos + 3 * k), a1, bb1, c1);
        // torque on c2 about its centre r2 from -f_l at lever Q_l - r2 = w - z_l,
        // w = Qk - r2, z_l = Qk - Q_l: sum_l (w - z_l) x (-g_l z_l) = Fk x w,
        // since z_l x z_l = 0 -- the force sum alone gives it
        const double3 T2 = cross(Fk, Qk - ld3(g.r + 3 * s.core_bead[k2]));
        double v[10] = {Fk.x, Fk.y, Fk.z, T1.x, T1.y, T1.z, T2.x, T2.y, T2.z, act * E};
#pragma unroll
        for (int q = 0; q < 10; q++)
            for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
        if (lane == 0) {
            double *slot = slots + 10 * ((size_t)(k1 * s.n_c + k2) * nsub + sub);
#pragma unroll
            for (int q = 0; q < 10; q++) slot[q] = v[q];
        }
    }
}

// Energy partials -> one double, fixed order (one workgroup).
__global__ void k_ereduce(const double *Epart, int ne, double *out) {
    __shared__ double red[16];
    double v = 0.0;
    for (int i = threadIdx.x; i < ne; i += blockDim.x) v += Epart[i];
    double t = block_sum<1024>(v, red);
    if (threadIdx.x == 0) out[0] = t;
}

__global__ void k_fill(double *p, long long n, double v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i(int *p, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

void launch_sites(const Stat &s, const Geo &g, const Sites &st, int *cc_count, hipStream_t q, bool fuse_chunks) {
    long long tot = (long long)s.n_c * s.Nq + (long long)s.n_lhc * s.n_lh_g;
    const int nsb = nblk(std::max(tot, 1LL), 256);
    const int nch = nchunks(s.n_tail) + nchunks(s.n_LH) + nchunks(s.n_lhc * s.n_lh_g) + nchunks(s.n_lnk);
    if (!st.use_cells && fuse_chunks) {
        // the chunk spheres in the same launch
        k_sites<<<nsb + nblk(nch, 4), 256, 0, q>>>(s, g, st, cc_count, nsb);
        return;
    }
    k_sites<<<nsb, 256, 0, q>>>(s, g, st, cc_count, nsb);
    if (!st.use_cells) {
        if (nch > 0) k_chunks<<<nblk(nch, 4), 256, 0, q>>>(s, g, st);
        return;
    }
    {
        // the grids are built from this pass's positions
        Sites sp = st;
        sp.cg[SET_TAIL].pos = g.r_t;
        sp.cg[SET_TAIL].idx = nullptr;
        sp.cg[SET_TAIL].n = s.n_tail;
        sp.cg[SET_LH].pos = g.r_lh;
        sp.cg[SET_LH].idx = nullptr;
        sp.cg[SET_LH].n = s.n_LH;
        sp.cg[SET_HEAD].pos = st.G;
        sp.cg[SET_HEAD].idx = nullptr;
        sp.cg[SET_HEAD].n = s.n_lhc * s.n_lh_g;
        sp.cg[SET_LINKER].pos = g.r;
        sp.cg[SET_LINKER].idx = s.lnk;
        sp.cg[SET_LINKER].n = s.n_lnk;
        k_cells<<<NSETS, kCellThreads, 0, q>>>(sp);
        k_cell_sort<<<512, 256, 0, q>>>(s, st);
    }
}

bool launch_frames_sites(const Stat &s, const Geo &g, const Sites &st, int *cc_count, hipStream_t q) {
    if (s.nfseg == 0 || st.use_cells) return false;
    const long long tot = (long long)s.n_c * s.Nq + (long long)s.n_lhc * s.n_lh_g;
    const int nsb = nblk(std::max(tot, 1LL), 256), nbeta = nblk((long long)s.n_tail + s.n_LH, 256);
    const int nch = nchunks(s.n_tail) + nchunks(s.n_LH) + nchunks(s.n_lhc * s.n_lh_g) + nchunks(s.n_lnk);
    k_frames_sites<<<s.nfseg + nbeta + nsb + nblk(nch, 4), 256, 0, q>>>(s, g, st, cc_count, nbeta, nsb);
    return true;
}

void launch_forces(const Stat &s, const Geo &g, const Sites &st, const Frc &f, const ForceScratch &w,
                   double *Epart, int terms, hipStream_t q, MarkFn mark, void *arg, const ForceStreams *fs,
                   bool sites_done) {
    const int off_mech = 0, off_lnk = s.n, off_tail = off_lnk + s.n_lnk, off_lh = off_tail + s.n_tail,
              off_core = off_lh + s.n_LH, off_cc = off_core + s.n_c;
    auto E = [&](int off) { return Epart ? Epart + off : nullptr; };
    auto M = [&](int id, int ph) { if (mark) mark(arg, id, ph); };
    // k_mech assigns F, T of every main bead, then k_linker adds to the
    // linker beads and k_core to the cores (disjoint beads, so the two run
    // side by side); k_tail and k_lh assign their own arrays; k_cc_gather
    // adds k_cc_work's slots to the cores last
    auto tail = [&](hipStream_t qq) {
        if (s.n_tail) { M(K_TAIL, 0); k_tail<<<s.n_tail, kTailThreads, 0, qq>>>(s, g, st, f, E(off_tail), terms); M(K_TAIL, 1); }
    };
    auto lh = [&](hipStream_t qq) {
        if (s.n_LH) { M(K_LH, 0); k_lh<<<s.n_LH, kLHThreads, 0, qq>>>(s, g, st, f, E(off_lh), terms); M(K_LH, 1); }
    };
    auto cc = [&](hipStream_t qq) {
        if (!(terms & TERM_ELEC)) return;
        M(K_CCLIST, 0);
        k_cc_list<<<nblk((long long)s.n_c * s.n_c, 256), 256, 0, qq>>>(s, g, w.cc_valid, w.cc_list, w.cc_count);
        M(K_CCLIST, 1);
        M(K_CCWORK, 0);
        if (Epart)
            k_cc_work<true><<<w.cc_grid, kCCThreads, 0, qq>>>(s, g, st.Qp, w.cc_list, w.cc_count, w.cc_nkb,
                                                              w.cc_nlc, w.cc_slots, kExpCoef6, kExp2Tab);
        else
            k_cc_work<false><<<w.cc_grid, kCCThreads, 0, qq>>>(s, g, st.Qp, w.cc_list, w.cc_count, w.cc_nkb,
                                                               w.cc_nlc, w.cc_slots, kExpCoef6, kExp2Tab);
        M(K_CCWORK, 1);
    };
    auto mech = [&]() { M(K_MECH, 0); k_mech<<<nblk(s.n, 64), 64, 0, q>>>(s, g, f, E(off_mech), terms); M(K_MECH, 1); };
    auto linker = [&](hipStream_t qq) {
        if (s.n_lnk) { M(K_LINKER, 0); k_linker<<<s.n_lnk, kLinkerThreads, 0, qq>>>(s, g, st, f, E(off_lnk), terms); M(K_LINKER, 1); }
    };
    auto core = [&]() {
        M(K_CORE, 0);
        const int *cs = w.core_split;
        k_core<<<dim3(s.n_c, cs[0] + cs[1] + cs[2]), kCoreThreads, 0, q>>>(s, g, st, f, w.core_part, w.core_cnt,
                                                                          E(off_core), terms, cs[0], cs[1]);
        M(K_CORE, 1);
    };
    // with branches (force pass 2, after the factor has joined) the chunk
    // spheres take the sites launch; in pass 1 beside the factor the extra
    // launch measured faster
    if (!sites_done) { M(K_SITES, 0); launch_sites(s, g, st, w.cc_count, q, fs != nullptr); M(K_SITES, 1); }
    // The enqueue orders were measured (cfg2): on one stream, beside the
    // factorisation, tail and LH first and core-core last; with branches,
    // core-core first and tail and LH last (enqueued first there, their
    // thousands of small workgroups took the CUs ahead of the longer chains).
    if (!fs) {
        tail(q);
        lh(q);
        mech();
        linker(q);
        core();
        cc(q);
    } else {
        (void)hipEventRecord(fs->fork, q);
        for (int i = 0; i < 3; i++) (void)hipStreamWaitEvent(fs->b[i], fs->fork, 0);
        cc(fs->b[2]);
        mech();
        (void)hipEventRecord(fs->mech, q);
        (void)hipStreamWaitEvent(fs->b[1], fs->mech, 0);
        linker(fs->b[1]);
        core();
        tail(fs->b[0]);
        lh(fs->b[1]);
        for (int i = 0; i < 3; i++) {
            (void)hipEventRecord(fs->join[i], fs->b[i]);
            (void)hipStreamWaitEvent(q, fs->join[i], 0);
        }
    }
    M(K_CCGATHER, 0);
    k_cc_gather<<<s.n_c, kGatherThreads, 0, q>>>(s, g, f, w.cc_valid, w.cc_nkb * w.cc_nlc, w.cc_slots, E(off_cc), terms);
    M(K_CCGATHER, 1);
}

int energy_slots(const Stat &s) { return s.n + s.n_lnk + s.n_tail + s.n_LH + 2 * s.n_c; }

void launch_energy_reduce(const double *Epart, int ne, double *out, hipStream_t q) {
    k_ereduce<<<1, 1024, 0, q>>>(Epart, ne, out);
}

void launch_fill(double *p, long long n, double v, hipStream_t q) {
    if (n > 0) k_fill<<<nblk(n, 256), 256, 0, q>>>(p, n, v);
}

void launch_fill_int(int *p, long long n, int v, hipStream_t q) {
    if (n > 0) k_fill_i<<<nblk(n, 256), 256, 0, q>>>(p, n, v);
}

}  // namespace bd
