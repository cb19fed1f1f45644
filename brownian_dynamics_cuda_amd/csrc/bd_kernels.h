// bd_kernels.h -- host-side launchers of the BD step kernels.
#pragma once
#include "bd_dev.h"

namespace bd {

// the message bd_last_error() returns (per host thread)
void set_last_error(const char *msg);

// Scratch of the core-core charge sums.
struct ForceScratch {
    unsigned char *cc_valid;  // n_c * n_c
    int *cc_list;             // 2 * n_c * (n_c - 1) / 2
    int *cc_count;            // 1
    double *cc_slots;         // n_c * n_c * nkb * nlc * 10
    int cc_nkb, cc_nlc;       // work units per pair: 64-charge blocks x chunks
    int cc_grid;
    // k_core runs core_split[p] workgroups per core on its phase p; each
    // leaves 8 partial sums in core_part and the last to finish (core_cnt)
    // adds them up in order
    int core_split[3];
    double *core_part;        // n_c * (sum of core_split) * 8
    int *core_cnt;            // n_c, zero between passes
};

// forces (bd_forces.hip)
// world charge and head sites, partner chunk spheres (with fuse_chunks in
// the sites launch itself); resets the core-core pair count of the pass
void launch_sites(const Stat &s, const Geo &g, const Sites &st, int *cc_count, hipStream_t q,
                  bool fuse_chunks = false);
// mark(arg, id, phase): optional timing hook called around each kernel
// (phase 0 before, 1 after); ids are the K_* values below.
enum KernelId { K_SITES, K_MECH, K_LINKER, K_TAIL, K_LH, K_CORE, K_CCLIST, K_CCWORK,
                K_RPY, K_CHOL_STEP, K_NOISE, K_FRAMES, K_BETAS, K_TRANSLATE,
                K_COMMIT, K_EREDUCE, K_CCGATHER, K_CHOL_TRAIL, K_CHOL_DIAG, K_REFINE, K_COUNT };
typedef void (*MarkFn)(void *, int, int);
// Extra streams of a force pass: after the charge sites, the tail, the
// linker-histone (then, once mechanics is done, linker) and the core-core
// branches run beside mechanics + core on the pass's stream, and the
// core-core gather waits for all of them.
struct ForceStreams {
    hipStream_t b[3];
    hipEvent_t fork, mech, join[3];
};
void launch_forces(const Stat &s, const Geo &g, const Sites &st, const Frc &f, const ForceScratch &w,
                   double *Epart, int terms, hipStream_t q, MarkFn mark = nullptr, void *arg = nullptr,
                   const ForceStreams *fs = nullptr, bool sites_done = false);
// the predicted state's frame passes and bending angles with force pass 2's
// sites and chunk spheres in one launch (false, nothing launched: the chain
// has no frame segments or the partner search uses cell grids)
bool launch_frames_sites(const Stat &s, const Geo &g, const Sites &st, int *cc_count, hipStream_t q);
int energy_slots(const Stat &s);
void launch_energy_reduce(const double *Epart, int ne, double *out, hipStream_t q);
void launch_fill(double *p, long long n, double v, hipStream_t q);
void launch_fill_int(int *p, long long n, int v, hipStream_t q);

// frames and integrator (bd_integrate.hip)
void launch_frames(const Stat &s, const Geo &g, hipStream_t q);         // update_Euler_Angle
void launch_betas(const Stat &s, const Geo &g, int tails, int lh, hipStream_t q);
// both, as one launch, where the chain cuts into frame segments or has at
// most 1025 beads; false (nothing launched) otherwise
bool launch_frames_betas(const Stat &s, const Geo &g, int tails, int lh, hipStream_t q);
// the translation from src to dst with forces f (f2.F set: the mean of f and
// f2) and, with rr set, the rotation of the main beads' frames in the same
// launch (torques likewise)
void launch_translate(const Stat &s, int hi_all, const Geo &src, const Geo &dst, const Frc &f, const Frc &f2,
                      const double *rd, const double *rad_all, const double *rr, hipStream_t q);
void launch_commit(const Stat &s, const Geo &from, const Geo &to, hipStream_t q);

// diffusion tensor, Cholesky, noise (bd_chol.hip)
struct CholWork {
    double *A;       // mpad x ld, row-major; lower triangle holds D then L
    int m, mpad, ld;
    double *Linv;    // (ld/64) tiles of 64x64: inverses of the diagonal factors, row-major
    int *neg;        // negative-pivot counter
    int mode;        // BD_CHOL_*
    // block columns with at least this many trailing tiles take the
    // two-launch form (k_panel + k_update): about two rounds of workgroups
    // at one per CU
    int split_tiles;
    // super-panel width in block columns: 0 chooses (8 from 64 block
    // columns up, none below), -1 never, k >= 2 always k
    int panel_tiles;
    // FP64 refinement of B xi in the FP32 / BF16 modes (1: on): xs, the
    // hydrodynamic-set positions D was last assembled from (ld doubles);
    // rv, two work vectors of ld doubles; pr, D inv(L~)^T p (ld)
    int refine;
    double *xs, *rv, *pr;
    // the super-panels' panel tiles packed once in matrix-core operand order
    // (FP64 / FP32 modes) or split into three BF16 slices (BF16 mode);
    // split_bytes(ld, sp_nb) bytes; null: the trailing update reads the
    // parked tiles directly
    void *sp;
    int sp_nb;
    // smallest trailing matrix (block rows) that takes the packed form; -1:
    // the mode's default
    int pack_min_rows;
};
// bytes of the packed / split panel buffer for leading dimension ld and
// panels of nb
size_t split_bytes(int ld, int nb);
// super-panel width the factor of c uses (0: one block column at a time)
int chol_panel_tiles(const CholWork &c);
// D into c.A; also clears the negative-pivot count of the factor that
// follows, and with factor_first factors and inverts the first diagonal tile.
// Where refine_active(c), the hydrodynamic-set positions go to c.xs too.
void launch_rpy(const Stat &s, int n_hi, const Geo &g, const double *rad_all, const CholWork &c,
                hipStream_t q, bool factor_first = false);
// the factor of the D that launch_rpy(factor_first) left in c.A; gate
// (optional) is recorded on q after the launch of block column gate_kb
// (after the last one when gate_kb is past the end)
void launch_cholesky(const CholWork &c, hipStream_t q, MarkFn mark = nullptr, void *mark_arg = nullptr,
                     hipEvent_t gate = nullptr, int gate_kb = -1);
// factor and inverse of the diagonal tile d alone (clear: restart the
// negative-pivot count first)
void launch_diag(const CholWork &c, int d, bool clear, hipStream_t q);
// a caller's matrix (lower triangle of row-major D) into c.A with the
// identity padding; the factor in c.A back out as row-major lower L
void launch_pack(const double *D, int ldd, const CholWork &c, hipStream_t q);
void launch_unpark(const CholWork &c, double *L, int ldl, hipStream_t q);
void launch_noise(const Stat &s, int n_hi, int n_D3, const CholWork &c, const double *p,
                  const double *rad_all, double *rd, hipStream_t q, const double *du = nullptr);
// whether the noise of factor c takes the FP64 refinement, and the
// refinement itself: c.pr = D inv(L~)^T p over the HI rows (the du of
// launch_noise)
bool refine_active(const CholWork &c);
void launch_refine(const Stat &s, const CholWork &c, const double *p, const double *rad_all, hipStream_t q);

}  // namespace bd
