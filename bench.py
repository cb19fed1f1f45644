"""Benchmark: BD steps/s of the 100-nucleosome fiber on one MI355X.

One step = one Ermak-McCammon predictor-corrector step of the whole fiber
(RPY tensor over cores + linker beads, its Cholesky factor, noise product,
two force passes, rotation/translation, frame rebuild), as main_cuda does
(func_cuda.cu:3434-3680).  Noise is drawn on the host in the reference
order (MT19937 seed 5489) for every step before timing and uploaded to
HBM, so the timed region reads inputs already resident on the GPU.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build runs one fiber on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from brownian_dynamics_cuda_amd import engine as E  # noqa: E402
from brownian_dynamics_cuda_amd import fiber as Fb  # noqa: E402

PEAK_FP64 = 78.6e12  # MI355X FP64 (matrix and vector), FLOP/s
PEAK_FP32_MFMA = 157.3e12
PEAK_BF16_MFMA = 2.5e15  # dense
PEAK_HBM = 8.0e12    # B/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", type=int, default=2, help="benchmark configuration (1-5)")
    ap.add_argument("--timing-steps", type=int, default=5)
    ap.add_argument("--host-io-steps", type=int, default=50,
                    help="steps timed through bd_step with host noise buffers (PCIe-inclusive rate)")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="reference CPU steps to time (0: skip; default 2 for the 6- to 100-nucleosome "
                         "fibers, 0 for the 500- and 1000-nucleosome ones, where one CPU step "
                         "takes many minutes)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--chol", choices=("fp64", "fp32", "bf16"), default="fp64",
                    help="factor arithmetic of the timed steps (the headline is fp64; the "
                         "others are for profiling the mixed modes)")
    ap.add_argument("--eager", action="store_true",
                    help="launch the step's kernels one by one instead of replaying the "
                         "captured hipGraph (same kernels; for profilers)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the timed steps, run the library's streaming copy kernel "
                         "once over MIB MiB (calibrates HBM traffic counters)")
    ap.add_argument("--traffic-file", default=None,
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary for the roofline traffic "
                         "(default: profiles/hbm_traffic_cfg<config>.txt; traffic is null "
                         "when no summary exists for the configuration)")
    ap.add_argument("--e2e-steps", type=int, default=200,
                    help="steps of the end-to-end rate with the host noise draw overlapped "
                         "(0: skip)")
    return ap.parse_args()


def read_traffic(path, kernel):
    """HBM bytes per launch of `kernel` from a committed counter summary
    (lines 'traffic <kernel> <bytes_per_launch>'); None if absent."""
    try:
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic" and w[1] == kernel:
                    return float(w[2])
    except OSError:
        pass
    return None


def noise_stream(fib, nsteps, seed=5489):
    """(rr, p) for nsteps steps in the reference draw order."""
    rng = E.HostRng(seed)
    rv = fib.r_var()
    n3 = 3 * fib.n
    nD3 = 3 * (fib.n + fib.n_tail + fib.n_LH)
    rr = np.zeros((nsteps, n3))
    p = np.zeros((nsteps, nD3))
    for s in range(nsteps):
        rr[s], p[s] = rng.step_noise(s, rv, nD3)
    return rr, p


def chol_flops(m):
    return m ** 3 / 3.0


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs one fiber on one GPU; --gpus {args.gpus} is not supported "
              "(several GPUs would take one fiber each, or split a fiber's force pass by "
              "receiver ranges)", file=sys.stderr)
        sys.exit(2)
    cfg = args.config
    fib, st = Fb.make_config(cfg)
    n_hi3 = 3 * fib.n_hi
    K, W = args.steps, args.warmup
    t0 = time.time()
    rr, p = noise_stream(fib, W + K)
    t_noise = time.time() - t0
    n3, nD3 = rr.shape[1], p.shape[1]

    mode = {"fp64": E.CHOL_FP64, "fp32": E.CHOL_FP32, "bf16": E.CHOL_BF16}[args.chol]
    eng = E.BDEngine(fib, chol_mode=mode, use_graph=not args.eager)
    eng.set_state(st, rebuild=True)
    d_rr = E.DeviceBuffer(rr.nbytes)
    d_p = E.DeviceBuffer(p.nbytes)
    d_rr.upload(rr)
    d_p.upload(p)
    d_e = E.DeviceBuffer(8 * (W + K))

    def run(s0, s1):
        for s in range(s0, s1):
            eng.step_device(d_rr.ptr + 8 * n3 * s, d_p.ptr + 8 * nD3 * s, d_e.ptr + 8 * s)

    run(0, W)
    eng.synchronize()
    t0 = time.perf_counter()
    run(W, W + K)
    eng.synchronize()
    dt = time.perf_counter() - t0
    energies = d_e.download(W + K)
    final = eng.get_state()
    finite = bool(np.all(np.isfinite(final.r)) and np.all(np.isfinite(energies)))
    ms = 1e3 * dt / K
    steps_per_s = K / dt
    if args.copy_probe > 0:
        nb = args.copy_probe << 20
        src, dst = E.DeviceBuffer(nb), E.DeviceBuffer(nb)
        dst.copy_from(src, nb)
        src.free()
        dst.free()

    # PCIe-inclusive rate: bd_step copies rr and p from host memory and the
    # energy back every step (the func_cuda.h path also copies r_lh back)
    host_io = None
    if args.host_io_steps > 0:
        nh = min(args.host_io_steps, W + K)
        t1 = time.perf_counter()
        for s in range(nh):
            eng.step(rr[s], p[s])
        host_io = nh / (time.perf_counter() - t1)

    # per-kernel HIP-event timing on the step's stream (a few instrumented steps)
    eng.set_timing(True)
    for s in range(args.timing_steps):
        eng.step(rr[s], p[s])
    kt = eng.timing()
    eng.set_timing(False)
    counts = eng.work_counts()
    per_step = {k: v[0] / args.timing_steps for k, v in kt.items()}
    dominant = max(per_step, key=per_step.get)

    # roofline of the dominant kernel
    tot_ms, launches = kt[dominant]
    avg_s = tot_ms / launches / 1e3
    # super-panel trailing updates: the lower triangle of each trailing
    # matrix (m_s rows) takes a rank-K update, K = 64 NB: K m_s (m_s + 1)
    # FLOPs per launch; the block-column launches carry the rest of n^3/3
    nb = eng.chol_schedule()
    nt = (n_hi3 + 63) // 64
    trail_fl = []
    k0 = 0
    while nb and k0 + nb < nt:
        mrows = n_hi3 - 64 * (k0 + nb)
        trail_fl.append(64.0 * nb * mrows * (mrows + 1))
        k0 += nb
    if dominant in ("chol_step", "potrf"):
        # the factorisation's algorithmic FLOPs less the trailing updates',
        # spread over the block-column launches (one per 64 columns after
        # the first); with super-panels these are the FP64 panel columns
        per_launch = (chol_flops(n_hi3) - sum(trail_fl)) / (launches / args.timing_steps)
        pipe = "FP64 MFMA" if args.chol == "fp64" or trail_fl else f"FP64 + {args.chol.upper()} MFMA"
        roof = dict(bound="mfma", achieved=per_launch / avg_s / 1e12, peak=PEAK_FP64 / 1e12,
                    unit="TFLOP/s", kernel=dominant, pipe=pipe,
                    flop_per_launch=per_launch, avg_launch_us=avg_s * 1e6)
    elif dominant == "chol_trail":
        # averaged over the launches of one factorisation, priced against the
        # matrix peak of the mode's arithmetic; the three-slice BF16 form
        # makes six BF16 products per algorithmic FLOP
        per_launch = sum(trail_fl) / len(trail_fl)
        peak, pipe, mult = {"fp64": (PEAK_FP64, "FP64 MFMA", 1.0),
                            "fp32": (PEAK_FP32_MFMA, "FP32 MFMA", 1.0),
                            "bf16": (PEAK_BF16_MFMA, "BF16 MFMA, six products per FLOP", 6.0)}[args.chol]
        roof = dict(bound="mfma", achieved=mult * per_launch / avg_s / 1e12, peak=peak / 1e12,
                    unit="TFLOP/s", kernel=dominant, pipe=pipe,
                    flop_per_launch=per_launch, avg_launch_us=avg_s * 1e6, panel_tiles=nb)
    else:
        # Debye-Hueckel charge sums: 38 FLOPs per charge pair, FP64 on the
        # vector ALUs, whose MI355X peak equals the FP64 matrix peak
        pair_flops = 38.0
        per_launch = counts["core_charge_pairs"] * pair_flops
        roof = dict(bound="mfma", achieved=per_launch / avg_s / 1e12, peak=PEAK_FP64 / 1e12,
                    unit="TFLOP/s", kernel=dominant, pipe="FP64 VALU",
                    flop_per_launch=per_launch, avg_launch_us=avg_s * 1e6)
    roof["frac"] = roof["achieved"] / roof["peak"]
    # the committed summaries are of FP64-mode steps: another factor mode
    # takes its traffic only from an explicit --traffic-file
    tfile = args.traffic_file or (os.path.join(ROOT, "profiles", f"hbm_traffic_cfg{cfg}.txt")
                                  if args.chol == "fp64" else None)
    roof["traffic"] = read_traffic(tfile, dominant) if tfile else None

    # end to end: the host draws each batch's noise (MT19937 + polar normals,
    # main.cpp:479-486) on a thread while the GPU runs the previous batch
    # (bd_run_steps: upload, steps, energies back), as the host driver does
    e2e = None
    if args.e2e_steps > 0:
        import threading
        bsz = 50
        nb_ = max(1, args.e2e_steps // bsz)
        rng2 = E.HostRng(12345)
        rv = fib.r_var()

        def draw(step0, out):
            a = np.zeros((bsz, n3))
            b = np.zeros((bsz, nD3))
            for k in range(bsz):
                a[k], b[k] = rng2.step_noise(step0 + k, rv, nD3)
            out.append((a, b))

        eng.set_state(st, rebuild=True)
        eng.synchronize()
        t1 = time.perf_counter()
        cur = []
        draw(0, cur)
        for bi in range(nb_):
            nxt = []
            th = threading.Thread(target=draw, args=((bi + 1) * bsz, nxt)) if bi + 1 < nb_ else None
            if th:
                th.start()
            eng.run_steps(*cur[0])
            if th:
                th.join()
            cur = nxt
        de = time.perf_counter() - t1
        e2e = dict(steps_per_s=nb_ * bsz / de, steps=nb_ * bsz, batch=bsz,
                   note="host noise drawn on one host thread, overlapped with the previous batch")

    # mixed-precision factors: same state and noise, one step each way;
    # displacement error against FP64, factor error ||L L^T - D|| / ||D||
    # (through random probes L (L^T v) - D v above 4096 rows), the covariance
    # error of the FP64-refined noise, steps/s (refinement on, the default)
    mixed = None
    if not args.no_mixed:
        mixed = {}
        eng64 = E.BDEngine(fib, chol_mode=E.CHOL_FP64, use_graph=False)
        eng64.set_state(st, rebuild=True)
        x0 = eng64.get_state()
        eng64.step(rr[0], p[0])
        a = eng64.get_state()
        disp64 = np.concatenate([a.r - x0.r, a.r_t - x0.r_t, a.r_lh - x0.r_lh])
        eng64.close()
        for mode, name in ((E.CHOL_FP32, "fp32"), (E.CHOL_BF16, "bf16x3")):
            engm = E.BDEngine(fib, chol_mode=mode, use_graph=True)
            engm.set_state(st, rebuild=True)
            engm.step(rr[0], p[0])
            b = engm.get_state()
            disp = np.concatenate([b.r - x0.r, b.r_t - x0.r_t, b.r_lh - x0.r_lh])
            rel = float(np.linalg.norm(disp - disp64) / np.linalg.norm(disp64))
            engm.set_state(st, rebuild=True)
            D = engm.get_tensor()
            engm.factor()
            L, _ = engm.get_factor()
            if D.shape[0] <= 4096:
                ferr = float(np.linalg.norm(L @ L.T - D) / np.linalg.norm(D))
            else:
                g = np.random.default_rng(0)
                ferr = 0.0
                for _ in range(3):
                    v = g.standard_normal(D.shape[0])
                    Dv = D @ v
                    ferr = max(ferr, float(np.linalg.norm(L @ (L.T @ v) - Dv) / np.linalg.norm(Dv)))
            # covariance of the refined noise, B = L (I + X / 2) with
            # X = inv(L) D inv(L)^T - I (what bd_set_refine applies):
            # ||B B^T v - D v|| / ||D v|| on random probes
            from scipy.linalg import solve_triangular
            g = np.random.default_rng(1)
            cov = 0.0
            for _ in range(3):
                v = g.standard_normal(D.shape[0])
                y = L.T @ v
                bt = 0.5 * (y + solve_triangular(L, D @ solve_triangular(L.T, y, lower=False), lower=True))
                bbt = 0.5 * (L @ bt + D @ solve_triangular(L.T, bt, lower=False))
                Dv = D @ v
                cov = max(cov, float(np.linalg.norm(bbt - Dv) / np.linalg.norm(Dv)))
            del D, L
            engm.set_state(st, rebuild=True)
            d_e2 = E.DeviceBuffer(8 * (W + K))
            for s in range(W):
                engm.step_device(d_rr.ptr + 8 * n3 * s, d_p.ptr + 8 * nD3 * s, d_e2.ptr + 8 * s)
            engm.synchronize()
            t1 = time.perf_counter()
            for s in range(W, W + K):
                engm.step_device(d_rr.ptr + 8 * n3 * s, d_p.ptr + 8 * nD3 * s, d_e2.ptr + 8 * s)
            engm.synchronize()
            dm = time.perf_counter() - t1
            mixed[name] = dict(steps_per_s=K / dm, ms_per_step=1e3 * dm / K,
                               rel_displacement_err_vs_fp64=rel, factor_rel_err=ferr,
                               refined_covariance_rel_err=cov)
            engm.close()
            print(f"[bench] {name}: {K / dm:.2f} steps/s, displacement err {rel:.2e}, "
                  f"factor err {ferr:.2e}, refined covariance err {cov:.2e}", file=sys.stderr, flush=True)

    # CPU baseline: the reference's own func.cpp composed into one step
    cpu = None
    if args.cpu_steps is None:
        args.cpu_steps = 0 if args.config in (3, 5) else 2
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        kind = "reference" if O.ref_available() else "port"
        orc = O.Oracle(fib, "ref" if kind == "reference" else "or")
        fs = orc.prepare(st)
        t1 = time.perf_counter()
        for s in range(args.cpu_steps):
            orc.step(fs, rr[s], p[s])
        dc = time.perf_counter() - t1
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as fcpu:
                model = next((ln.split(":", 1)[1].strip() for ln in fcpu if ln.startswith("model name")),
                             model)
        except OSError:
            pass
        cpu = dict(value=args.cpu_steps / dc, unit="steps/s", cores=1, kind=kind,
                   sample=f"{args.cpu_steps} steps of the same fiber, noise and start state "
                          f"({dc:.1f} s), one thread",
                   host=dict(cpu=model, threads_available=len(os.sched_getaffinity(0))))

    out = {
        "metric": "BD steps/s, 100-nucleosome fiber: FP64 vs mixed Cholesky at <=1e-6 rel err",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.chol == "fp64" else f"f64 (factor trailing updates {args.chol})",
        "data": "synthetic (seeded zigzag fiber, host MT19937 noise)",
        "config": {"workload": f"cfg{cfg}: {fib.n_c}-nucleosome fiber, {fib.n} cores+linker beads, "
                               f"{fib.n_tail} tail beads, {fib.n_LH} free LH beads, "
                               f"D {n_hi3}x{n_hi3} FP64", "n_c": fib.n_c, "n_hi3": n_hi3},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "kernels_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items(), key=lambda x: -x[1])},
        "work": counts,
        "finite": finite,
        "pcie_inclusive_steps_per_s": host_io,
        "end_to_end": e2e,
        "noise_gen_s": round(t_noise, 2),
    }
    if cpu:
        out["speedup_vs_cpu"] = steps_per_s / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
