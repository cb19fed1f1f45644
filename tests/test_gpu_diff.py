"""The differentiable trajectory (brownian_dynamics_cuda_amd/diff.py) on the
GPU: the hand-written HIP core-core kernels and their adjoint against a
dense torch evaluation of the same pair sums (func.cpp:1489-1566); the
torch step against the oracle's FP64 step; the adjoint gradient against
central finite differences (h = 1e-4) of the oracle's own FP64 trajectory
(an independent FP64 CPU gradient of the reference's map); and the
50-nucleosome fiber over 200 steps with checkpoints every 20 steps, FP32
forces and the FP64 factor, against central differences of the FP64-force
trajectory and the FP64-force gradient.

The start states are taken a few dozen steps into a trajectory: the
generated fibers have straight linkers, where the Euler angles sit at the
acos(+-1) kink of update_Euler_Angle and their derivatives are unbounded."""
import copy

import numpy as np
import pytest

import oracle_py as O
from brownian_dynamics_cuda_amd import engine as E
from brownian_dynamics_cuda_amd import fiber as Fb
from brownian_dynamics_cuda_amd import system as S
from decks import rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PKEYS = ("q_core", "t_q", "h", "g", "s")


def dense_cc(Q, q, pairs, consts):
    """Core-core charge forces and per-core half energies by dense torch ops."""
    ke, kap, kex, sig = consts
    f = torch.zeros_like(Q)
    e = Q.new_zeros(Q.shape[0])
    for i, j in pairs:
        z = Q[i][:, None, :] - Q[j][None, :, :]
        d = torch.sqrt((z * z).sum(-1))
        qq = q[:, None] * q[None, :]
        ex = torch.exp(-kap * d)
        g = ke * qq * ex * (kap * d + 1) / (d * d)
        En = ke * qq * ex / d
        ev = d <= 8.0
        sd = sig / d
        g = g + torch.where(ev, kex * (12 * sd ** 12 - 6 * sd ** 6) / d, torch.zeros_like(d))
        En = En + torch.where(ev, kex * (sd ** 12 - sd ** 6), torch.zeros_like(d))
        fz = (g / d)[..., None] * z
        f = f.index_add(0, torch.tensor([i], device=Q.device), fz.sum(1)[None])
        f = f.index_add(0, torch.tensor([j], device=Q.device), -fz.sum(0)[None])
        e = e.index_add(0, torch.tensor([i, j], device=Q.device), 0.5 * En.sum().expand(2))
    return f, e


def test_core_core_kernel_and_adjoint():
    from brownian_dynamics_cuda_amd import diff as Df
    fib, st = Fb.make_fiber(6, seed=3)
    ph = fib.phys
    dev = torch.device("cuda")
    r = torch.tensor(st.r, dtype=torch.float64, device=dev).reshape(-1, 3)
    a = torch.tensor(st.a, dtype=torch.float64, device=dev).reshape(-1, 3)
    b = torch.tensor(st.b, dtype=torch.float64, device=dev).reshape(-1, 3)
    c = torch.tensor(st.c, dtype=torch.float64, device=dev).reshape(-1, 3)
    cores = np.nonzero(fib.types == 1)[0]
    p = torch.tensor(fib.core_pos, dtype=torch.float64, device=dev).reshape(-1, 3)
    k = torch.tensor(cores, device=dev)
    Q0 = r[k][:, None] + a[k][:, None] * p[None, :, 0:1] + b[k][:, None] * p[None, :, 1:2] \
        + c[k][:, None] * p[None, :, 2:3]
    q0 = torch.tensor(fib.core_q, dtype=torch.float64, device=dev)
    pairs = [(i, j) for i in range(len(cores)) for j in range(i + 1, len(cores))
             if np.linalg.norm(st.r.reshape(-1, 3)[cores[i]] - st.r.reshape(-1, 3)[cores[j]]) <= 25.0]
    assert len(pairs) >= 3
    nb = [[] for _ in cores]
    for i, j in pairs:
        nb[i].append(j)
        nb[j].append(i)
    off = torch.tensor(np.r_[0, np.cumsum([len(x) for x in nb])], dtype=torch.int32, device=dev)
    nbr = torch.tensor([v for x in nb for v in sorted(x)], dtype=torch.int32, device=dev)
    consts = (ph.k_e, ph.debye, ph.k_ex, ph.sig_cc)
    g = torch.Generator(device="cpu").manual_seed(5)
    gf = torch.randn(Q0.shape, generator=g, dtype=torch.float64).to(dev)
    ge = torch.randn(len(cores), generator=g, dtype=torch.float64).to(dev)
    outs = {}
    for name, fn in (("hip", lambda Q, q: Df.CoreCorePairs.apply(Q, q, off, nbr, consts)),
                     ("dense", lambda Q, q: dense_cc(Q, q, pairs, consts))):
        Q = Q0.clone().requires_grad_(True)
        q = q0.clone().requires_grad_(True)
        f, e = fn(Q, q)
        gQ, gq = torch.autograd.grad((f * gf).sum() + (e * ge).sum(), [Q, q])
        outs[name] = [t.detach().cpu().numpy() for t in (f, e, gQ, gq)]
    for k_, (x, y) in enumerate(zip(outs["hip"], outs["dense"])):
        assert rel_err(x, y) <= 1e-12, k_
    # FP32 kernels against the FP64 ones
    f32, e32 = Df.CoreCorePairs.apply(Q0.float(), q0.float(), off, nbr, consts)
    assert rel_err(f32.double().cpu().numpy(), outs["dense"][0]) <= 1e-5
    assert rel_err(e32.double().cpu().numpy(), outs["dense"][1]) <= 1e-5


def thermalized(fib, st, steps, seed=777):
    """positions and frames `steps` oracle steps into a trajectory, and the
    generator to continue it"""
    orc = O.Oracle(fib)
    fs = orc.prepare(st)
    g = O.Rng(seed)
    nD3 = 3 * (fib.n + fib.n_tail)
    for s in range(steps):
        rr, p = g.step_noise(s, fib.r_var(), nD3)
        orc.step(fs, rr, p)
    st0 = S.State(fs.view("r").copy(), fs.view("a").copy(), fs.view("b").copy(),
                  fs.view("c").copy(), fs.view("r_t").copy(), np.zeros(0))
    return st0, g


def oracle_loss(fib, st0, rr, p, loss):
    o = O.Oracle(fib)
    f = o.prepare(st0)
    for k in range(len(rr)):
        o.step(f, rr[k], p[k])
    if loss == "end_to_end":
        r = f.view("r").reshape(-1, 3)
        return float(((r[-1] - r[0]) ** 2).sum())
    return o.forces(f)["E"]


def perturbed(fib, v, eps):
    f2 = copy.deepcopy(fib)
    f2.core_q = fib.core_q + eps * v["q_core"]
    f2.t_q = fib.t_q + eps * v["t_q"]
    ph = copy.deepcopy(fib.phys)
    ph.h, ph.g, ph.s = fib.phys.h + eps * v["h"], fib.phys.g + eps * v["g"], fib.phys.s + eps * v["s"]
    f2.phys = ph
    return f2


def directions(fib, k, seed):
    rng = np.random.default_rng(seed)
    return [dict(q_core=rng.standard_normal(fib.Nq) / np.sqrt(fib.Nq),
                 t_q=rng.standard_normal(fib.n_tail) / np.sqrt(fib.n_tail),
                 h=rng.standard_normal() * fib.phys.h, g=rng.standard_normal() * fib.phys.g,
                 s=rng.standard_normal() * fib.phys.s) for _ in range(k)]


def test_forward_and_gradient_against_oracle():
    """3-core fiber with tails, 10 steps: the FP64 torch/HIP trajectory
    against the oracle's, and the adjoint gradient of both losses against
    central differences of the oracle's trajectory."""
    from brownian_dynamics_cuda_amd import diff as Df
    print("case: 3-core fiber, 10 steps, FP64 forces, gradient vs oracle FD", flush=True)
    fib, st = Fb.make_fiber(3, seed=11, with_lh=False)
    st0, g = thermalized(fib, st, 30)
    nD3 = 3 * (fib.n + fib.n_tail)
    noise = [g.step_noise(30 + s, fib.r_var(), nD3) for s in range(10)]
    rr, p = [x[0] for x in noise], [x[1] for x in noise]
    m = Df.DiffBD(fib, force_dtype=torch.float64)
    # forward against the oracle
    prm = m.params(False)
    x0 = m.initial_state(st0)
    rrt, pt = m.noise(rr, p)
    xT = m.trajectory(x0, rrt, pt, prm, checkpoint_every=5)
    orc = O.Oracle(fib)
    fs = orc.prepare(st0)
    for k in range(10):
        orc.step(fs, rr[k], p[k])
    for k in ("r", "a", "b", "c", "r_t"):
        assert rel_err(xT[k].detach().cpu().numpy().reshape(-1), fs.view(k)) <= 1e-10, k
    for loss in ("end_to_end", "energy"):
        L, grads = m.gradients(st0, rr, p, loss=loss, checkpoint_every=5)
        assert L == pytest.approx(oracle_loss(fib, st0, rr, p, loss), rel=1e-10)
        for v in directions(fib, 3, 1):
            dd = sum(float(np.sum(grads[k] * v[k])) for k in PKEYS)
            fd = (oracle_loss(perturbed(fib, v, 1e-4), st0, rr, p, loss)
                  - oracle_loss(perturbed(fib, v, -1e-4), st0, rr, p, loss)) / 2e-4
            print(f"{loss}: adjoint {dd:.10e} oracle FD {fd:.10e}", flush=True)
            assert abs(dd - fd) <= 1e-6 * abs(fd)


def test_cfg4_trajectory_gradient():
    """The 50-nucleosome fiber, 200 steps, checkpoints every 20, FP32
    forces and the FP64 factor: the gradient of the squared end-to-end
    distance against central differences (h = 1e-4) of the FP64-force
    trajectory and against the FP64-force gradient."""
    from brownian_dynamics_cuda_amd import diff as Df
    print("case: cfg4, 50 nucleosomes, 200 steps, checkpoints every 20", flush=True)
    fib, st = Fb.make_config(4)
    nD3 = 3 * (fib.n + fib.n_tail)
    # start 40 steps in, from the FP64 product step
    rng = E.HostRng(5489)
    rv = fib.r_var()
    warm = [rng.step_noise(s, rv, nD3) for s in range(40)]
    with E.BDEngine(fib) as eng:
        eng.set_state(st, rebuild=True)
        eng.run_steps(np.array([w[0] for w in warm]), np.array([w[1] for w in warm]))
        st0 = eng.get_state()
    noise = [rng.step_noise(40 + s, rv, nD3) for s in range(200)]
    rr, p = [x[0] for x in noise], [x[1] for x in noise]
    m32 = Df.DiffBD(fib, force_dtype=torch.float32)
    L32, g32 = m32.gradients(st0, rr, p, loss="end_to_end", checkpoint_every=20)
    m64 = Df.DiffBD(fib, force_dtype=torch.float64)
    L64, g64 = m64.gradients(st0, rr, p, loss="end_to_end", checkpoint_every=20)
    assert all(np.all(np.isfinite(g32[k])) for k in PKEYS)
    print(f"L32 {L32:.10e} L64 {L64:.10e}", flush=True)
    assert abs(L32 - L64) <= 1e-5 * abs(L64)
    prm0 = m64.params(False)
    x0 = m64.initial_state(st0)
    rrt, pt = m64.noise(rr, p)

    def loss_at(v, eps):
        pr = {k: prm0[k] + eps * torch.as_tensor(v[k], dtype=torch.float64, device=x0["r"].device)
              for k in PKEYS}
        with torch.no_grad():
            return float(m64.end_to_end2(m64.trajectory(x0, rrt, pt, pr, checkpoint_every=20)))

    for v in directions(fib, 2, 2):
        d64 = sum(float(np.sum(g64[k] * v[k])) for k in PKEYS)
        d32 = sum(float(np.sum(g32[k] * v[k])) for k in PKEYS)
        fd = (loss_at(v, 1e-4) - loss_at(v, -1e-4)) / 2e-4
        print(f"cfg4 end-to-end: adjoint FP64 {d64:.8e} FP32 {d32:.8e} FD {fd:.8e}", flush=True)
        assert abs(d64 - fd) <= 1e-4 * abs(fd)
        assert abs(d32 - d64) <= 1e-2 * abs(d64)
