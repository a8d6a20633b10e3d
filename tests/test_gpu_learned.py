"""GPU tests of the learned preconditioner: the network's forward and
backward passes (BF16 MFMA kernels) against the float64 numpy network with
BF16 rounding at the same points, training on the GPU, and flexible PCG with
the trained network inside (with GMG-PCG fallback)."""
import numpy as np
import pytest

from oracle import conv_oracle as C
from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _from_bf16(t, shape):
    import torch
    return t.view(torch.bfloat16).float().cpu().numpy().astype(np.float64).reshape(shape)


def test_network_vs_oracle():
    import torch
    from pyfantom_amd.learned import ConvNet, NetRunner
    _torch()
    print("network vs oracle", flush=True)
    grid = (16, 8, 8)
    net = ConvNet(seed=3)
    # non-trivial biases
    g = torch.Generator().manual_seed(4)
    for name in net.b:
        net.b[name] = (torch.randn(net.b[name].shape, generator=g) * 0.05).cuda()
    net.b["out"][3:] = 0.0
    run = NetRunner(net, grid)
    D, H, W = run.shape[0]
    n = D * H * W * 3
    r = torch.from_numpy(np.random.default_rng(0).standard_normal(n)).cuda()
    x_el = torch.from_numpy(np.random.default_rng(1).uniform(0, 1, int(np.prod(grid)))).cuda()
    run.set_density(x_el)
    y = run.forward(r).cpu().numpy().astype(np.float64).reshape(D, H, W, 16)
    Wn = {k: C.bf16_round(v.cpu().numpy().astype(np.float64)) for k, v in net.w.items()}
    Bn = {k: v.cpu().numpy().astype(np.float64) for k, v in net.b.items()}
    x = _from_bf16(run.act["x"], (D, H, W, 16))
    # the input encodes r / rms(r) and the nodal density
    rr = r.cpu().numpy().reshape(D, H, W, 3)
    assert np.abs(x[..., :3] - C.bf16_round(rr / np.sqrt(np.mean(rr ** 2)))).max() < 1e-6
    yo, cache = U.forward(x, Wn, Bn, rounding=True)
    err = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print(f"forward rel err {err:.2e}", flush=True)
    assert err < 1e-2
    t = C.bf16_round(np.random.default_rng(2).standard_normal((D, H, W, 16)))
    t[..., 3:] = 0
    tdev = torch.from_numpy(t.astype(np.float32).reshape(-1)).cuda().to(torch.bfloat16).view(torch.int16)
    gw, gb = run.backward(tdev)
    gWo, gBo = U.backward(t, Wn, cache, rounding=True)
    for name in Wn:
        a = gw[name].cpu().numpy().astype(np.float64)
        b = gWo[name]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = np.sum(a * b) / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print(f"grad {name}: rel {rel:.2e} cos {cos:.6f}", flush=True)
        assert rel < 3e-2 and cos > 0.999, name
        ab = gb[name].cpu().numpy().astype(np.float64)
        assert np.linalg.norm(ab - gBo[name]) <= 3e-2 * max(np.linalg.norm(gBo[name]), 1e-30)


def test_training_and_fpcg(oracle):
    """Train briefly on (K e, e) pairs of a 32x16x16 cantilever and use the
    network in flexible PCG: it must converge to the FP64 GMG solution, alone
    (through the fallback if it stalls) and as a correction after a V-cycle."""
    import torch
    import pyfantom_amd as pf
    from pyfantom_amd.learned import ConvNet, NetRunner, LearnedPreconditioner, fpcg_solve, _Vec
    from tests.test_gpu_solver import cantilever
    _torch()
    print("training + fpcg", flush=True)
    nel = (32, 16, 16)
    mesh, K, solver, FE = cantilever(nel, tol=1e-8, n_level=3)
    rng = np.random.default_rng(0)
    xd = rng.uniform(0.01, 1, mesh.n_elements)
    from tests.test_gpu_solver import filtered_field
    E = torch.from_numpy(filtered_field(oracle, nel)).cuda()
    K.set_rho(E)
    net = ConvNet(seed=0)
    run = NetRunner(net, K.grid)
    run.set_density(torch.from_numpy(xd).cuda())
    vec = _Vec(K)
    free = (~K.constraints).to(torch.float64)
    # errors from restarted GMG iterates: e = x* - x_k, r = K e
    ustar, _ = solver.solve(FE.rhs, use_last=False)
    ustar = ustar.clone()
    pairs = []
    x = torch.zeros_like(ustar)
    for k in range(6):
        e = (ustar - x) * free
        r = torch.empty_like(e)
        vec.ku(e, r)
        pairs.append((r.clone(), e.clone()))
        solver.last_x0 = x.clone()
        solver.maxiter = 1
        x, _ = solver.solve(FE.rhs, use_last=True)
        x = x.clone()
    solver.maxiter = 1000
    losses = []
    for step in range(80):
        r, e = pairs[step % len(pairs)]
        loss = float(run.loss_and_grad(r, e).item())
        run.adam_step(lr=2e-3)
        losses.append(loss)
    first, last = np.mean(losses[:6]), np.mean(losses[-6:])
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < first
    mg = pf.MultiGrid(mesh, K, maxiter=500, tol=1e-8, n_smooth=2, omega=0.6, n_level=3, cycle="V")
    for mode in ("hybrid", "learned"):
        M = LearnedPreconditioner(net, K, mg=mg, mode=mode)
        M.set_density(torch.from_numpy(xd).cuda())
        fb = pf.MultiGrid(mesh, K, maxiter=500, tol=1e-8, n_smooth=2, omega=0.6, n_level=3, cycle="V")
        u, info = fpcg_solve(K, FE.rhs, M, tol=1e-8, maxiter=300, fallback=fb)
        c = float(torch.dot(FE.rhs, u))
        c0 = float(torch.dot(FE.rhs, ustar))
        print(f"{mode}: iterations {info['iterations']} fallback {info['fallback']} "
              f"({info['fallback_iterations']}) residual {info['residual']:.2e} |dc|/c {abs(c - c0) / abs(c0):.2e}",
              flush=True)
        assert info["residual"] < 1e-8 * 1.5
        assert abs(c - c0) <= 1e-6 * abs(c0)


def test_trained_network_ab(oracle):
    """A/B of the committed trained network (pyfantom_amd/data, trained on the
    GPU by pyfantom_amd/train_learned.py on 96x48x48 SIMP runs) on a held-out
    problem: a 64x32x32 cantilever after 8 OC iterations.  Compares flexible
    PCG with the V-cycle + learned correction against the same with an
    untrained network and against plain GMG-PCG, all to 1e-8; prints the
    iteration counts and times (the gain, or loss, of the learned part) and
    checks that every variant reaches the FP64 solution."""
    import os
    import time
    import torch
    import pyfantom_amd as pf
    from pyfantom_amd.learned import ConvNet, LearnedPreconditioner, fpcg_solve
    from pyfantom_amd.train_learned import cantilever_problem
    _torch()
    path = os.path.join(os.path.dirname(pf.__file__), "data", "learned_precond_96x48x48.npz")
    print("trained network A/B", flush=True)
    nel = (64, 32, 32)
    mesh, K, solver, FE, filt, prob = cantilever_problem(nel, vf=0.3, load="edge", n_level=4)
    opt = pf.OC(prob, move=0.2)
    for _ in range(8):
        opt.iter()
    x = prob.get_desvars()
    plain = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=4, cycle="V")
    plain.solve(FE.rhs, use_last=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u0, r0 = plain.solve(FE.rhs, use_last=False)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    u0 = u0.clone()
    c0 = float(torch.dot(FE.rhs, u0))
    results = {"plain": (plain.last_iterations, t_plain)}
    for label, net in (("trained", ConvNet.from_file(path)), ("untrained", ConvNet(seed=0))):
        mg = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=4, cycle="V")
        fb = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=4, cycle="V")
        M = LearnedPreconditioner(net, K, mg=mg, mode="hybrid")
        M.set_density(x)
        fpcg_solve(K, FE.rhs, M, tol=1e-8, maxiter=3, fallback=None)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, info = fpcg_solve(K, FE.rhs, M, tol=1e-8, maxiter=300, fallback=fb)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        c = float(torch.dot(FE.rhs, u))
        results[label] = (info["iterations"] + info["fallback_iterations"], t)
        assert info["residual"] < 1.5e-8
        assert abs(c - c0) <= 1e-6 * abs(c0)
    print("A/B (iterations, seconds): " + ", ".join(f"{k} {v[0]} / {v[1] * 1e3:.1f} ms" for k, v in results.items()),
          flush=True)
    print(f"trained vs untrained iterations: {results['trained'][0]} vs {results['untrained'][0]}; "
          f"trained vs plain GMG-PCG time ratio {results['trained'][1] / results['plain'][1]:.2f}", flush=True)
