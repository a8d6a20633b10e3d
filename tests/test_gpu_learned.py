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


def _mid_design(nel, n_level=4, simp_iters=8):
    """An edge-loaded cantilever after a few OC iterations (held out: the
    network was trained on point loads at 96x48x48)."""
    import pyfantom_amd as pf
    from pyfantom_amd.train_learned import cantilever_problem
    mesh, K, solver, FE, filt, prob = cantilever_problem(nel, vf=0.3, load="edge", n_level=n_level)
    opt = pf.OC(prob, move=0.2)
    for _ in range(simp_iters):
        opt.iter()
    return mesh, K, solver, FE


def _committed_net():
    import os
    import pyfantom_amd as pf
    from pyfantom_amd.learned import ConvNet
    path = os.path.join(os.path.dirname(pf.__file__), "data", "learned_precond_96x48x48.npz")
    with np.load(path, allow_pickle=False) as z:
        meta = {k: z[k] for k in ("pairs", "grid", "loss_kind", "pair_mode") if k in z.files}
    return ConvNet.from_file(path), meta


def test_trained_network_gains_energy(oracle):
    """What training optimised, on held-out data: on post-V-cycle pairs of an
    edge-loaded 64x32x32 design (the network saw point-loaded 96x48x48 runs),
    one application of the committed network removes more of the remaining
    error energy than an untrained network of the same architecture."""
    import torch
    import pyfantom_amd as pf
    from pyfantom_amd.learned import ConvNet, NetRunner, _Vec, density_from_moduli
    from pyfantom_amd.train_learned import design_pairs
    _torch()
    net, meta = _committed_net()
    print(f"trained network energy gain (trained on {int(meta['pairs'])} {meta['pair_mode']} pairs, "
          f"{meta['loss_kind']} loss, grid {list(meta['grid'])})", flush=True)
    assert int(meta["pairs"]) >= 20000
    mesh, K, solver, FE = _mid_design((64, 32, 32))
    mg = pf.MultiGrid(mesh, K, maxiter=1, tol=1e-8, n_smooth=2, omega=0.6, n_level=4, cycle="V",
                      omega_schedule="fixed", precision="mixed", fallback=False)
    pairs = design_pairs(K, solver, FE, mg, n_pairs=12)
    vec = _Vec(K)
    free = (~K.constraints).to(torch.float64)
    out = {}
    for label, nn in (("trained", net), ("untrained", ConvNet(seed=12345))):
        run = NetRunner(nn, K.grid)
        run.set_density(density_from_moduli(K.rho))
        ls = [run.energy_loss_and_grad(t, e, vec, free) for t, e in pairs]
        out[label] = float(np.mean(ls))
    print(f"  energy left after one correction: trained {out['trained']:.3f}, untrained {out['untrained']:.3f}",
          flush=True)
    assert out["trained"] < out["untrained"]


def test_trained_network_ab(oracle):
    """A/B on the held-out 64x32x32 design, all to 1e-8: plain FP64 GMG-PCG,
    and flexible PCG with the FP32 V-cycle plus the learned correction for the
    committed and for an untrained network (LearnedSolver).  Checks that every
    variant reaches the FP64 solution; prints iterations and times, the gain
    (or loss) of the learned part."""
    import time
    import torch
    import pyfantom_amd as pf
    from pyfantom_amd.learned import ConvNet, LearnedSolver
    _torch()
    print("trained network A/B", flush=True)
    net, _ = _committed_net()
    mesh, K, solver, FE = _mid_design((64, 32, 32))
    plain = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=4, cycle="V",
                         omega_schedule="fixed")
    plain.solve(FE.rhs, use_last=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u0, r0 = plain.solve(FE.rhs, use_last=False)
    torch.cuda.synchronize()
    results = {"plain": (plain.last_iterations, time.perf_counter() - t0)}
    u0 = u0.clone()
    c0 = float(torch.dot(FE.rhs, u0))
    for label, nn in (("trained", net), ("untrained", ConvNet(seed=12345))):
        ls = LearnedSolver(mesh, K, nn, maxiter=400, tol=1e-8, precision="mixed", n_level=4)
        ls.solve(FE.rhs, use_last=False)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, res = ls.solve(FE.rhs, use_last=False)
        torch.cuda.synchronize()
        c = float(torch.dot(FE.rhs, u))
        results[label] = (ls.last_iterations + ls.last_fallback_iterations, time.perf_counter() - t0)
        assert res < 1.5e-8
        assert abs(c - c0) <= 1e-6 * abs(c0)
    print("A/B (iterations, seconds): " + ", ".join(f"{k} {v[0]} / {v[1] * 1e3:.1f} ms" for k, v in results.items()),
          flush=True)
    test_trained_network_ab.results = results


@pytest.mark.xfail(strict=False, reason="measured: the committed network wins its training objective (energy left "
                                        "after one correction) but not the PCG iteration count -- 51-78 trained "
                                        "against 48 untrained and 44 plain at 64x32x32, 52 / 47 / 46 at "
                                        "192x96x96: a stronger nonlinear correction disturbs the conjugacy of the "
                                        "Krylov recurrence more")
def test_trained_network_fewer_iterations(oracle):
    """The claim the A/B is for: the trained network takes fewer flexible-PCG
    iterations than the untrained one (held-out 64x32x32 design)."""
    res = getattr(test_trained_network_ab, "results", None)
    if res is None:
        test_trained_network_ab(oracle)
        res = test_trained_network_ab.results
    assert res["trained"][0] < res["untrained"][0]


def test_c4_learned_solver(oracle):
    """192x96x96 (5.4M DOF) edge-loaded cantilever after 10 OC iterations:
    LearnedSolver (FP32 V-cycle + learned correction in flexible PCG, FP64
    GMG-PCG fallback) reaches 1e-8 with compliance within 1e-6 of the plain
    FP64 GMG-PCG solve; iterations and times of plain, mixed and learned are
    printed."""
    import time
    import torch
    import pyfantom_amd as pf
    from pyfantom_amd.learned import LearnedSolver
    _torch()
    print("C4 learned solver 192x96x96", flush=True)
    net, _ = _committed_net()
    mesh, K, solver, FE = _mid_design((192, 96, 96), simp_iters=10)
    out = {}
    for label in ("plain", "mixed", "learned"):
        if label == "learned":
            s = LearnedSolver(mesh, K, net, maxiter=400, tol=1e-8, precision="mixed", n_level=4)
        else:
            s = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=4, cycle="V",
                             omega_schedule="fixed", precision="fp64" if label == "plain" else "mixed")
        s.solve(FE.rhs, use_last=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, res = s.solve(FE.rhs, use_last=False)
        torch.cuda.synchronize()
        out[label] = (s.last_iterations, time.perf_counter() - t0, float(torch.dot(FE.rhs, u)), res,
                      getattr(s, "last_fallback", False))
        assert res < 1.5e-8, label
    c0 = out["plain"][2]
    for label, (it, t, c, res, fb) in out.items():
        print(f"  {label}: {it} iterations, {t * 1e3:.1f} ms, residual {res:.2e}, fallback {fb}, "
              f"|c - c_fp64|/c {abs(c - c0) / abs(c0):.2e}", flush=True)
        assert abs(c - c0) <= 1e-6 * abs(c0), label
