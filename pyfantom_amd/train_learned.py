"""Data generation, training and evaluation of the learned preconditioner.

Training data: SIMP runs (modified SIMP, p = 3, OC, filter r = 1.5) on
96 x 48 x 48 cantilevers with a unit downward point load at a seeded random
node of the x = L face and a volume fraction drawn from U[0.2, 0.5].  For
each design the FP64 GMG-PCG solution x* is computed to 1e-10, and restarted
one-step GMG-PCG iterates x_k give errors e_k = x* - x_k.  The network is
used after a V-cycle (LearnedPreconditioner, mode 'hybrid'), so each error is
turned into the pair the network sees there: r = K e_k, z = V(r) (the same
FP32 V-cycle as the solver), input t = r - K z, target e_k - z.  The density
channel comes from the moduli (density_from_moduli), as in LearnedSolver.
Pairs are consumed online, one Adam step each (never stored to disk).

The loss ('energy', default) is the fraction of the error energy e.Ke that
the Galerkin-scaled correction leaves, 1 - (t.c)^2 / ((c.Kc)(t.e)): what the
hybrid preconditioner gains from one application ('mse' is the Euclidean
error of the normalised prediction, dominated by the void regions).

Usage (GPU):
    python -m pyfantom_amd.train_learned --train 96 48 48 --pairs 20000 \
        --eval 192 96 96 --out pyfantom_amd/data/learned_precond_96x48x48.npz
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

import pyfantom_amd as pf
from pyfantom_amd import workloads
from pyfantom_amd.learned import (ConvNet, LearnedPreconditioner, LearnedSolver, NetRunner, _Vec,
                                  density_from_moduli, fpcg_solve)


def cantilever_problem(nel, seed=0, vf=0.3, load="edge", tol=1e-8, n_level=3, precision="fp64"):
    """Cantilever with the x = 0 face clamped.  load='edge': unit downward load
    spread over the bottom edge of the x = L face; load='random': unit downward
    point load at a seeded random node of the x = L face."""
    return workloads.cantilever(nel, vf=vf, r=1.5, load=load, seed=seed, maxiter=1000, tol=tol, n_smooth=2,
                                omega=0.6, n_level=n_level, cycle="V", precision=precision, omega_schedule="fixed")


def design_pairs(K, solver, FE, mg, n_pairs=8, mode="hybrid"):
    """Pairs on the current E from the errors e_k of restarted one-step
    GMG-PCG iterates, r = K e_k (free DOFs only).  mode 'learned': (r, e_k),
    what the network alone sees; mode 'hybrid': z = V(r) from `mg`,
    (t, e_k - z) with t = r - K z, what it sees after a V-cycle."""
    vec = _Vec(K)
    free = (~K.constraints).to(torch.float64)
    old = (solver.maxiter, solver.tol)
    solver.tol = 1e-10
    ustar, _ = solver.solve(FE.rhs, use_last=False)
    ustar = ustar.clone()
    x = torch.zeros_like(ustar)
    pairs = []
    solver.maxiter = 1
    for _ in range(n_pairs):
        e = (ustar - x) * free
        r = torch.empty_like(e)
        vec.ku(e, r)
        r.mul_(free)
        if mode == "learned":
            pairs.append((r, e))
        else:
            z = torch.empty_like(e)
            mg.precondition_native(r, z)
            z.mul_(free)
            kz = torch.empty_like(e)
            vec.ku(z, kz)
            pairs.append(((r - kz) * free, e - z))
        solver.last_x0 = x.clone()
        x, _ = solver.solve(FE.rhs, use_last=True)
        x = x.clone()
    solver.maxiter, solver.tol = old
    return pairs


def trajectory_pairs(K, solver, FE, mg, net, n_pairs=8):
    """On-policy pairs: the iterates x_k of flexible PCG preconditioned by the
    V-cycle plus the current network (hybrid mode) from x = 0; for each,
    e_k = x* - x_k, r = K e_k, z = V(r), pair (r - K z, e_k - z)."""
    vec = _Vec(K)
    free = (~K.constraints).to(torch.float64)
    old = solver.tol
    solver.tol = 1e-10
    ustar, _ = solver.solve(FE.rhs, use_last=False)
    ustar = ustar.clone()
    solver.tol = old
    M = LearnedPreconditioner(net, K, mg=mg, mode="hybrid")
    M.set_density(density_from_moduli(K.rho64))
    xs = []
    fpcg_solve(K, FE.rhs, M, tol=1e-30, maxiter=n_pairs, fallback=None, stall_window=10 ** 6,
               record=lambda x: xs.append(x.clone()))
    pairs = []
    for x in [torch.zeros_like(ustar)] + xs[:-1]:
        e = (ustar - x) * free
        r = torch.empty_like(e)
        vec.ku(e, r)
        r.mul_(free)
        z = torch.empty_like(e)
        mg.precondition_native(r, z)
        z.mul_(free)
        kz = torch.empty_like(e)
        vec.ku(z, kz)
        pairs.append(((r - kz) * free, e - z))
    return pairs


def train(nel, target_pairs, simp_iters=10, pairs_per_design=20, lr=5e-4, seed=0, n_level=4, log=print,
          max_seconds=None, loss="energy", mode="hybrid"):
    """Online training until target_pairs (t, target) pairs have been used
    (or max_seconds have passed, checked after each run)."""
    net = ConvNet(seed=seed)
    rng = np.random.default_rng(seed)
    t0 = time.time()
    losses = []
    used = 0
    run = 0
    while used < target_pairs:
        vf = float(rng.uniform(0.2, 0.5))
        mesh, K, solver, FE, filt, prob = cantilever_problem(nel, seed=seed * 1000 + run, vf=vf, load="random",
                                                             n_level=n_level)
        mg = pf.MultiGrid(mesh, K, maxiter=1, tol=1e-8, n_smooth=2, omega=0.6, n_level=n_level, cycle="V",
                          omega_schedule="fixed", precision="mixed", fallback=False)
        opt = pf.OC(prob, move=0.2)
        runner = NetRunner(net, K.grid)
        vec = _Vec(K)
        free = (~K.constraints).to(torch.float64)
        for it in range(simp_iters):
            opt.iter()
            runner.set_density(density_from_moduli(K.rho64))
            pairs = (trajectory_pairs(K, solver, FE, mg, net, pairs_per_design) if mode == "trajectory" else
                     design_pairs(K, solver, FE, mg, pairs_per_design, mode=mode))
            for t, e in pairs:
                if loss == "energy":
                    lv = runner.energy_loss_and_grad(t, e, vec, free)
                    if not math.isfinite(lv):
                        continue
                else:
                    lv = float(runner.loss_and_grad(t, e).item())
                losses.append(lv)
                runner.adam_step(lr=lr)
                used += 1
        log(f"run {run}: vf {vf:.3f} c {prob.f():.4e} pairs {used} loss {np.mean(losses[-pairs_per_design * simp_iters:]):.5f} "
            f"({time.time() - t0:.1f} s)")
        run += 1
        del mesh, K, solver, FE, filt, prob, mg, opt, runner
        if max_seconds is not None and time.time() - t0 > max_seconds:
            log(f"stopping after {time.time() - t0:.0f} s with {used} pairs")
            break
    return net, losses, run


def evaluate(net, nel, n_level=4, simp_iters=10, log=print):
    """On a mid-optimisation design of an edge-loaded cantilever, all to 1e-8:
    plain FP64 GMG-PCG, the mixed-precision GMG-PCG (FP32 V-cycle), and
    flexible PCG with the FP32 V-cycle plus the learned correction
    (LearnedSolver) for `net` and for an untrained network; iterations,
    seconds and compliance against the FP64 solve."""
    mesh, K, solver, FE, filt, prob = cantilever_problem(nel, vf=0.3, load="edge", n_level=n_level)
    opt = pf.OC(prob, move=0.2)
    for _ in range(simp_iters):
        opt.iter()
    E = K.rho64.clone()
    res = {"grid": list(nel), "simp_iters": simp_iters}

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t

    opts = dict(n_smooth=2, omega=0.6, n_level=n_level, cycle="V", omega_schedule="fixed")
    plain = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, **opts)
    K.set_rho(E)
    plain.solve(FE.rhs, use_last=False)  # warm-up (graph capture, setup)
    (u0, r0), t_plain = timed(lambda: plain.solve(FE.rhs, use_last=False))
    u0 = u0.clone()
    c0 = float(torch.dot(FE.rhs, u0))
    res["plain_fp64"] = {"iterations": plain.last_iterations, "seconds": t_plain, "residual": r0}
    mx = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, precision="mixed", **opts)
    mx.solve(FE.rhs, use_last=False)
    (um, rm), t_mx = timed(lambda: mx.solve(FE.rhs, use_last=False))
    cm = float(torch.dot(FE.rhs, um))
    res["mixed"] = {"iterations": mx.last_iterations, "seconds": t_mx, "residual": rm,
                    "compliance_rel_err": abs(cm - c0) / abs(c0)}
    for label, nn in (("learned_trained", net), ("learned_untrained", ConvNet(seed=12345))):
        ls = LearnedSolver(mesh, K, nn, maxiter=400, tol=1e-8, precision="mixed", **{
            k: v for k, v in opts.items() if k != "cycle"})
        ls.solve(FE.rhs, use_last=False)  # warm-up
        (u, r), t = timed(lambda: ls.solve(FE.rhs, use_last=False))
        c = float(torch.dot(FE.rhs, u))
        res[label] = {"iterations": ls.last_iterations, "fallback": ls.last_fallback,
                      "fallback_iterations": ls.last_fallback_iterations, "seconds": t, "residual": r,
                      "compliance_rel_err": abs(c - c0) / abs(c0)}
    log(json.dumps(res))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, nargs=3, default=[96, 48, 48])
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--simp-iters", type=int, default=10)
    ap.add_argument("--pairs-per-design", type=int, default=20)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--eval", type=int, nargs=3, default=[192, 96, 96])
    ap.add_argument("--out", default=None)
    ap.add_argument("--max-seconds", type=float, default=None)
    ap.add_argument("--loss", default="energy", choices=["energy", "mse"])
    ap.add_argument("--mode", default="hybrid", choices=["hybrid", "learned", "trajectory"],
                    help="pairs: post-V-cycle (hybrid), raw (learned: the network alone) or on-policy FPCG iterates")
    args = ap.parse_args()
    log = lambda m: print(m, flush=True)
    net, losses, runs = train(tuple(args.train), args.pairs, args.simp_iters, args.pairs_per_design, lr=args.lr,
                              log=log, max_seconds=args.max_seconds, loss=args.loss, mode=args.mode)
    if args.out:
        net.save(args.out, losses=np.array(losses, dtype=np.float32), pairs=np.array(len(losses)),
                 runs=np.array(runs), grid=np.array(args.train), lr=np.array(args.lr),
                 loss_kind=np.array(args.loss), pair_mode=np.array(args.mode))
    evaluate(net, tuple(args.eval), log=log)


if __name__ == "__main__":
    sys.exit(main())
