"""Data generation, training and evaluation of the learned preconditioner.

Training data: SIMP runs (modified SIMP, p = 3, OC, filter r = 1.5) on
cantilevers with a unit downward point load at a seeded random node of the
x = L face and a volume fraction drawn from U[0.2, 0.5].  For each design the
FP64 GMG-PCG solution x* is computed, and the errors of restarted one-step
GMG-PCG iterates e_k = x* - x_k give residual -> error pairs (K e_k, e_k):
the residuals a preconditioner actually sees inside a solve.  Pairs are
consumed online (never stored to disk).

Usage (GPU):
    python -m pyfantom_amd.train_learned --train 96 48 48 --runs 6 --steps 400 \
        --eval 192 96 96 --out learned.npz
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
from pyfantom_amd.learned import ConvNet, LearnedPreconditioner, NetRunner, _Vec, fpcg_solve


def cantilever_problem(nel, seed=0, vf=0.3, load="edge", tol=1e-8, n_level=3, precision="fp64"):
    """Cantilever with the x = 0 face clamped.  load='edge': unit downward load
    spread over the bottom edge of the x = L face; load='random': unit downward
    point load at a seeded random node of the x = L face."""
    mesh = pf.StructuredMesh3D(*nel, *nel, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    K = pf.StructuredStiffnessKernel(mesh)
    solver = pf.MultiGrid(mesh, K, maxiter=1000, tol=tol, n_smooth=2, omega=0.6, n_level=n_level, cycle="V",
                          precision=precision)
    FE = pf.FiniteElement(mesh, K, solver)
    nx, ny, nz = (n + 1 for n in nel)
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    FE.add_dirichlet_boundary_condition(node_ids=(k + nz * nx * j).ravel(), rhs=0)
    if load == "edge":
        FE.add_point_forces(forces=np.array([[0.0, -1.0 / nz, 0.0]]), node_ids=np.arange(nz) + nz * (nx - 1))
    else:
        rng = np.random.default_rng(seed)
        jj, kk = int(rng.integers(0, ny)), int(rng.integers(0, nz))
        FE.add_point_forces(forces=np.array([[0.0, -1.0, 0.0]]), node_ids=np.array([kk + nz * (nx - 1) + nz * nx * jj]))
    filt = pf.StructuredFilter3D(mesh, 1.5)
    prob = pf.MinimumCompliance(FE, filt, E_mul=[1.0], void=1e-9, penalty=3.0, volume_fraction=[vf],
                                heavyside=False, law="modified")
    return mesh, K, solver, FE, filt, prob


def design_pairs(K, solver, FE, n_pairs=8):
    """(r, e) pairs from restarted one-step GMG-PCG iterates on the current E."""
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
        pairs.append((r, e))
        solver.last_x0 = x.clone()
        x, _ = solver.solve(FE.rhs, use_last=True)
        x = x.clone()
    solver.maxiter, solver.tol = old
    return pairs


def train(nel, runs, steps, simp_iters=6, pairs_per_design=6, lr=2e-3, seed=0, log=print):
    net = ConvNet(seed=seed)
    rng = np.random.default_rng(seed)
    t0 = time.time()
    losses = []
    step = 0
    data = []
    for run in range(runs):
        vf = float(rng.uniform(0.2, 0.5))
        mesh, K, solver, FE, filt, prob = cantilever_problem(nel, seed=seed * 1000 + run, vf=vf, load="random")
        opt = pf.OC(prob, move=0.2)
        runner = NetRunner(net, K.grid)
        for it in range(simp_iters):
            opt.iter()
            runner.set_density(prob.get_desvars())
            for r, e in design_pairs(K, solver, FE, pairs_per_design):
                data.append((runner, r, e))
        log(f"run {run}: vf {vf:.3f} c {prob.f():.4e} pairs {len(data)} ({time.time() - t0:.1f} s)")
        # train on everything gathered so far, newest first
        per_run = max(1, steps // runs)
        for s in range(per_run):
            rn, r, e = data[int(rng.integers(0, len(data)))]
            loss = float(rn.loss_and_grad(r, e).item())
            rn.adam_step(lr=lr)
            losses.append(loss)
            step += 1
        log(f"  step {step}: loss {np.mean(losses[-per_run:]):.5f}")
    return net, losses


def evaluate(net, nel, n_level=4, simp_iters=10, log=print):
    """Plain FP64 GMG-PCG vs the mixed-precision solve vs flexible PCG with the
    learned preconditioner (alone and after a V-cycle), on a mid-optimisation
    design of an edge-loaded cantilever, all to 1e-8."""
    mesh, K, solver, FE, filt, prob = cantilever_problem(nel, vf=0.3, load="edge", n_level=n_level)
    opt = pf.OC(prob, move=0.2)
    for _ in range(simp_iters):
        opt.iter()
    x_design = prob.get_desvars()
    E = K.rho.clone()
    res = {"grid": list(nel), "simp_iters": simp_iters}

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t

    plain = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=n_level, cycle="V")
    K.set_rho(E)
    plain.solve(FE.rhs, use_last=False)  # warm-up (graph capture, setup)
    (u0, r0), t_plain = timed(lambda: plain.solve(FE.rhs, use_last=False))
    u0 = u0.clone()
    c0 = float(torch.dot(FE.rhs, u0))
    res["plain_fp64"] = {"iterations": plain.last_iterations, "seconds": t_plain, "residual": r0}
    mx = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=n_level, cycle="V",
                      precision="mixed", inner_tol=1e-4)
    mx.solve(FE.rhs, use_last=False)
    (um, rm), t_mx = timed(lambda: mx.solve(FE.rhs, use_last=False))
    cm = float(torch.dot(FE.rhs, um))
    res["mixed"] = {"outer": mx.last_outer_iterations, "inner_iterations": mx.last_iterations, "seconds": t_mx,
                    "residual": rm, "compliance_rel_err": abs(cm - c0) / abs(c0),
                    "u_rel_err": float(torch.linalg.norm(um - u0) / torch.linalg.norm(u0))}
    mg = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=n_level, cycle="V")
    for mode in ("hybrid", "learned"):
        M = LearnedPreconditioner(net, K, mg=mg, mode=mode)
        M.set_density(x_design)
        fb = pf.MultiGrid(mesh, K, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=n_level, cycle="V")
        fpcg_solve(K, FE.rhs, M, tol=1e-8, maxiter=5, fallback=None)  # warm-up
        (u, info), t = timed(lambda: fpcg_solve(K, FE.rhs, M, tol=1e-8, maxiter=400, fallback=fb))
        c = float(torch.dot(FE.rhs, u))
        res[f"fpcg_{mode}"] = {"iterations": info["iterations"], "fallback": info["fallback"],
                               "fallback_iterations": info["fallback_iterations"], "seconds": t,
                               "residual": info["residual"], "compliance_rel_err": abs(c - c0) / abs(c0)}
    # the network alone without GMG: flexible PCG with z = N(r) and no fallback
    M = LearnedPreconditioner(net, K, mode="learned")
    M.set_density(x_design)
    (u, info), t = timed(lambda: fpcg_solve(K, FE.rhs, M, tol=1e-8, maxiter=400, fallback=None))
    res["fpcg_learned_nofallback"] = {"iterations": info["iterations"], "seconds": t, "residual": info["residual"]}
    log(json.dumps(res))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, nargs=3, default=[96, 48, 48])
    ap.add_argument("--runs", type=int, default=6)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--eval", type=int, nargs=3, default=[192, 96, 96])
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    log = lambda m: print(m, flush=True)
    net, losses = train(tuple(args.train), args.runs, args.steps, log=log)
    if args.out:
        net.save(args.out, losses=np.array(losses))
    evaluate(net, tuple(args.eval), log=log)


if __name__ == "__main__":
    sys.exit(main())
