"""Benchmark of the SIMP finite-element solve path on one MI355X.

Workload (default): 3D cantilever, 256 x 128 x 128 hex8 (12,830,211 DOF,
FP64), x = 0 face clamped, unit downward load along the bottom edge of the
x = L face, modified SIMP (E0 = 1, Emin = 1e-9, p = 3), volume fraction 0.3,
density filter r = 1.5, OC (move 0.2).  Solver: GMG-PCG to 1e-8 relative
residual, 5 coarsenings (coarsest 8 x 4 x 4 elements), 2 pre + 2 post Jacobi
sweeps with fixed omega = 0.6, warm start.  --config c2 runs the same at
128 x 64 x 64 (4 coarsenings).  A step is one SIMP iteration: filter ->
interpolation -> solve -> compliance -> sensitivity -> chain rule -> filter
adjoint -> OC update.  The same steps are then run with the mixed-precision
solver (FP64 PCG, FP32 V-cycle).

Also measured: the isolated K.u of the dominant kernel (seeded density
x ~ U[0.01, 1], default_rng(0)) timed with HIP events on the stream it runs
on, its roofline fraction, and the FP64 CPU oracle on the host cores.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
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

HBM_PEAK_GBS = 8000.0
METRIC = "s/SIMP iter to 1e-8 rel res, mixed+learned vs FP64 GMG-PCG; GDOF/s K·u"

CONFIGS = {
    "c2": dict(nel=(128, 64, 64), n_level=4, name="cantilever_128x64x64"),
    "c3": dict(nel=(256, 128, 128), n_level=4, name="cantilever_256x128x128"),
    "small": dict(nel=(32, 16, 16), n_level=3, name="cantilever_32x16x16"),
    # combined run: bridge, MMA, filter r = 2, volume fraction 0.15
    "c5": dict(nel=(384, 192, 192), n_level=6, name="bridge_384x192x192", kind="bridge", optimizer="mma",
               vf=0.15, r=2.0, omega=0.5),
    "bridge_small": dict(nel=(48, 24, 24), n_level=3, name="bridge_48x24x24", kind="bridge", optimizer="mma",
                         vf=0.15, r=2.0, omega=0.5),
}


# SIMP legs: name -> MultiGrid options.  "fp64" is the plain FP64 GMG-PCG of
# the specification (fixed-weight Jacobi); the others change only the solver.
LEGS = {
    "fp64": dict(precision="fp64", smoother="jacobi"),
    "mixed": dict(precision="mixed", smoother="jacobi"),
    "fp64_cheb": dict(precision="fp64", smoother="chebyshev"),
    "mixed_cheb": dict(precision="mixed", smoother="chebyshev"),
    # FP32 V-cycle under flexible (Polak-Ribiere) PCG
    "mixed_flex": dict(precision="mixed", smoother="jacobi", flexible=True),
    # V-cycle fine and level-1 sweeps on the matrix cores (BF16 operands,
    # split node values), flexible PCG
    "bf16": dict(precision="bf16", smoother="jacobi"),
    # flexible PCG: FP32 V-cycle + the committed network's correction, FP64 GMG-PCG fallback
    "mixed_learned": dict(precision="mixed", smoother="jacobi", learned=True),
}


def build_problem(pf, nel, n_level, tol=1e-8, precision="fp64", kind="cantilever", vf=0.3, r=1.5, omega=0.6,
                  omega_schedule="reference", smoother="jacobi", learned=False, flexible=None):
    """cantilever: x = 0 face clamped, total load -1 in y over the nodes of
    the bottom edge (y = 0) of the x = L face.  bridge: the four bottom
    corner nodes (y = 0, x in {0, L}, z in {0, W}) pinned in all DOFs, total
    load -1 in y spread evenly over the nodes of the top surface y = H.
    learned: the solver is LearnedSolver with the committed network."""
    from pyfantom_amd import workloads
    build = workloads.cantilever if kind == "cantilever" else workloads.bridge
    mesh, K, solver, FE, filt, prob = build(nel, vf=vf, r=r, maxiter=1000, tol=tol, n_smooth=2, omega=omega,
                                            n_level=n_level, cycle="V", precision=precision,
                                            omega_schedule=omega_schedule, smoother=smoother, flexible=flexible)
    if learned:
        from pyfantom_amd.learned import ConvNet, LearnedSolver
        net = ConvNet.from_file(os.path.join(ROOT, "pyfantom_amd", "data", "learned_precond_96x48x48.npz"))
        solver = LearnedSolver(mesh, K, net, maxiter=1000, tol=tol, precision=precision, n_level=n_level,
                               n_smooth=2, omega=omega, smoother=smoother, omega_schedule=omega_schedule)
        FE.solver = solver   # the optimiser's first evaluation already uses it
    return mesh, K, solver, FE, filt, prob


def time_ku(torch, pf, K, mesh, reps=100):
    """Average duration of the K.u kernel from HIP events on its stream.

    The launches rotate over enough independent (E, u, y) sets that their
    footprint is three times the 256 MB Infinity Cache, so every launch reads
    its inputs from HBM: with one set reused, part of the 240 MB working set
    of the 256x128x128 K.u stays in that cache and the kernel measures 60 us
    instead of 80.  The first set is the seeded field (x ~ U[0.01, 1],
    default_rng(0), E = Emin + x^3; u ~ N(0, 1), default_rng(1), zero on the
    constrained DOFs); the others use seeds 2, 3, ...  Before the timed
    launches the kernel runs back to back for 0.3 s (steady clocks)."""
    from pyfantom_amd import _lib
    n_dof, n_el = K.shape[0], mesh.n_elements
    bytes_alg = 16.0 * n_dof + 8.0 * n_el
    n_sets = max(3, int(np.ceil(3 * 256e6 / bytes_alg)))
    kid, k0 = K.op_args()
    sets, args = [], []
    for s in range(n_sets):
        x = np.random.default_rng(0 if s == 0 else 2 * s).uniform(0.01, 1.0, n_el)
        E = torch.from_numpy(1e-9 + x ** 3 * (1 - 1e-9)).cuda()
        u = torch.from_numpy(np.random.default_rng(1 if s == 0 else 2 * s + 1).standard_normal(n_dof)).cuda()
        u[K.constraints] = 0
        y = torch.empty_like(u)
        sets.append((E, u, y))
        args.append((kid, *K.grid, K.k0_ptr(k0), E.data_ptr(), K.mask.data_ptr(), u.data_ptr(), y.data_ptr()))
    K.set_rho(sets[0][0])
    stream = torch.cuda.current_stream()
    t_warm = time.perf_counter() + 0.3
    i = 0
    while time.perf_counter() < t_warm:
        for _ in range(50):
            _lib.call("fantom_ku_f64", *args[i % n_sets])
            i += 1
        torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for r in range(reps):
        _lib.call("fantom_ku_f64", *args[r % n_sets])
    ev1.record(stream)
    ev1.synchronize()
    t = ev0.elapsed_time(ev1) / 1e3 / reps
    return dict(seconds=t, gdof_s=n_dof / t / 1e9, achieved_gbs=bytes_alg / t / 1e9, bytes=bytes_alg,
                flops_dense=1152.0 * n_el, buffer_sets=n_sets,
                kernel="ku3d_tile_f64" if kid == _lib.KU_TILE else "ku3d_node_f64")


def time_construct(torch, K, mesh, reps):
    """Explicit CSR assembly of the seeded field: construct() once (row
    lengths, scan, fill), then the fill kernel alone timed with HIP events on
    its stream.  Algorithmic bytes: 12 B per nonzero (value, int32 column)
    written, the 8 B row pointers and the 8 B moduli read."""
    from pyfantom_amd import _lib
    x = np.random.default_rng(0).uniform(0.01, 1.0, mesh.n_elements)
    E = torch.from_numpy(1e-9 + x ** 3 * (1 - 1e-9)).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A = K.construct(E)
    torch.cuda.synchronize()
    t_construct = time.perf_counter() - t0
    _, indptr, indices, _ = K._csr_pattern
    data = A.values()
    args = (*K.grid, K.K0.data_ptr(), E.data_ptr(), K.mask.data_ptr(), indptr.data_ptr(), indices.data_ptr(),
            data.data_ptr())
    for _ in range(5):
        _lib.call("fantom_csr_assemble_f64", *args)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, min(reps, 20))
    ev0.record(stream)
    for _ in range(reps):
        _lib.call("fantom_csr_assemble_f64", *args)
    ev1.record(stream)
    ev1.synchronize()
    t = ev0.elapsed_time(ev1) / 1e3 / reps
    nnz = int(indices.shape[0])
    alg = 12.0 * nnz + 8.0 * (K.shape[0] + 1) + 8.0 * mesh.n_elements
    return {"construct_seconds": t_construct, "fill_seconds": t, "nnz": nnz, "algorithmic_bytes": alg,
            "achieved_gbs": alg / t / 1e9, "hbm_fraction": alg / t / 8e12, "kernel": "csr_fill_k<3>"}


def time_unstructured(torch, pf, reps=50, shape=(48, 32, 64), cpu=True):
    """Unstructured K.u (UniformStiffnessKernel, fantom_ku_general_f64) on a
    jittered tetrahedral mesh of the box [0,2] x [0,1] x [0,1], shape cells cut
    into 6 tetrahedra each (default 589,824 tet4, 12 x 12 element matrices),
    x = 0 face clamped, moduli U[0.1, 1.1] and u ~ N(0, 1) from
    default_rng(0).  HIP events on the stream the kernels run on; the element
    matrices (680 MB) exceed the 256 MB Infinity Cache, so every product reads
    them from HBM.  Algorithmic bytes per product: the element matrices once
    (8 m^2 per element), the moduli and connectivity (8 + 4 npe per element),
    u read and y written (16 per DOF).  The per-element products written and
    re-read by the gather (16 m per element) are implementation traffic, not
    counted.  CPU: the oracle's node-basis product on the same mesh, on the
    host's OpenMP threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from unstructured_meshes import tet_mesh
    X, el = tet_mesh(*shape, jitter=0.1, seed=2)
    mesh = pf.GeneralMesh(X, el)
    K = pf.UniformStiffnessKernel(mesh)
    K.set_constraints(np.concatenate([3 * np.flatnonzero(mesh.nodes[:, 0] < 1e-9) + d for d in range(3)]))
    rng = np.random.default_rng(0)
    E_h = rng.uniform(0.1, 1.1, mesh.n_elements)
    u_h = rng.standard_normal(K.shape[0])
    K.set_rho(torch.from_numpy(E_h).cuda())
    u = torch.from_numpy(u_h).cuda()
    for _ in range(5):
        K.dot(u)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        K.dot(u)
    ev1.record(stream)
    ev1.synchronize()
    t = ev0.elapsed_time(ev1) / 1e3 / reps
    m = K.groups[0]["m"]
    n_el, n_dof = mesh.n_elements, K.shape[0]
    alg = n_el * (8.0 * m * m + 8.0 + 4.0 * m // 3) + 16.0 * n_dof
    out = {"workload": f"tet4_{shape[0]}x{shape[1]}x{shape[2]}x6", "elements": n_el, "dofs": n_dof,
           "seconds": t, "gdof_s": n_dof / t / 1e9, "algorithmic_bytes": alg,
           "roofline": {"bound": "hbm", "achieved": alg / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / t / 1e9 / HBM_PEAK_GBS},
           "kernels": ["kg_elem_k<4, 3, 1>", "kg_gather_k<0>"]}
    if cpu:
        from oracle import oracle as O
        O.build()
        con = K.constraints.cpu().numpy()
        args = (mesh.K_flat, mesh.K_ptr, mesh.elements_flat, mesh.elements_ptr, 3)
        O.matvec_flat(*args, E_h, u_h, con)
        n_cpu = 5
        t0 = time.perf_counter()
        for _ in range(n_cpu):
            y_cpu = O.matvec_flat(*args, E_h, u_h, con)
        tc = (time.perf_counter() - t0) / n_cpu
        y_gpu = K.dot(u).cpu().numpy()
        out["cpu_baseline"] = {"value": n_dof / tc / 1e9, "unit": "GDOF/s", "seconds": tc,
                               "cores": int(O.lib().or_num_threads()), "kind": "port",
                               "sample": f"{n_cpu} products on the same mesh"}
        out["max_abs_diff_vs_oracle"] = float(np.abs(y_gpu - y_cpu).max())
        out["max_abs_y"] = float(np.abs(y_cpu).max())
    return out


def headline(legs):
    """The fastest leg that reached the FP64 compliance to 1e-6: its time
    per SIMP iteration and its ratio to the plain FP64 GMG-PCG."""
    ok = {k: v for k, v in legs.items() if v["compliance_rel_diff_vs_fp64"] <= 1e-6}
    if not ok:
        return None
    k = max(ok, key=lambda n: ok[n]["speedup_vs_fp64"])
    return {"leg": k, "s_per_iter": ok[k]["s_per_iter"], "speedup_vs_fp64": ok[k]["speedup_vs_fp64"],
            "compliance_rel_diff_vs_fp64": ok[k]["compliance_rel_diff_vs_fp64"]}


def precision_errors(torch, pf, mesh, FE_m, FE, n_level, smoother="jacobi", precision="mixed", flexible=None):
    """One cold solve of the seeded field (x ~ U[0.01, 1], default_rng(0),
    E = Emin + x^3) with the mixed-precision solver and with the FP64 one:
    relative errors of u (L2), compliance and the element sensitivity
    -u_e^T K0 u_e (L2) against FP64.  Neighbouring moduli differ by up to
    1e6 in this field and the spectral radius of D^-1 K reaches about 4.4,
    so a Jacobi weight above 2 / 4.4 amplifies the highest modes: both
    solves use a fixed omega = 0.4."""
    x = np.random.default_rng(0).uniform(0.01, 1.0, mesh.n_elements)
    E = torch.from_numpy(1e-9 + x ** 3 * (1 - 1e-9)).cuda()
    out = {}
    for tag, fe, prec in (("mixed", FE_m, precision), ("fp64", FE, "fp64")):
        s = pf.MultiGrid(mesh, fe.kernel, maxiter=1000, tol=1e-8, n_smooth=2, omega=0.4, n_level=n_level,
                         cycle="V", precision=prec, omega_schedule="fixed", fallback=False, smoother=smoother,
                         flexible=flexible if prec != "fp64" else None)
        fe.kernel.set_rho(E)
        U, res = s.solve(fe.rhs, use_last=False)
        out[tag] = (U.clone(), float(torch.dot(fe.rhs, U)), fe.kernel.process_grad(U).clone(), res,
                    s.last_iterations)
        del s
    (um, cm, dm, rm, im), (u6, c6, d6, r6, i6) = out["mixed"], out["fp64"]
    nrm = torch.linalg.vector_norm
    return {"u_rel_l2": float(nrm(um - u6) / nrm(u6)), "compliance_rel": abs(cm - c6) / abs(c6),
            "dc_rel_l2": float(nrm(dm - d6) / nrm(d6)), "rel_residual_mixed": float(rm),
            "rel_residual_fp64": float(r6), "pcg_iterations_mixed": im, "pcg_iterations_fp64": i6,
            "solver": ("GMG-PCG, 2+2 Jacobi sweeps, fixed omega 0.4" if smoother == "jacobi" else
                       "GMG-PCG, fourth-kind Chebyshev 2+2") + ", cold start, tol 1e-8"}


def measured_traffic(workload):
    """HBM bytes per launch of the fine K.u for this workload, from the
    committed counter runs (profiles/ku_tile_f64_traffic.md), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "ku_tile_f64_traffic.json")) as fh:
            return json.load(fh)[workload]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nel=(128, 64, 64), n_level=4, steps=3):
    """Whole SIMP iterations of the FP64 oracle (C/OpenMP restatement of the
    reference's CPU path) on the host: filter -> modified SIMP -> GMG-PCG to
    1e-8 (warm-started from the previous solution, hierarchy rebuilt every
    solve as the reference does) -> compliance -> dc -> chain rule -> filter
    adjoint -> OC, on the 128x64x64 cantilever from x0 = 0.3.  One untimed
    step, then the median of `steps` timed steps, all host threads; K.u and one
    PCG iteration (with its setup) also at 1 thread.  The oracle's coarse
    operators are rediscretised (2^3-averaged moduli), where the reference's
    CPU MultiGrid builds Galerkin CSR levels (solvers/CPU/_solvers.py:956-1147)."""
    from oracle import oracle as O
    O.build()
    threads = O.num_threads()
    nx, ny, nz = (n + 1 for n in nel)
    n_el, n_dof = int(np.prod(nel)), nx * ny * nz * 3
    K0 = O.k0_exact(3, 0.3)
    el = O.elements(nel)
    offs, w = O.filter_taps(3, 1.5)
    con = np.zeros(n_dof, bool)
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    clamp = (k + nz * nx * j).ravel()
    for d in range(3):
        con[clamp * 3 + d] = True
    f = np.zeros(n_dof)
    f[(np.arange(nz) + nz * (nx - 1)) * 3 + 1] = -1.0 / nz
    dg = np.full(n_el, 1.0 / n_el)

    def step(x, u):
        rho = O.filter_fwd(nel, offs, w, x)
        E = O.penalize_modified(rho, 3.0, 1.0, 1e-9)
        u, it, res = O.mgpcg_solve(nel, K0, E, con, f, n_level=n_level, n_smooth=2, omega=0.6, tol=1e-8, x0=u)
        c = float(f @ u)
        dx = O.filter_adj(nel, offs, w, O.process_dk(K0, el, u) * O.penalize_grad_modified(rho, 3.0, 1.0, 1e-9))
        xn, _ = O.oc_update(x, dx, dg, 0.3, 0.2)
        return xn, u, it, c, E

    x, u = np.full(n_el, 0.3), None
    x, u, it0, c, E = step(x, u)          # untimed
    times, its = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        x, u, it, c, E = step(x, u)
        times.append(time.perf_counter() - t0)
        its.append(it)
    t_ku = {}
    t_it = {}
    for nth in (threads, 1):
        O.set_num_threads(nth)
        v = np.random.default_rng(1).standard_normal(n_dof)
        O.matvec(K0, el, E, v, con)
        t0 = time.perf_counter()
        O.matvec(K0, el, E, v, con)
        t_ku[nth] = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.mgpcg_solve(nel, K0, E, con, f, n_level=n_level, n_smooth=2, omega=0.6, tol=1e-30, maxiter=1)
        t_it[nth] = time.perf_counter() - t0
    O.set_num_threads(threads)
    med = float(np.median(times))
    return dict(value=med, unit="s/SIMP-iter", cores=threads, kind="port",
                workload=f"cantilever_{nel[0]}x{nel[1]}x{nel[2]}_simp_oc_gmgpcg_1e-8",
                sample=(f"oracle SIMP iterations on the {nel[0]}x{nel[1]}x{nel[2]} cantilever (filter, SIMP law, "
                        f"GMG-PCG to 1e-8 warm-started, {n_level} coarsenings, rediscretised coarse operators, "
                        f"2+2 Jacobi omega 0.6, dc, filter adjoint, OC): median of {steps} after 1 untimed, "
                        f"{threads} threads"),
                step_seconds=times, pcg_iterations=its, compliance=c, cpu_model=_cpu_model(),
                ku_seconds={str(kk): vv for kk, vv in t_ku.items()},
                pcg_iteration_with_setup_seconds={str(kk): vv for kk, vv in t_it.items()},
                ku_gdof_s={str(kk): n_dof / vv / 1e9 for kk, vv in t_ku.items()})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--ku-reps", type=int, default=100)
    ap.add_argument("--legs", default="fp64,mixed,mixed_flex,bf16,mixed_cheb,mixed_learned",
                    help="SIMP legs to time after the FP64 baseline: " + ", ".join(sorted(LEGS)))
    ap.add_argument("--no-seeded", action="store_true", help="skip the seeded-field precision comparison")
    ap.add_argument("--progress", action="store_true", help="print compliance and PCG iterations every step")
    ap.add_argument("--omega-schedule", default=None, choices=["reference", "fixed"],
                    help="Jacobi weight: the reference's adaptive schedule (omega is its maximum) or fixed omega "
                         "(default: fixed for the cantilevers, as specified for them; reference for the bridge)")
    ap.add_argument("--ku-only", action="store_true",
                    help="only the isolated K.u of the workload and a 256 MiB calibration copy (for counter runs)")
    ap.add_argument("--construct-only", action="store_true",
                    help="only the explicit CSR assembly K(E) @ I of the workload (construct())")
    ap.add_argument("--unstructured-only", action="store_true",
                    help="only the unstructured tet4 K.u (UniformStiffnessKernel) and its CPU oracle")
    ap.add_argument("--no-unstructured", action="store_true", help="skip the unstructured K.u measurement")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (the sharding is a design, not code)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import pyfantom_amd as pf

    cfg = CONFIGS[args.config]
    nel, n_level = cfg["nel"], cfg["n_level"]
    if args.omega_schedule is None:
        args.omega_schedule = "reference" if cfg.get("kind") == "bridge" else "fixed"

    if args.construct_only:
        mesh = pf.StructuredMesh3D(*nel, *nel, physics=pf.LinearElasticity(E=1.0, nu=0.3))
        K = pf.StructuredStiffnessKernel(mesh)
        j, k = np.meshgrid(np.arange(nel[1] + 1), np.arange(nel[2] + 1), indexing="ij")
        clamp = (k + (nel[2] + 1) * (nel[0] + 1) * j).ravel()
        K.set_constraints(np.concatenate([3 * clamp + d for d in range(3)]))
        print(json.dumps(time_construct(torch, K, mesh, args.ku_reps)), flush=True)
        return

    if args.unstructured_only:
        res = time_unstructured(torch, pf, args.ku_reps, cpu=not args.no_cpu)
        # one device copy of a known size: calibrates FETCH_SIZE / WRITE_SIZE in counter runs
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        res["calibration_copy_bytes_each_way"] = src.numel() * 8
        print(json.dumps(res), flush=True)
        return

    if args.ku_only:
        # hardware-counter runs: the K.u launches, then one device copy of a
        # known size whose FETCH_SIZE / WRITE_SIZE calibrate the counters
        mesh = pf.StructuredMesh3D(*nel, *nel, physics=pf.LinearElasticity(E=1.0, nu=0.3))
        K = pf.StructuredStiffnessKernel(mesh)
        j, k = np.meshgrid(np.arange(nel[1] + 1), np.arange(nel[2] + 1), indexing="ij")
        clamp = (k + (nel[2] + 1) * (nel[0] + 1) * j).ravel()
        K.set_constraints(np.concatenate([3 * clamp + d for d in range(3)]))
        ku = time_ku(torch, pf, K, mesh, args.ku_reps)
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"ku_seconds": ku["seconds"], "ku_algorithmic_bytes": ku["bytes"], "kernel": ku["kernel"],
                          "calibration_copy_bytes_each_way": src.numel() * 8}), flush=True)
        return

    fallbacks = []

    def simp_run(leg):
        """W untimed + K timed SIMP iterations of one leg; returns (seconds,
        PCG iterations, outer iterations, problem objects)."""
        opts = LEGS[leg]
        mesh, K, solver, FE, filt, prob = build_problem(pf, nel, n_level, kind=cfg.get("kind", "cantilever"),
                                                        vf=cfg.get("vf", 0.3), r=cfg.get("r", 1.5),
                                                        omega=cfg.get("omega", 0.6),
                                                        omega_schedule=args.omega_schedule, **opts)
        print(f"# {leg}: setup done: {cfg['name']} dofs={K.shape[0]}", file=sys.stderr, flush=True)
        opt = pf.MMA(prob) if cfg.get("optimizer") == "mma" else pf.OC(prob, move=0.2)
        for w in range(args.warmup):
            opt.iter()
            print(f"# {leg}: warmup {w}: c={prob.f():.6e} iters={solver.last_iterations}", file=sys.stderr,
                  flush=True)
        torch.cuda.synchronize()
        iters, outer = [], []
        fallbacks.clear()
        t0 = time.perf_counter()
        for s in range(args.steps):
            opt.iter()
            iters.append(solver.last_iterations)
            outer.append(getattr(solver, "last_outer_iterations", None))
            if solver.last_fallback:
                fallbacks.append((s, solver.last_fallback_iterations))
            if args.progress:
                print(f"# {leg}: step {s}: c={prob.f():.6e} iters={solver.last_iterations}", file=sys.stderr,
                      flush=True)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        print(f"# {leg}: timed {args.steps} steps: {t / args.steps * 1e3:.1f} ms/step, pcg iters {iters}",
              file=sys.stderr, flush=True)
        return t, iters, outer, (mesh, K, solver, FE, filt, prob)

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    t, iters, _, (mesh, K, solver, FE, filt, prob) = simp_run("fp64")
    ms = t / args.steps * 1e3
    legs = {}
    for leg in [l for l in args.legs.split(",") if l and l != "fp64"]:
        tl, it_l, out_l, objs_l = simp_run(leg)
        fb_l = list(fallbacks)
        prob_l = objs_l[5]
        legs[leg] = {"options": LEGS[leg], "s_per_iter": tl / args.steps, "speedup_vs_fp64": t / tl,
                     "pcg_iterations_mean": float(np.mean(it_l)), "pcg_iterations": it_l,
                     "compliance": prob_l.f(),
                     # end state of the same number of SIMP steps from the same start
                     "compliance_rel_diff_vs_fp64": abs(prob_l.f() - prob.f()) / abs(prob.f()),
                     "solve_rel_residual": float(prob_l._residual),
                     "u_rel_l2_vs_fp64": rel(prob_l._U, prob._U),
                     "dc_rel_l2_vs_fp64": rel(prob_l.nabla_f(), prob.nabla_f()),
                     "fp64_fallbacks": [{"step": s, "fp64_iterations": n} for s, n in fb_l]}
        if LEGS[leg]["precision"] in ("mixed", "bf16") and not LEGS[leg].get("learned") and not args.no_seeded:
            legs[leg]["seeded_field"] = precision_errors(torch, pf, mesh, objs_l[3], FE, n_level,
                                                         smoother=LEGS[leg]["smoother"],
                                                         precision=LEGS[leg]["precision"],
                                                         flexible=LEGS[leg].get("flexible"))
        del objs_l, prob_l
        torch.cuda.empty_cache()
    ku = time_ku(torch, pf, K, mesh, args.ku_reps)
    print(f"# K.u: {ku['seconds']*1e6:.1f} us, {ku['achieved_gbs']:.0f} GB/s", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": t / args.steps,
        "unit": "s/SIMP-iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": cfg["name"] + ("_simp_mma" if cfg.get("optimizer") == "mma" else "_simp_oc")
                   + "_gmgpcg_1e-8", "elements": list(nel), "dofs": K.shape[0],
                   "coarsenings": n_level,
                   "smoother": f"jacobi 2+2 omega {cfg.get('omega', 0.6)} ({args.omega_schedule} schedule)",
                   "precision": "fp64 plain GMG-PCG"},
        "pcg_iterations_mean": float(np.mean(iters)),
        "compliance": prob.f(),
        "solve_rel_residual": float(prob._residual),
        "ku": {"gdof_s": ku["gdof_s"], "seconds": ku["seconds"], "kernel": ku["kernel"],
               "buffer_sets": ku["buffer_sets"]},
        "legs": legs,
        "headline": headline(legs),
        "roofline": {"bound": "hbm", "achieved": ku["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ku["achieved_gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic(cfg["name"]),
                     "kernel": ku["kernel"], "algorithmic_bytes": ku["bytes"],
                     "traffic_source": "profiles/ku_tile_f64_traffic.md (FETCH_SIZE/WRITE_SIZE, copy-calibrated)"},
    }
    if not args.no_unstructured:
        torch.cuda.empty_cache()
        out["unstructured_ku"] = time_unstructured(torch, pf, cpu=not args.no_cpu)
        print(f"# unstructured K.u: {out['unstructured_ku']['seconds']*1e6:.1f} us, "
              f"{out['unstructured_ku']['roofline']['achieved']:.0f} GB/s", file=sys.stderr, flush=True)
    if not args.no_cpu and cfg.get("kind", "cantilever") == "cantilever":
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cb = cpu_baseline()
        # the GPU's FP64 leg on the same 128x64x64 workload, same start
        cfg2 = CONFIGS["c2"]
        _, K2, solver2, _, _, prob2 = build_problem(pf, cfg2["nel"], cfg2["n_level"], omega_schedule="fixed")
        opt2 = pf.OC(prob2, move=0.2)
        opt2.iter()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(cb["pcg_iterations"].__len__()):
            opt2.iter()
        torch.cuda.synchronize()
        g2 = (time.perf_counter() - t0) / len(cb["pcg_iterations"])
        cb["gpu_fp64_same_workload_s_per_iter"] = g2
        cb["gpu_speedup_same_workload"] = cb["value"] / g2
        print(f"# cpu baseline: {cb['value']:.2f} s/SIMP-iter on {cb['cores']} threads ({cb['cpu_model']}), "
              f"GPU {g2 * 1e3:.1f} ms", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cb
        del K2, solver2, prob2, opt2
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
