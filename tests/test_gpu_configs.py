"""The benchmark configurations at their full sizes, on the GPU, checked
against the FP64 oracle through size-independent properties:

* 128x64x64 and 256x128x128: the fine K.u of the seeded field against the
  oracle's element-loop matvec (core/CPU/_ops.py:330-368) to 1e-13;
* 256x128x128 precision form: FP64 and mixed-precision (FP32 and BF16
  V-cycles) GMG-PCG to 1e-8 on the seeded field; the oracle's K.u of each solution gives its true
  residual and a second-order estimate of the exact compliance
  (c* = f.u + u.r + r.K^-1 r, r = f - K u), which both solves must meet to
  1e-6; u and dc differences are printed;
* 384x192x192 bridge: one MMA step with the FP64 and the mixed solver, the
  oracle's true residual and the compliance identity c = f.u = u.K u;
* 48x24x24 bridge: three MMA SIMP steps against the oracle pipeline (filter,
  SIMP law, GMG-PCG to 1e-12, dc, filter adjoint, mmasub).
The 192x96x96 learned-preconditioner configuration is in test_gpu_learned.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _cantilever_con_f(nel):
    from pyfantom_amd.workloads import cantilever_nodes
    nx, ny, nz = (n + 1 for n in nel)
    clamp, edge = cantilever_nodes(nel)
    con = np.zeros(nx * ny * nz * 3, bool)
    for d in range(3):
        con[clamp * 3 + d] = True
    f = np.zeros(nx * ny * nz * 3)
    f[edge * 3 + 1] = -1.0 / nz
    return con, f


@pytest.mark.parametrize("nel", [(128, 64, 64), (256, 128, 128)])
def test_full_size_ku_vs_oracle(oracle, nel):
    """Fine K.u of the seeded field (x ~ U[0.01, 1], default_rng(0); u ~
    N(0, 1), default_rng(1), zero on the clamped DOFs) against the oracle."""
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import seeded_moduli
    torch = _torch()
    print(f"full-size K.u vs oracle {nel}", flush=True)
    mesh = pf.StructuredMesh3D(*nel, *nel, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    K = pf.StructuredStiffnessKernel(mesh)
    con, _ = _cantilever_con_f(nel)
    K.set_constraints(np.flatnonzero(con))
    E = seeded_moduli(mesh.n_elements)
    u = np.random.default_rng(1).standard_normal(con.size)
    u[con] = 0.0
    K.set_rho(torch.from_numpy(E).cuda())
    y = K.dot(torch.from_numpy(u).cuda()).cpu().numpy()
    yo = oracle.matvec(oracle.k0_exact(3, 0.3), oracle.elements(nel), E, u, con)
    err = np.abs(y - yo).max() / np.abs(yo).max()
    print(f"  max |Ku - Ku_oracle| / max |Ku_oracle| = {err:.2e}", flush=True)
    assert err <= 1e-13


def _oracle_checks(oracle, nel, E, con, f, u):
    """True relative residual of u under the oracle's operator and the
    second-order compliance estimate f.u + u.r."""
    el = oracle.elements(nel)
    K0 = oracle.k0_exact(3, 0.3)
    r = f - oracle.matvec(K0, el, E, u, con)
    return float(np.linalg.norm(r) / np.linalg.norm(f)), float(f @ u + u @ r), el, K0


def test_c3_precision_form_vs_fp64_and_oracle(oracle):
    """256x128x128 cantilever, seeded field, 1e-8: FP64 and mixed-precision
    (FP64 PCG with an FP32 V-cycle, and with the BF16 matrix-core V-cycle)
    solutions against each other and the oracle."""
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import cantilever, seeded_moduli
    torch = _torch()
    nel = (256, 128, 128)
    print("C3 precision form", flush=True)
    mesh, K, _, FE, _, _ = cantilever(nel, n_level=5)
    con, f = _cantilever_con_f(nel)
    assert np.array_equal(K.constraints.cpu().numpy(), con) and np.array_equal(FE.rhs.cpu().numpy(), f)
    E = seeded_moduli(mesh.n_elements)
    K.set_rho(torch.from_numpy(E).cuda())
    sol = {}
    for prec in ("fp64", "mixed", "bf16"):
        # the seeded field's D^-1 K reaches a spectral radius of about 4.4, so
        # a fixed Jacobi weight above 2 / 4.4 amplifies its highest modes
        s = pf.MultiGrid(mesh, K, maxiter=500, tol=1e-8, n_smooth=2, omega=0.4, n_level=5, cycle="V",
                         omega_schedule="fixed", precision=prec, fallback=False)
        U, res = s.solve(FE.rhs, use_last=False)
        u = U.cpu().numpy()
        dc = K.process_grad(U).cpu().numpy()
        sol[prec] = (u, res, s.last_iterations, dc)
        print(f"  {prec}: {s.last_iterations} PCG iterations, residual {res:.2e}", flush=True)
        del s
    u6, r6, i6, d6 = sol["fp64"]
    tr6, cstar, el, K0 = _oracle_checks(oracle, nel, E, con, f, u6)
    c6 = float(f @ u6)
    print(f"  fp64: oracle true residual {tr6:.2e}, compliance {c6!r}, oracle estimate {cstar!r}", flush=True)
    assert r6 < 1e-8 and tr6 < 1.5e-8
    assert abs(c6 - cstar) <= 1e-6 * abs(cstar)
    for prec in ("mixed", "bf16"):   # FP32 and BF16 V-cycles
        um, rm, im, dm = sol[prec]
        trm, _, _, _ = _oracle_checks(oracle, nel, E, con, f, um)
        cm = float(f @ um)
        dco = oracle.process_dk(K0, el, um)
        du = np.linalg.norm(um - u6) / np.linalg.norm(u6)
        ddc = np.linalg.norm(dm - d6) / np.linalg.norm(d6)
        print(f"  {prec}: oracle true residual {trm:.2e}, compliance {cm!r}; vs fp64: |du|/|u| {du:.2e} "
              f"|dc|/c {abs(cm - c6) / abs(c6):.2e} |d(dc)|/|dc| {ddc:.2e}; vs oracle: |c - c*|/c* "
              f"{abs(cm - cstar) / abs(cstar):.2e}", flush=True)
        assert rm < 1e-8 and trm < 1.5e-8
        assert abs(cm - cstar) <= 1e-6 * abs(cstar)
        assert abs(cm - c6) <= 1e-6 * abs(c6)
        # the device sensitivity of the solution is the oracle's on the same u
        assert np.linalg.norm(dm - dco) <= 1e-12 * np.linalg.norm(dco)


def _bridge_con_f(nel):
    from pyfantom_amd.workloads import bridge_nodes
    nx, ny, nz = (n + 1 for n in nel)
    corners, top = bridge_nodes(nel)
    con = np.zeros(nx * ny * nz * 3, bool)
    for d in range(3):
        con[corners * 3 + d] = True
    f = np.zeros(nx * ny * nz * 3)
    f[top * 3 + 1] = -1.0 / top.size
    return con, f


def test_c5_bridge_mma_step(oracle):
    """384x192x192 bridge (43M DOF), volume fraction 0.15, filter r = 2, one
    MMA step from the uniform design with the FP64 solver and one with the
    mixed solver (FP64 fallback on stagnation): each solution meets 1e-8 under
    the oracle's operator and c = f.u = u.K u; the two compliances agree."""
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import bridge
    torch = _torch()
    nel = (384, 192, 192)
    con, f = _bridge_con_f(nel)
    out = {}
    for prec in ("fp64", "mixed", "bf16"):
        print(f"C5 bridge MMA step, {prec}", flush=True)
        mesh, K, solver, FE, filt, prob = bridge(nel, n_level=6, omega=0.5, omega_schedule="reference",
                                                 precision=prec)
        if prec == "fp64":
            assert np.array_equal(K.constraints.cpu().numpy(), con)
            assert np.array_equal(FE.rhs.cpu().numpy(), f)
        opt = pf.MMA(prob)
        opt.iter()
        x = prob.get_desvars()
        assert float(x.min()) >= 0.0 and float(x.max()) <= 1.0
        c = prob.f()     # iter() evaluates the updated design: c, U and E belong to it
        U = prob._U.cpu().numpy()
        E = K.rho.cpu().numpy()
        out[prec] = (c, solver.last_iterations, float(prob._residual), U, E, opt.last_newton_steps,
                     getattr(solver, "last_fallback", False))
        del mesh, K, solver, FE, filt, prob, opt
        torch.cuda.empty_cache()
    c6, i6, r6, u6, E6, n6, _ = out["fp64"]
    cm, im, rm, um, Em, nm, fbm = out["mixed"]
    el = oracle.elements(nel)
    K0 = oracle.k0_exact(3, 0.3)
    # the two runs start from the same design; their updated designs differ
    # by the MMA step's response to the (slightly different) sensitivities
    for tag, u, c, E in (("fp64", u6, c6, E6), ("mixed", um, cm, Em)):
        Ku = oracle.matvec(K0, el, E, u, con)
        tr = np.linalg.norm(f - Ku) / np.linalg.norm(f)
        cu = float(u @ Ku)
        print(f"  {tag}: c {c!r} oracle true residual {tr:.2e} |f.u - u.Ku|/c {abs(c - cu) / abs(c):.2e}",
              flush=True)
        assert tr < 1.5e-8
        assert abs(float(f @ u) - c) <= 1e-12 * abs(c)
        assert abs(c - cu) <= 1e-6 * abs(c)
    print(f"  PCG iterations fp64 {i6} mixed {im} (fallback {fbm}); Newton steps {n6} / {nm}; "
          f"|c_mixed - c_fp64|/c {abs(cm - c6) / abs(c6):.2e}", flush=True)
    assert r6 < 1e-8 and rm < 1e-8
    assert abs(cm - c6) <= 1e-6 * abs(c6)


def test_bridge_small_mma_pipeline_vs_oracle(oracle):
    """48x24x24 bridge, three MMA steps: the device pipeline (filter r = 2,
    modified SIMP, GMG-PCG, dc, filter adjoint, MMA) against the oracle fed
    the device's design and asymptote state each step."""
    from oracle import mma_oracle as M
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import bridge
    torch = _torch()
    nel = (48, 24, 24)
    print("bridge 48x24x24 MMA pipeline", flush=True)
    mesh, K, solver, FE, filt, prob = bridge(nel, n_level=3, tol=1e-12, omega=0.5, omega_schedule="reference")
    con, f = _bridge_con_f(nel)
    assert np.array_equal(K.constraints.cpu().numpy(), con) and np.array_equal(FE.rhs.cpu().numpy(), f)
    opt = pf.MMA(prob)
    K0 = oracle.k0_exact(3, 0.3)
    el = oracle.elements(nel)
    offs, w = oracle.filter_taps(3, 2.0)
    n = prob.N()
    for it in range(1, 4):
        x = prob.get_desvars().cpu().numpy()
        rho = oracle.filter_fwd(nel, offs, w, x)
        E = oracle.penalize_modified(rho, 3.0, 1.0, 1e-9)
        u, its, res = oracle.mgpcg_solve(nel, K0, E, con, f, n_level=3, n_smooth=2, omega=0.5, tol=1e-12)
        assert res < 1e-11
        c = float(f @ u)
        dx = oracle.filter_adj(nel, offs, w, oracle.process_dk(K0, el, u)
                               * oracle.penalize_grad_modified(rho, 3.0, 1.0, 1e-9))
        assert abs(prob.f() - c) <= 1e-10 * abs(c), (it, prob.f(), c)
        df = prob.nabla_f().cpu().numpy()
        assert np.abs(df - dx).max() <= 1e-8 * np.abs(dx).max()
        g = prob.g().cpu().numpy().reshape(1)
        assert abs(g[0] - (x.mean() - 0.15)) <= 1e-12
        dg = prob.nabla_g().cpu().numpy().reshape(1, n)
        low, upp = opt.low.cpu().numpy(), opt.upp.cpu().numpy()
        x1, x2 = opt.x_1.cpu().numpy(), opt.x_2.cpu().numpy()
        xr, _, _, _ = M.mmasub(it, x, np.zeros(n), np.ones(n), x1, x2, df, g, dg, low, upp, a0=1.0, move=0.5)
        opt.iter()
        xn = prob.get_desvars().cpu().numpy()
        err = np.abs(xn - xr).max()
        print(f"  step {it}: c {c:.6e} (oracle PCG {its} its), max |x - x_oracle| {err:.2e}", flush=True)
        assert err <= 1e-7
