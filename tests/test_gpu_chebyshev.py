"""Fourth-kind Chebyshev smoothing in the native drivers (the reference's
CPU MultiGrid smoother, solvers/CPU/_solvers.py:884-954), through the
reference-shaped MultiGrid API:

* the power-iteration estimate of rho(D^-1 A) on the fine level against the
  top generalised eigenvalue of (A, diag A) from scipy;
* FP64 and mixed-precision solves against the FP64 oracle's GMG-PCG;
* the seeded 128x64x64 field, on which fixed-weight Jacobi (omega = 0.6)
  diverges: iteration counts of Jacobi and Chebyshev;
* graph-replayed iterations equal eager ones bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _solver(pf, mesh, K, **kw):
    args = dict(maxiter=500, tol=1e-10, n_smooth=2, omega=0.6, n_level=3, cycle="V", omega_schedule="fixed",
                smoother="chebyshev")
    args.update(kw)
    return pf.MultiGrid(mesh, K, **args)


def test_chebyshev_rho_estimate(oracle):
    """16x8x8 clamped cantilever, seeded moduli: the fine-level estimate after
    40 power iterations lies within 3% below the top eigenvalue of D^-1 A
    (A with the Dirichlet rows and columns of the reference's convention)."""
    import scipy.linalg as sla
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import cantilever, seeded_moduli
    torch = _torch()
    nel = (16, 8, 8)
    print("chebyshev rho estimate", flush=True)
    mesh, K, _, FE, _, _ = cantilever(nel, n_level=2)
    E = seeded_moduli(mesh.n_elements, seed=3)
    K.set_rho(torch.from_numpy(E).cuda())
    s = _solver(pf, mesh, K, n_level=2, cheb_power_iters=(40, 4))
    s.solve(FE.rhs, use_last=False)
    rho = s.smoother_rho()
    con = K.constraints.cpu().numpy()
    A = oracle.assemble(oracle.k0_exact(3, 0.3), oracle.elements(nel), E, len(con), con).toarray()
    lam = sla.eigh(A, np.diag(np.diag(A)), eigvals_only=True)[-1]
    print(f"  rho estimates {rho}, top eigenvalue of D^-1 A {lam:.6f}", flush=True)
    assert 0.97 * lam <= rho[0] <= lam * (1 + 1e-9)
    assert len(rho) == 2 and rho[1] > 1.0


def test_chebyshev_solve_vs_oracle(oracle):
    """64x32x32 cantilever, filtered seeded field: the FP64 Chebyshev-smoothed
    GMG-PCG to 1e-12 matches the oracle's GMG-PCG to 1e-10 in compliance; the
    mixed-precision solve with the same smoother reaches 1e-8 without the
    FP64 fallback and matches to 1e-6."""
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import cantilever
    from test_gpu_solver import filtered_field, oracle_cantilever_inputs
    torch = _torch()
    nel = (64, 32, 32)
    print("chebyshev solve vs oracle", flush=True)
    mesh, K, _, FE, _, _ = cantilever(nel, n_level=3)
    E = filtered_field(oracle, nel)
    K.set_rho(torch.from_numpy(E).cuda())
    s64 = _solver(pf, mesh, K, tol=1e-12)
    u64, r64 = s64.solve(FE.rhs, use_last=False)
    u64 = u64.clone()
    con, f = oracle_cantilever_inputs(nel)
    uo, it, ro = oracle.mgpcg_solve(nel, oracle.k0_exact(3, 0.3), E, con, f, n_level=3, n_smooth=2, omega=0.6,
                                    tol=1e-12)
    c64, co = float(f @ u64.cpu().numpy()), float(f @ uo)
    jac = _solver(pf, mesh, K, tol=1e-12, smoother="jacobi")
    jac.solve(FE.rhs, use_last=False)
    print(f"  FP64: chebyshev {s64.last_iterations} its, jacobi {jac.last_iterations} its, oracle {it} its; "
          f"|c - c_oracle|/c {abs(c64 - co) / abs(co):.2e}; rho {s64.smoother_rho()}", flush=True)
    assert r64 < 1e-12
    assert abs(c64 - co) <= 1e-10 * abs(co)
    mx = _solver(pf, mesh, K, tol=1e-8, precision="mixed")
    um, rm = mx.solve(FE.rhs, use_last=False)
    cm = float(torch.dot(FE.rhs, um))
    print(f"  mixed: {mx.last_iterations} its, fallback {mx.last_fallback}, |c - c64|/c {abs(cm - c64) / c64:.2e}, "
          f"rho {mx.smoother_rho()}", flush=True)
    assert not mx.last_fallback and rm < 1e-8
    assert abs(cm - c64) <= 1e-6 * abs(c64)


def test_chebyshev_on_seeded_field():
    """128x64x64 seeded field (x ~ U[0.01, 1], E = Emin + x^3): fixed-weight
    Jacobi at omega = 0.6 does not reach 1e-8 in 200 iterations (rho(D^-1 A)
    is about 4.4, so 0.6 amplifies the top modes); the Chebyshev smoother,
    which scales itself by the estimated rho, converges, and its solution
    agrees with the omega = 0.4 Jacobi solve."""
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import cantilever, seeded_moduli
    torch = _torch()
    nel = (128, 64, 64)
    print("chebyshev on the seeded 128x64x64 field", flush=True)
    mesh, K, _, FE, _, _ = cantilever(nel, n_level=4)
    K.set_rho(torch.from_numpy(seeded_moduli(mesh.n_elements)).cuda())
    out = {}
    for tag, kw in (("jacobi 0.6", dict(smoother="jacobi", omega=0.6)),
                    ("jacobi 0.4", dict(smoother="jacobi", omega=0.4)),
                    ("chebyshev", dict(smoother="chebyshev")),
                    ("chebyshev mixed", dict(smoother="chebyshev", precision="mixed", fallback=False))):
        s = _solver(pf, mesh, K, tol=1e-8, maxiter=200, n_level=4, **kw)
        U, res = s.solve(FE.rhs, use_last=False)
        out[tag] = (s.last_iterations, res, float(torch.dot(FE.rhs, U)))
        print(f"  {tag}: {s.last_iterations} its, residual {res:.2e}" +
              (f", rho {['%.3f' % v for v in s.smoother_rho()]}" if "chebyshev" in tag else ""), flush=True)
        del s
    assert not out["jacobi 0.6"][1] < 1e-8
    for tag in ("jacobi 0.4", "chebyshev", "chebyshev mixed"):
        assert out[tag][1] < 1e-8, tag
    c4 = out["jacobi 0.4"][2]
    assert abs(out["chebyshev"][2] - c4) <= 1e-6 * abs(c4)
    assert abs(out["chebyshev mixed"][2] - c4) <= 1e-6 * abs(c4)


def test_chebyshev_graph_matches_eager(oracle):
    """Graph-replayed Chebyshev iterations equal eagerly launched ones bit for
    bit over several solves with new moduli (each setup refreshes rho)."""
    import pyfantom_amd as pf
    from pyfantom_amd.workloads import cantilever
    from test_gpu_solver import filtered_field
    torch = _torch()
    nel = (32, 16, 16)
    print("chebyshev graph vs eager", flush=True)
    mesh, K, _, FE, _, _ = cantilever(nel, n_level=3)
    sg = _solver(pf, mesh, K, use_graph=True)
    se = _solver(pf, mesh, K, use_graph=False)
    for trial in range(3):
        K.set_rho(torch.from_numpy(filtered_field(oracle, nel, seed=trial)).cuda())
        xg, rg = sg.solve(FE.rhs, use_last=False)
        xg = xg.clone()
        xe, re = se.solve(FE.rhs, use_last=False)
        assert sg.last_iterations == se.last_iterations
        assert torch.equal(xg, xe), trial
        assert sg.smoother_rho() == se.smoother_rho()
        assert rg < 1e-10
