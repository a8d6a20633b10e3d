"""Element post-processing (MinimumCompliance.FEA, fantom_fea_locals_f64)
against the oracle: thresholded design, direct solve, then strain, stress,
strain energy, von Mises and compliance per element (3D cantilever, 2D MBB)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _problem(pf, dim):
    if dim == 3:
        nel = (12, 6, 6)
        mesh = pf.StructuredMesh3D(*nel, *nel, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    else:
        nel = (24, 8)
        mesh = pf.StructuredMesh2D(*nel, *nel, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    K = pf.StructuredStiffnessKernel(mesh)
    solver = pf.MultiGrid(mesh, K, maxiter=400, tol=1e-12, n_smooth=2, omega=0.6, n_level=1, cycle="V")
    FE = pf.FiniteElement(mesh, K, solver)
    if dim == 3:
        nx, ny, nz = (n + 1 for n in nel)
        j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
        FE.add_dirichlet_boundary_condition(node_ids=(k + nz * nx * j).ravel(), rhs=0)
        FE.add_point_forces(forces=np.array([[0.0, -1.0 / nz, 0.0]]), node_ids=np.arange(nz) + nz * (nx - 1))
        filt = pf.StructuredFilter3D(mesh, 1.5)
    else:
        nelx, nely = nel
        nx = nelx + 1
        FE.add_dirichlet_boundary_condition(node_ids=np.arange(nely + 1) * nx, dofs=np.array([[1, 0]]), rhs=0)
        FE.add_dirichlet_boundary_condition(node_ids=np.array([nelx]), dofs=np.array([[0, 1]]), rhs=0)
        FE.add_point_forces(forces=np.array([[0.0, -1.0]]), node_ids=np.array([nely * nx]))
        filt = pf.StructuredFilter2D(mesh, 1.5)
    prob = pf.MinimumCompliance(FE, filt, E_mul=[1.0], void=1e-9, penalty=3.0, volume_fraction=[0.4],
                                heavyside=False, law="modified")
    return nel, K, FE, prob


@pytest.mark.parametrize("dim", [3, 2])
def test_fea_matches_oracle(oracle, dim):
    import scipy.sparse.linalg as spla
    import pyfantom_amd as pf
    torch = _torch()
    print(f"fea {dim}d", flush=True)
    nel, K, FE, prob = _problem(pf, dim)
    # a connected design: solid (0.9) with scattered void elements (0.1), so
    # that the thresholded structure is well posed
    x = np.where(np.random.default_rng(4).uniform(size=int(np.prod(nel))) < 0.25, 0.1, 0.9)
    prob.set_desvars(torch.from_numpy(x).cuda())
    out = prob.FEA()
    # oracle: thresholded moduli, direct solve, element quantities
    x = prob.get_desvars().cpu().numpy()
    w = (x > 0.5).astype(float) + 1e-9
    K0 = oracle.k0_exact(dim, 0.3)
    el = oracle.elements(nel)
    con = K.constraints.cpu().numpy()
    f = FE.rhs.cpu().numpy()
    A = oracle.assemble(K0, el, w, len(f), con)
    u = spla.spsolve(A.tocsc(), f)
    strain, stress, energy, vm = oracle.fea_locals(K0, oracle.d_iso(dim, 0.3), oracle.b_center(dim), el, w, u)
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    assert abs(out["compliance"] - f @ u) <= 1e-9 * abs(f @ u)
    assert rel(out["Displacements"].cpu().numpy(), u) <= 1e-8
    assert rel(out["strain"].cpu().numpy(), strain) <= 1e-8
    assert rel(out["stress"].cpu().numpy(), stress) <= 1e-8
    assert rel(out["strain_energy"].cpu().numpy(), energy) <= 1e-8
    assert rel(out["von_mises"].cpu().numpy(), vm) <= 1e-8
    # the solver's iteration limit is restored after the 4x boost of the solve
    assert FE.solver.maxiter == 400
    print(f"fea {dim}d compliance {out['compliance']!r}", flush=True)
