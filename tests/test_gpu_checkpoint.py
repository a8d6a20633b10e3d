"""Checkpoint / resume on the device: a SIMP run saved after a few steps
and resumed in freshly built objects continues bit for bit like the
uninterrupted run (OC on the 2D MBB beam, MMA on a 3D cantilever)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _cantilever(pf, nel):
    mesh = pf.StructuredMesh3D(*nel, *nel, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    K = pf.StructuredStiffnessKernel(mesh)
    solver = pf.MultiGrid(mesh, K, maxiter=500, tol=1e-10, n_smooth=2, omega=0.6, n_level=2, cycle="V")
    FE = pf.FiniteElement(mesh, K, solver)
    nx, ny, nz = (n + 1 for n in nel)
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    FE.add_dirichlet_boundary_condition(node_ids=(k + nz * nx * j).ravel(), rhs=0)
    FE.add_point_forces(forces=np.array([[0.0, -1.0 / nz, 0.0]]), node_ids=np.arange(nz) + nz * (nx - 1))
    filt = pf.StructuredFilter3D(mesh, 1.5)
    return pf.MinimumCompliance(FE, filt, E_mul=[1.0], void=1e-9, penalty=3.0, volume_fraction=[0.3],
                                heavyside=False, law="modified")


def _mbb(pf, nelx=60, nely=20):
    mesh = pf.StructuredMesh2D(nelx, nely, nelx, nely, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    K = pf.StructuredStiffnessKernel(mesh)
    solver = pf.MultiGrid(mesh, K, maxiter=500, tol=1e-10, n_smooth=2, omega=0.6, n_level=2, cycle="V")
    FE = pf.FiniteElement(mesh, K, solver)
    nx = nelx + 1
    FE.add_dirichlet_boundary_condition(node_ids=np.arange(nely + 1) * nx, dofs=np.array([[1, 0]]), rhs=0)
    FE.add_dirichlet_boundary_condition(node_ids=np.array([nelx]), dofs=np.array([[0, 1]]), rhs=0)
    FE.add_point_forces(forces=np.array([[0.0, -1.0]]), node_ids=np.array([nely * nx]))
    filt = pf.StructuredFilter2D(mesh, 1.5)
    return pf.MinimumCompliance(FE, filt, E_mul=[1.0], void=1e-9, penalty=3.0, volume_fraction=[0.5],
                                heavyside=False, law="modified")


@pytest.mark.parametrize("case", ["oc_mbb", "mma_cantilever"])
def test_resume_is_bit_identical(tmp_path, case):
    import pyfantom_amd as pf
    torch = _torch()
    print(f"checkpoint {case}", flush=True)
    make = (lambda: _mbb(pf)) if case == "oc_mbb" else (lambda: _cantilever(pf, (16, 8, 8)))
    opt_cls = (lambda p: pf.OC(p, move=0.2)) if case == "oc_mbb" else pf.MMA
    prob = make()
    opt = opt_cls(prob)
    for _ in range(3):
        opt.iter()
    path = tmp_path / "run.npz"
    pf.checkpoint.save(path, prob, opt)
    for _ in range(2):
        opt.iter()
    x_ref, f_ref = prob.get_desvars().clone(), prob.f()
    # a fresh run restored from the file
    prob2 = make()
    opt2 = opt_cls(prob2)
    pf.checkpoint.load(path, prob2, opt2)
    assert prob2.iteration == 3
    for _ in range(2):
        opt2.iter()
    assert torch.equal(prob2.get_desvars(), x_ref)
    assert prob2.f() == f_ref
    print(f"resumed run matches: f = {f_ref!r}", flush=True)
