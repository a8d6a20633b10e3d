"""dtype=float32 meshes (the reference dispatches f4 kernels:
core/CUDA/Kernels/_matvec.py:25-31, core/CPU/_ops.py:212-245): K.dot runs the
FP32 tiled operator and returns float32; diagonal and process_grad return the
FP64 results in float32; the FE / problem / solver stack stays FP64, so a SIMP
step on a float32 mesh matches the float64 mesh to the rounding of K0."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def test_float32_kernel_ops(oracle):
    import pyfantom_amd as pf
    torch = _torch()
    nel = (24, 12, 20)
    print("float32 kernel ops", flush=True)
    mesh = pf.StructuredMesh3D(*nel, *nel, dtype=np.float32, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    K = pf.StructuredStiffnessKernel(mesh)
    nx, ny, nz = (n + 1 for n in nel)
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    clamp = (k + nz * nx * j).ravel()
    con = np.zeros(nx * ny * nz * 3, bool)
    for d in range(3):
        con[clamp * 3 + d] = True
    K.set_constraints(np.flatnonzero(con))
    E = np.random.default_rng(0).uniform(0.01, 1.0, mesh.n_elements) ** 3
    K.set_rho(torch.from_numpy(E).cuda())
    assert K.rho.dtype == torch.float32 and K.rho64.dtype == torch.float64
    u = np.random.default_rng(1).standard_normal(con.size)
    u[con] = 0.0
    y = K.dot(torch.from_numpy(u).cuda())
    assert y.dtype == torch.float32
    K0 = oracle.k0_exact(3, 0.3)
    assert np.abs(mesh.K_single - K0).max() <= 1e-7 * np.abs(K0).max()
    yo = oracle.matvec(K0, oracle.elements(nel), E.astype(np.float32).astype(np.float64),
                       u.astype(np.float32).astype(np.float64), con)
    err = np.abs(y.cpu().numpy() - yo).max() / np.abs(yo).max()
    print(f"  FP32 K.u vs oracle: {err:.2e}", flush=True)
    assert err <= 2e-6
    d = K.diagonal()
    assert d.dtype == torch.float32
    do = oracle.diag(K0, oracle.elements(nel), E, con.size, con)
    assert np.abs(d.cpu().numpy() - do).max() <= 1e-6 * np.abs(do).max()
    g = K.process_grad(torch.from_numpy(u).cuda())
    assert g.dtype == torch.float32
    go = oracle.process_dk(K0, oracle.elements(nel), u)
    assert np.abs(g.cpu().numpy() - go).max() <= 1e-6 * np.abs(go).max()


def test_float32_mesh_simp_step():
    """One OC step on float32 and float64 meshes of a 32x16x16 cantilever:
    compliance and sensitivities agree to the FP32 rounding of K0."""
    import pyfantom_amd as pf
    torch = _torch()
    print("float32 mesh SIMP step", flush=True)
    nel = (32, 16, 16)
    out = {}
    for dt in (np.float64, np.float32):
        mesh = pf.StructuredMesh3D(*nel, *nel, dtype=dt, physics=pf.LinearElasticity(E=1.0, nu=0.3))
        K = pf.StructuredStiffnessKernel(mesh)
        solver = pf.MultiGrid(mesh, K, maxiter=500, tol=1e-10, n_smooth=2, omega=0.6, n_level=3, cycle="V",
                              omega_schedule="fixed")
        FE = pf.FiniteElement(mesh, K, solver)
        nx, ny, nz = (n + 1 for n in nel)
        j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
        FE.add_dirichlet_boundary_condition(node_ids=(k + nz * nx * j).ravel(), rhs=0)
        FE.add_point_forces(forces=np.array([[0.0, -1.0 / nz, 0.0]]), node_ids=np.arange(nz) + nz * (nx - 1))
        filt = pf.StructuredFilter3D(mesh, 1.5)
        prob = pf.MinimumCompliance(FE, filt, E_mul=[1.0], void=1e-9, penalty=3.0, volume_fraction=[0.3],
                                    heavyside=False, law="modified")
        opt = pf.OC(prob, move=0.2)
        opt.iter()
        out[dt] = (prob.f(), prob.nabla_f().double().cpu().numpy())
    (c64, d64), (c32, d32) = out[np.float64], out[np.float32]
    print(f"  compliance float64 {c64!r} float32 mesh {c32!r}", flush=True)
    assert abs(c32 - c64) <= 1e-6 * abs(c64)
    assert np.linalg.norm(d32 - d64) <= 1e-5 * np.linalg.norm(d64)


def test_float32_2d_is_rejected():
    import pyfantom_amd as pf
    _torch()
    mesh = pf.StructuredMesh2D(8, 4, 8, 4, dtype=np.float32, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    with pytest.raises(NotImplementedError):
        pf.StructuredStiffnessKernel(mesh)
