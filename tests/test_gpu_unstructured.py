"""Unstructured meshes on the device: UniformStiffnessKernel /
GeneralStiffnessKernel (node-basis gather kernels of libfantom_hip.so)
against the oracle's flat restatements and a scipy assembly, on triangle,
mixed quad/triangle, tetrahedron, distorted hex and mixed hex/tet meshes,
with and without Dirichlet DOFs; CG solves against scipy's direct solve; and
size-independent properties on a 0.59M-tetrahedron mesh.

Tolerances: FP64 products summed in a different order from the oracle's
(same per-row order, different element-matrix layout on the device) agree to
1e-12 relative to the largest entry; the oracle and the kernel visit a node's
elements in the same ascending order, so most entries are bit-equal."""
import numpy as np
import pytest

from unstructured_meshes import hex_mesh, hex_tet_mesh, quad_tri_mesh, tet_mesh, tri_mesh

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


MESHES = {
    "tri": lambda: tri_mesh(40, 20),
    "quad_tri": lambda: quad_tri_mesh(40, 20),
    "tet": lambda: tet_mesh(8, 5, 5),
    "hex": lambda: hex_mesh(8, 5, 5),
    "hex_tet": lambda: hex_tet_mesh(8, 5, 5),
}


def _kernel(pf, m):
    return pf.UniformStiffnessKernel(m) if m.is_uniform else pf.GeneralStiffnessKernel(m)


def _flat(m):
    return m.K_flat, m.K_ptr, m.elements_flat, m.elements_ptr, m.dof


def _close(a, b, rtol=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    np.testing.assert_allclose(a, b, rtol=0, atol=rtol * scale)


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("constrained", [False, True])
def test_operators_match_oracle(oracle, name, constrained):
    torch = _torch()
    import pyfantom_amd as pf

    print(f"\n[unstructured] mesh={name} constrained={constrained}", flush=True)
    X, el = MESHES[name]()
    m = pf.GeneralMesh(X, el)
    K = _kernel(pf, m)
    n_dof = K.shape[0]
    rng = np.random.default_rng(5)
    E = rng.uniform(1e-3, 1.0, m.n_elements)
    u = rng.normal(size=n_dof)
    con = None
    if constrained:
        con = np.zeros(n_dof, bool)
        con[rng.choice(n_dof, n_dof // 6, replace=False)] = True
        K.set_constraints(np.flatnonzero(con))
    K.set_rho(torch.from_numpy(E).cuda())
    ud = torch.from_numpy(u).cuda()
    y = K.dot(ud).cpu().numpy()
    _close(y, oracle.matvec_flat(*_flat(m), E, u, con))
    d = K.diagonal().cpu().numpy()
    _close(d, oracle.diag_flat(*_flat(m), E, n_dof, con))
    dk = K.process_grad(ud).cpu().numpy()
    _close(dk, oracle.process_dk_flat(*_flat(m), u))
    # deterministic: a second product is bit-identical
    np.testing.assert_array_equal(K.dot(ud).cpu().numpy(), y)
    # explicit matrix (construct) against the brute-force scipy assembly
    A = K.construct(torch.from_numpy(E).cuda()).to_dense().cpu().numpy()
    B = oracle.assemble_flat(*_flat(m), E, n_dof, con).toarray()
    _close(A, B)


def test_solver_facing_modes(oracle):
    """The residual and K u . u modes the CG solver uses."""
    torch = _torch()
    import pyfantom_amd as pf
    from pyfantom_amd import _lib

    X, el = MESHES["hex_tet"]()
    m = pf.GeneralMesh(X, el)
    K = _kernel(pf, m)
    n_dof = K.shape[0]
    rng = np.random.default_rng(6)
    E = rng.uniform(0.1, 1.0, m.n_elements)
    K.set_constraints(np.arange(0, n_dof, 7))
    K.set_rho(torch.from_numpy(E).cuda())
    u = torch.from_numpy(rng.normal(size=n_dof)).cuda()
    b = torch.from_numpy(rng.normal(size=n_dof)).cuda()
    y = K.dot(u).clone()
    out = torch.empty_like(u)
    K.ku_residual(u, b, out)
    _close(out.cpu().numpy(), (b - y).cpu().numpy())
    part = torch.zeros(K.n_partials(), dtype=torch.float64, device=u.device)
    q = torch.empty_like(u)
    npart = K.ku_dot(u, q, part)
    assert npart == K.n_partials() == _lib.load().fantom_ku_general_partials(m.n_nodes, m.dof)
    np.testing.assert_array_equal(q.cpu().numpy(), y.cpu().numpy())
    assert float(part.sum()) == pytest.approx(float(u @ y), rel=1e-12)


def _cantilever_fe(pf, m, maxiter=20000):
    K = _kernel(pf, m)
    solver = pf.CG(K, maxiter=maxiter, tol=1e-12)
    FE = pf.FiniteElement(m, K, solver)
    X = m.nodes
    fixed = np.flatnonzero(X[:, 0] < 1e-9)
    FE.add_dirichlet_boundary_condition(node_ids=fixed, rhs=0)
    tip = X[:, 0].max()
    load = np.flatnonzero(X[:, 0] > tip - 1e-9)
    f = np.zeros(m.dof)
    f[1] = -1.0 / len(load)
    FE.add_point_forces(forces=f[None], node_ids=load)
    return K, FE, fixed, load


@pytest.mark.parametrize("name", ["quad_tri", "tet", "hex_tet"])
def test_cg_solve_matches_direct(oracle, name):
    torch = _torch()
    import scipy.sparse.linalg as spla
    import pyfantom_amd as pf

    print(f"\n[unstructured-cg] mesh={name}", flush=True)
    X, el = MESHES[name]()
    m = pf.GeneralMesh(X, el)
    K, FE, fixed, load = _cantilever_fe(pf, m)
    rng = np.random.default_rng(8)
    E = rng.uniform(0.2, 1.0, m.n_elements)
    U, res = FE.solve(torch.from_numpy(E).cuda())
    assert res < 1e-10   # true residual; CG stops on the recursive one
    con = K.constraints.cpu().numpy()
    A = oracle.assemble_flat(*_flat(m), E, K.shape[0], con)
    rhs = FE.rhs.cpu().numpy()
    Ud = spla.spsolve(A.tocsc(), rhs)
    np.testing.assert_allclose(U.cpu().numpy(), Ud, rtol=0, atol=1e-8 * np.abs(Ud).max())
    # compliance and its sensitivity: sum_e E_e dk_e = -U^T K U = -f.U
    dk = K.process_grad(U).cpu().numpy()
    assert float(E @ dk) == pytest.approx(-float(rhs @ Ud), rel=1e-8)


def test_positions_map_to_nearest_nodes():
    torch = _torch()
    import pyfantom_amd as pf

    X, el = MESHES["tet"]()
    m = pf.GeneralMesh(X, el)
    K = _kernel(pf, m)
    FE = pf.FiniteElement(m, K, pf.CG(K))
    rng = np.random.default_rng(9)
    P = m.nodes[rng.choice(m.n_nodes, 50, replace=False)] + rng.normal(scale=1e-3, size=(50, 3))
    got = FE.nearest_nodes(P).cpu().numpy()
    want = np.argmin(((P[:, None, :] - m.nodes[None]) ** 2).sum(-1), axis=1)
    np.testing.assert_array_equal(got, want)
    assert FE.is_3D


def test_large_tet_mesh_properties():
    """0.59M tetrahedra (~0.31M DOF): symmetry <Ku, v> = <u, Kv>, rigid
    translations in the null space without Dirichlet DOFs, and the diagonal
    equals e_i^T K e_i on sampled rows."""
    torch = _torch()
    import pyfantom_amd as pf

    print("\n[unstructured-large] tet 48x32x64 x6", flush=True)
    X, el = tet_mesh(48, 32, 64, jitter=0.1, seed=2)
    m = pf.GeneralMesh(X, el)
    K = pf.UniformStiffnessKernel(m)
    n = K.shape[0]
    g = torch.Generator(device="cuda").manual_seed(0)
    E = torch.rand(m.n_elements, dtype=torch.float64, device="cuda", generator=g) + 0.1
    K.set_rho(E)
    u = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    v = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    ku = K.dot(u).clone()
    kv = K.dot(v).clone()
    assert float(ku @ v) == pytest.approx(float(u @ kv), rel=1e-11)
    scale = float(ku.abs().max())
    for a in range(3):
        t = torch.zeros(n, dtype=torch.float64, device="cuda")
        t[a::3] = 1.0
        assert float(K.dot(t).abs().max()) < 1e-11 * scale
    d = K.diagonal().clone()
    for i in [0, 1, n // 3, n // 2 + 1, n - 1]:
        e = torch.zeros(n, dtype=torch.float64, device="cuda")
        e[i] = 1.0
        assert float(K.dot(e)[i]) == pytest.approx(float(d[i]), rel=1e-13)
