"""CPU pins of the unstructured path: simplex element matrices, GeneralMesh
and the oracle's flat (mixed-size) node-basis restatements."""
import numpy as np
import pytest

from pyfantom_amd.geom import GeneralMesh
from pyfantom_amd.physics import LinearElasticity
from unstructured_meshes import hex_mesh, hex_tet_mesh, quad_tri_mesh, tet_mesh, tri_mesh, _KUHN


def _rigid_modes(X):
    """Rigid-body displacement fields of the nodes X (n x d), flattened."""
    n, d = X.shape
    modes = []
    for a in range(d):
        t = np.zeros((n, d))
        t[:, a] = 1.0
        modes.append(t.ravel())
    pairs = [(0, 1)] if d == 2 else [(0, 1), (1, 2), (0, 2)]
    for a, b in pairs:
        r = np.zeros((n, d))
        r[:, a], r[:, b] = -X[:, b], X[:, a]
        modes.append(r.ravel())
    return np.array(modes)


@pytest.mark.parametrize("dim", [2, 3])
def test_simplex_matrix_symmetric_rank_rigid(dim):
    rng = np.random.default_rng(1)
    X = rng.uniform(size=(dim + 1, dim)) + np.eye(dim + 1, dim) * 2.0
    if np.linalg.det(X[1:] - X[0]) < 0:
        X[[1, 2]] = X[[2, 1]]
    K = LinearElasticity(E=2.0, nu=0.3).K(X)
    assert K.shape == ((dim + 1) * dim,) * 2
    np.testing.assert_allclose(K, K.T, atol=1e-12)
    # rank = DOF - rigid modes (3 in 2D, 6 in 3D)
    n_rigid = 3 if dim == 2 else 6
    assert np.linalg.matrix_rank(K, tol=1e-9 * np.abs(K).max()) == K.shape[0] - n_rigid
    R = _rigid_modes(X)
    assert np.abs(K @ R.T).max() < 1e-10 * np.abs(K).max()
    assert np.all(np.linalg.eigvalsh(K) > -1e-10 * np.abs(K).max())


def test_simplex_measures():
    phys = LinearElasticity()
    assert phys.volume(np.array([[0, 0], [1, 0], [0, 1.0]])) == pytest.approx(0.5)
    assert phys.volume(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])) == pytest.approx(1 / 6)
    with pytest.raises(Exception, match="Node Order"):
        phys.K(np.array([[0, 0], [0, 1], [1, 0.0]]))


def test_tri_pair_energy_equals_quad_for_linear_fields():
    """A linear displacement has constant strain, which the bilinear quad and
    the two constant-strain triangles both integrate exactly: u^T K u agrees."""
    phys = LinearElasticity(E=1.0, nu=0.3)
    Xq = np.array([[0, 0], [2.0, 0], [2.0, 1.0], [0, 1.0]])
    Kq = phys.K(Xq)
    Kt = np.zeros((8, 8))
    for tri in ([0, 1, 2], [0, 2, 3]):
        Ke = phys.K(Xq[tri])
        g = np.array([[2 * a, 2 * a + 1] for a in tri]).ravel()
        Kt[np.ix_(g, g)] += Ke
    rng = np.random.default_rng(3)
    for _ in range(5):
        A, c = rng.normal(size=(2, 2)), rng.normal(size=2)
        u = (Xq @ A.T + c).ravel()
        assert u @ Kt @ u == pytest.approx(u @ Kq @ u, rel=1e-12)


def test_tets_energy_equals_hex_for_linear_fields():
    """Six Kuhn tetrahedra of a box against the exactly integrated hex8
    (trilinear, 2x2x2 Gauss) for linear displacement fields."""
    phys = LinearElasticity(E=1.0, nu=0.3)
    Xh = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0],
                   [0, 0, 1.5], [2, 0, 1.5], [2, 1, 1.5], [0, 1, 1.5]], dtype=np.float64)
    Kh = phys.K(Xh)
    Kt = np.zeros((24, 24))
    vol = 0.0
    for tet in _KUHN:
        tet = list(tet)
        if np.linalg.det(Xh[tet[1:]] - Xh[tet[0]]) < 0:
            tet[2], tet[3] = tet[3], tet[2]
        Ke = phys.K(Xh[tet])
        vol += phys.volume(Xh[tet])
        g = np.array([[3 * a, 3 * a + 1, 3 * a + 2] for a in tet]).ravel()
        Kt[np.ix_(g, g)] += Ke
    assert vol == pytest.approx(3.0)
    rng = np.random.default_rng(4)
    for _ in range(5):
        A, c = rng.normal(size=(3, 3)), rng.normal(size=3)
        u = (Xh @ A.T + c).ravel()
        assert u @ Kt @ u == pytest.approx(u @ Kh @ u, rel=1e-12)


def test_general_mesh_drops_unused_nodes():
    X, el = tri_mesh(3, 2)
    extra = np.vstack([X, [[9.0, 9.0], [8.0, 8.0]]])
    # put two unused nodes first so every id shifts
    Xs = np.vstack([extra[-2:], extra[:-2]])
    m0 = GeneralMesh(X, el)
    m1 = GeneralMesh(Xs, el + 2)
    assert m1.n_nodes == X.shape[0]
    np.testing.assert_array_equal(m1.nodes, m0.nodes)
    np.testing.assert_array_equal(m1.elements_flat, m0.elements_flat)
    np.testing.assert_allclose(m1.K_flat, m0.K_flat)
    with pytest.raises(ValueError):
        GeneralMesh(X, el + X.shape[0])


def test_general_mesh_layout():
    X, el = quad_tri_mesh(4, 2)
    m = GeneralMesh(X, el)
    assert not m.is_uniform and m.dof == 2
    sizes = np.diff(m.elements_ptr)
    assert set(sizes.tolist()) == {3, 4}
    np.testing.assert_array_equal(np.diff(m.K_ptr), (sizes * 2) ** 2)
    assert m.K_ptr[-1] == m.K_flat.size
    assert m.volume == pytest.approx(2.0)
    Xt, elt = tet_mesh(2, 1, 1)
    mt = GeneralMesh(Xt, elt)
    assert mt.is_uniform and mt.Ks.shape == (12, 12, 12)
    np.testing.assert_allclose(mt.K_flat, mt.Ks.reshape(-1))


def test_general_mesh_of_box_hexes_matches_structured_k0(oracle):
    """Undistorted hexes of a grid: every element matrix is the box K0."""
    nel = (3, 2, 2)
    X = oracle.nodes(nel, (3.0, 2.0, 2.0))
    m = GeneralMesh(X, oracle.elements(nel))
    K0 = oracle.k0_exact(3, 1.0 / 3.0)
    for Ke in m.Ks:
        np.testing.assert_allclose(Ke, K0, rtol=1e-10, atol=1e-12 * np.abs(K0).max())


def test_kernels_reject_wrong_mesh_kind():
    from pyfantom_amd.stiffness import GeneralStiffnessKernel, UniformStiffnessKernel

    Xm, elm = quad_tri_mesh(2, 2)
    with pytest.raises(ValueError, match="not uniform"):
        UniformStiffnessKernel(GeneralMesh(Xm, elm))
    Xt, elt = tri_mesh(2, 2)
    with pytest.raises(ValueError, match="is uniform"):
        GeneralStiffnessKernel(GeneralMesh(Xt, elt))


MESHES = {
    "tri": lambda: tri_mesh(5, 3),
    "quad_tri": lambda: quad_tri_mesh(4, 3),
    "tet": lambda: tet_mesh(3, 2, 2),
    "hex": lambda: hex_mesh(3, 2, 2),
    "hex_tet": lambda: hex_tet_mesh(4, 2, 2),
}


def _flat(m):
    return m.K_flat, m.K_ptr, m.elements_flat, m.elements_ptr, m.dof


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("constrained", [False, True])
def test_oracle_flat_vs_assembly(oracle, name, constrained):
    X, el = MESHES[name]()
    m = GeneralMesh(X, el)
    rng = np.random.default_rng(7)
    n_dof = m.n_nodes * m.dof
    E = rng.uniform(0.1, 1.0, m.n_elements)
    u = rng.normal(size=n_dof)
    con = None
    if constrained:
        con = np.zeros(n_dof, bool)
        con[rng.choice(n_dof, n_dof // 5, replace=False)] = True
    K = oracle.assemble_flat(*_flat(m), E, n_dof, con)
    y = oracle.matvec_flat(*_flat(m), E, u, con)
    np.testing.assert_allclose(y, K @ u, rtol=1e-12, atol=1e-12 * np.abs(K @ u).max())
    np.testing.assert_allclose(oracle.diag_flat(*_flat(m), E, n_dof, con), K.diagonal(), rtol=1e-13, atol=1e-300)
    v = rng.normal(size=n_dof)
    assert v @ oracle.matvec_flat(*_flat(m), E, u, con) == pytest.approx(u @ oracle.matvec_flat(*_flat(m), E, v, con),
                                                                            rel=1e-11)
    dk = oracle.process_dk_flat(*_flat(m), u)
    for e in range(m.n_elements):
        nodes = m.elements_flat[m.elements_ptr[e]:m.elements_ptr[e + 1]].astype(np.int64)
        g = (nodes[:, None] * m.dof + np.arange(m.dof)).ravel()
        Ke = m.K_flat[m.K_ptr[e]:m.K_ptr[e + 1]].reshape(g.size, g.size)
        assert dk[e] == pytest.approx(-u[g] @ Ke @ u[g], rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("nel", [(3, 2, 2), (4, 3)])
def test_oracle_flat_equals_shared_k0_on_grid(oracle, nel):
    """On a structured grid the per-element form with copies of K0 is the
    shared-K0 node-basis product, diagonal and sensitivity."""
    dim = len(nel)
    K0 = oracle.k0_exact(dim, 0.3)
    elems = oracle.elements(nel)
    npe = elems.shape[1]
    m = K0.shape[0]
    n_nodes = int(np.prod([n + 1 for n in nel]))
    n_dof = n_nodes * dim
    rng = np.random.default_rng(11)
    E = rng.uniform(0.1, 1, len(elems))
    u = rng.normal(size=n_dof)
    con = np.zeros(n_dof, bool)
    con[: 3 * dim] = True
    K_flat = np.tile(K0.ravel(), len(elems))
    K_ptr = np.arange(len(elems) + 1, dtype=np.int64) * m * m
    eptr = np.arange(len(elems) + 1, dtype=np.int32) * npe
    args = (K_flat, K_ptr, elems.ravel(), eptr, dim)
    np.testing.assert_array_equal(oracle.matvec_flat(*args, E, u, con), oracle.matvec(K0, elems, E, u, con))
    np.testing.assert_array_equal(oracle.diag_flat(*args, E, n_dof, con), oracle.diag(K0, elems, E, n_dof, con))
    np.testing.assert_allclose(oracle.process_dk_flat(*args, u), oracle.process_dk(K0, elems, u), rtol=1e-12)
