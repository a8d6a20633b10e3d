"""CPU checks that pin the FP64 oracle (oracle/oracle.c + oracle/oracle.py).

Every oracle function is compared with something independent of its own
formula: brute-force enumeration, exact rational integration, explicit sparse
matrices, dense products, finite differences or a direct solve.  Each test
would fail for the plausible mistakes noted in its docstring.
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rng(s=0):
    return np.random.default_rng(s)


# --------------------------------------------------------------------------- A1
@pytest.mark.parametrize("nel", [(1, 1, 1), (2, 1, 3), (3, 2, 2), (4, 3, 5), (1, 1), (3, 2), (5, 4)])
def test_connectivity_bruteforce(oracle, nel):
    """Index formulas vs coordinate lookup (catches swapped axes / wrong strides)."""
    assert np.array_equal(oracle.elements(nel), oracle.elements_bruteforce(nel))


def test_connectivity_golden_hash(oracle):
    """Bit-exact connectivity at the benchmark sizes against committed hashes."""
    with open(os.path.join(GOLDEN, "connectivity_sha256.json")) as f:
        ref = json.load(f)
    for key, h in ref.items():
        nel = tuple(int(v) for v in key.split("x"))
        el = oracle.elements(nel)
        assert hashlib.sha256(el.tobytes()).hexdigest() == h, key


@pytest.mark.parametrize("nel", [(3, 2, 4), (4, 3)])
def test_nodes_numbering_and_spacing(oracle, nel):
    """Node n(i, j, k) = k + nz i + nz nx j (2D: j nx + i) sits at (i h_x,
    j h_y, k h_z), the last one of each axis exactly at the length (catches a
    transposed meshgrid or a step formed as L * (i / n))."""
    L = (1.3, 0.7, 0.1)[:len(nel)]
    for dt in (np.float64, np.float32):
        X = oracle.nodes(nel, L, dt)
        n = [v + 1 for v in nel]
        for idx in np.ndindex(*n):
            if len(nel) == 3:
                i, j, k = idx
                node = k + n[2] * i + n[2] * n[0] * j
            else:
                i, j = idx
                node = j * n[0] + i
            for a, c in enumerate(idx):
                step = dt(dt(L[a]) / dt(nel[a]))
                want = dt(L[a]) if c == nel[a] else dt(dt(c) * step)
                assert X[node, a] == want


def test_element_volumes_positive(oracle):
    """Every hex has positive Jacobian volume with the reference node order."""
    nel = (3, 2, 2)
    X = oracle.nodes(nel, nel)
    el = oracle.elements(nel)
    for e in el:
        P = X[e]
        # signed volume of the tetra (0,1,3,4) is positive for a right-handed hex
        v = np.linalg.det(np.stack([P[1] - P[0], P[3] - P[0], P[4] - P[0]]))
        assert v > 0


# --------------------------------------------------------------------------- A3
def test_k0_exact_properties(oracle):
    """Symmetric, PSD, rigid-body nullspace of dim 6 (3D) / 3 (2D)."""
    for dim, nrb in ((3, 6), (2, 3)):
        K = oracle.k0_exact(dim, 0.3)
        assert np.allclose(K, K.T, atol=0)
        w = np.linalg.eigvalsh(K)
        assert np.sum(np.abs(w) < 1e-12 * w.max()) == nrb
        assert w.min() > -1e-13


def test_k0_exact_rigid_modes(oracle):
    """Translations and rotations are in the nullspace (catches a sign slip in B)."""
    K = oracle.k0_exact(3, 0.3)
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    modes = []
    for c in range(3):
        t = np.zeros((8, 3))
        t[:, c] = 1
        modes.append(t.ravel())
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1
        modes.append(np.cross(w[None, :], X).ravel())
    for m in modes:
        assert np.abs(K @ m).max() < 1e-14


def test_k0_2d_matches_top88(oracle):
    """2D K0 equals the published top88 KE under the node permutation that maps
    top88's (y-down) node order onto the reference order."""
    K = oracle.k0_exact(2, 0.3)
    T = oracle.k0_top88(0.3)
    import itertools
    found = False
    for perm in itertools.permutations(range(4)):
        # flips of the y displacement component for the y-down grid
        for sy in (1.0, -1.0):
            P = np.zeros((8, 8))
            for a, b in enumerate(perm):
                P[2 * a, 2 * b] = 1.0
                P[2 * a + 1, 2 * b + 1] = sy
            if np.allclose(P @ T @ P.T, K, atol=1e-14):
                found = True
                break
        if found:
            break
    assert found


def _top3d_ke(nu):
    """The published unit-cube hex8 KE of top3d (Liu & Tovey 2014, lk_H8):
    14 coefficients k = A^T [1, nu] / 144 arranged in six 6x6 blocks, times
    1 / ((1 + nu)(1 - 2 nu))."""
    A = np.array([[32, 6, -8, 6, -6, 4, 3, -6, -10, 3, -3, -3, -4, -8],
                  [-48, 0, 0, -24, 24, 0, 0, 0, 12, -12, 0, 12, 12, 12]], float)
    k = np.concatenate([[0.0], A.T @ np.array([1.0, nu]) / 144.0])   # 1-based as published
    B = lambda rows: np.array([[k[i] for i in r] for r in rows])
    K1 = B([[1, 2, 2, 3, 5, 5], [2, 1, 2, 4, 6, 7], [2, 2, 1, 4, 7, 6], [3, 4, 4, 1, 8, 8], [5, 6, 7, 8, 1, 2],
            [5, 7, 6, 8, 2, 1]])
    K2 = B([[9, 8, 12, 6, 4, 7], [8, 9, 12, 5, 3, 5], [10, 10, 13, 7, 4, 6], [6, 5, 11, 9, 2, 10],
            [4, 3, 5, 2, 9, 12], [11, 4, 6, 12, 10, 13]])
    K3 = B([[6, 7, 4, 9, 12, 8], [7, 6, 4, 10, 13, 10], [5, 5, 3, 8, 12, 9], [9, 10, 2, 6, 11, 5],
            [12, 13, 10, 11, 6, 4], [2, 12, 9, 4, 5, 3]])
    K4 = B([[14, 11, 11, 13, 10, 10], [11, 14, 11, 12, 9, 8], [11, 11, 14, 12, 8, 9], [13, 12, 12, 14, 7, 7],
            [10, 9, 8, 7, 14, 11], [10, 8, 9, 7, 11, 14]])
    K5 = B([[1, 2, 8, 3, 5, 4], [2, 1, 8, 4, 6, 11], [8, 8, 1, 5, 11, 6], [3, 4, 5, 1, 8, 2], [5, 6, 11, 8, 1, 8],
            [4, 11, 6, 2, 8, 1]])
    K6 = B([[14, 11, 7, 13, 10, 12], [11, 14, 7, 12, 9, 2], [7, 7, 14, 10, 2, 9], [13, 12, 10, 14, 7, 11],
            [10, 9, 2, 7, 14, 7], [12, 2, 9, 11, 7, 14]])
    KE = np.block([[K1, K2, K3, K4], [K2.T, K5, K6, K3.T], [K3.T, K6, K5.T, K2.T], [K4, K3, K2, K1.T]])
    return KE / ((nu + 1) * (1 - 2 * nu))


@pytest.mark.parametrize("nu", [0.3, 0.25, 0.45])
def test_k0_3d_matches_top3d(oracle, nu):
    """3D K0 (exact rational Gauss integration) equals top3d's published KE
    entry for entry: top3d's local node order is the reference's
    (physics/LinearElasticity.py:700-834 with Gauss points at +-1/sqrt(3))."""
    K = oracle.k0_exact(3, nu)
    T = _top3d_ke(nu)
    assert np.abs(K - T).max() <= 1e-15 * np.abs(T).max()


def test_simp_law_special_values(oracle):
    """Hand values of both SIMP laws (Problem/CPU/MinimumCompliance.py:358-377):
    reference law E = max(clip(H(rho)^p, void, 1) * E_mul, void) without and
    with the Heaviside projection (H(0) = 0, H(eta) = eta for eta = 0.5,
    H(1) = 1), and the modified law Emin + rho^p (E0 - Emin)."""
    rho = np.array([0.0, 0.5, 1.0, 0.1, 1.2])
    # p = 3, void 1e-6, E_mul 1
    np.testing.assert_allclose(oracle.penalize_reference(rho, 3.0, 1e-6, 1.0, heavyside=False),
                               [1e-6, 0.125, 1.0, 1e-3, 1.0], rtol=1e-15)
    # E_mul = 2, void = 0.01: rho = 0.1 -> 1e-3 clipped up to void before the scaling
    np.testing.assert_allclose(oracle.penalize_reference(rho, 3.0, 0.01, 2.0, heavyside=False),
                               [0.02, 0.25, 2.0, 0.02, 2.0], rtol=1e-15)
    # E_mul below void: the second clip raises everything to void
    np.testing.assert_allclose(oracle.penalize_reference(rho, 3.0, 0.01, 0.001, heavyside=False),
                               [0.01] * 5, rtol=1e-15)
    # p = 1 is the plain clip
    np.testing.assert_allclose(oracle.penalize_reference(rho, 1.0, 0.05, 1.0, heavyside=False),
                               [0.05, 0.5, 1.0, 0.1, 1.0], rtol=1e-15)
    # Heaviside, beta = 2 and 8, eta = 0.5
    for beta in (2.0, 8.0):
        e = oracle.penalize_reference(np.array([0.0, 0.5, 1.0]), 3.0, 1e-6, 1.0, heavyside=True, beta=beta)
        np.testing.assert_allclose(e, [1e-6, 0.125, 1.0], rtol=1e-14)
    np.testing.assert_allclose(oracle.penalize_modified(np.array([0.0, 0.5, 1.0]), 3.0, 1.0, 1e-9),
                               [1e-9, 1e-9 + (1 - 1e-9) / 8, 1.0], rtol=1e-15)
    np.testing.assert_allclose(oracle.penalize_modified(np.array([0.0, 0.5, 1.0]), 2.0, 10.0, 0.1),
                               [0.1, 0.1 + 9.9 / 4, 10.0], rtol=1e-15)
    # gradients: p H^(p-1) H' E_mul (the clip is ignored); H' = beta (1 - tanh^2(beta (rho - eta))) / den
    np.testing.assert_allclose(oracle.penalize_grad_reference(np.array([0.5, 1.0]), 3.0, 2.0, heavyside=False),
                               [1.5, 6.0], rtol=1e-15)
    den = 2 * np.tanh(1.0)
    np.testing.assert_allclose(oracle.penalize_grad_reference(np.array([0.5]), 3.0, 1.0, heavyside=True, beta=2.0),
                               [3 * 0.25 * 2.0 / den], rtol=1e-14)
    np.testing.assert_allclose(oracle.penalize_grad_modified(np.array([0.5]), 3.0, 1.0, 1e-9),
                               [0.75 * (1 - 1e-9)], rtol=1e-15)


def test_product_k0_matches_exact(oracle):
    """Product host K0 (Gauss quadrature) equals exact integration, 3D and 2D,
    unit and anisotropic spacing."""
    from pyfantom_amd.physics import LinearElasticity
    for nu in (0.3, 1 / 3):
        phys = LinearElasticity(E=1.0, nu=nu)
        for h in ((1, 1, 1), (0.5, 1, 2)):
            X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                         float) * np.array(h)[None, :]
            Kp = phys.K(X)
            Ke = oracle.k0_exact(3, nu, *h)
            assert np.abs(Kp - Ke).max() <= 1e-14 * np.abs(Ke).max()
        X2 = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
        assert np.abs(phys.K(X2) - oracle.k0_exact(2, nu)).max() <= 1e-14


def test_k0_golden(oracle):
    """Committed K0 fixtures (hex8 and quad4, nu = 0.3)."""
    d = np.load(os.path.join(GOLDEN, "k0.npz"))
    assert np.abs(d["hex8"] - oracle.k0_exact(3, 0.3)).max() < 1e-15
    assert np.abs(d["quad4"] - oracle.k0_exact(2, 0.3)).max() < 1e-15


# --------------------------------------------------------------------------- A5 A6 A14
def _mesh_case(oracle, nel, seed=0, bc=True):
    dim = len(nel)
    K0 = oracle.k0_exact(dim, 0.3)
    el = oracle.elements(nel)
    n_nodes = int(np.prod([n + 1 for n in nel]))
    dof = dim
    rng = _rng(seed)
    E = rng.uniform(0.01, 1.0, len(el)) ** 3
    u = rng.standard_normal(n_nodes * dof)
    con = np.zeros(n_nodes * dof, bool)
    if bc:
        con[rng.choice(n_nodes * dof, size=max(1, n_nodes * dof // 7), replace=False)] = True
    return K0, el, E, u, con


@pytest.mark.parametrize("nel", [(2, 3, 2), (4, 2, 3), (3, 4), (5, 2)])
@pytest.mark.parametrize("bc", [False, True])
def test_matvec_vs_assembly(oracle, nel, bc):
    """Node-basis gather vs scipy COO assembly with the Dirichlet convention."""
    K0, el, E, u, con = _mesh_case(oracle, nel, bc=bc)
    K = oracle.assemble(K0, el, E, len(u), con if bc else None)
    y = oracle.matvec(K0, el, E, u, con if bc else None)
    assert np.abs(y - K @ u).max() <= 1e-13 * np.abs(K @ u).max()
    if bc:
        assert np.array_equal(y[con], u[con])


def test_matvec_symmetry(oracle):
    """<Ku, v> = <u, Kv> on a mid-size grid with BCs."""
    K0, el, E, u, con = _mesh_case(oracle, (6, 5, 4))
    v = _rng(3).standard_normal(len(u))
    a = np.dot(oracle.matvec(K0, el, E, u, con), v)
    b = np.dot(u, oracle.matvec(K0, el, E, v, con))
    assert abs(a - b) <= 1e-12 * abs(a)


@pytest.mark.parametrize("nel", [(2, 3, 2), (3, 4)])
def test_diag_vs_assembly(oracle, nel):
    K0, el, E, u, con = _mesh_case(oracle, nel)
    K = oracle.assemble(K0, el, E, len(u), con)
    d = oracle.diag(K0, el, E, len(u), con)
    assert np.abs(d - K.diagonal()).max() <= 1e-14 * np.abs(d).max()


@pytest.mark.parametrize("nel", [(2, 3, 2), (3, 4)])
def test_dc_vs_dense_and_fd(oracle, nel):
    """-u_e^T K0 u_e vs einsum, and vs a finite difference of u^T K(E) u in E_e."""
    K0, el, E, u, _ = _mesh_case(oracle, nel)
    dof = len(nel)
    dc = oracle.process_dk(K0, el, u)
    ue = u[(el[:, :, None] * dof + np.arange(dof)).reshape(len(el), -1)]
    assert np.abs(dc + np.einsum("ei,ij,ej->e", ue, K0, ue)).max() <= 1e-13 * np.abs(dc).max()
    e0, h = 3, 1e-3
    K1 = oracle.assemble(K0, el, E, len(u))
    E2 = E.copy()
    E2[e0] += h
    K2 = oracle.assemble(K0, el, E2, len(u))
    fd = (u @ (K2 @ u) - u @ (K1 @ u)) / h
    assert abs(fd + dc[e0]) <= 1e-9 * abs(dc[e0])


# --------------------------------------------------------------------------- A7 A8
@pytest.mark.parametrize("nel", [(4, 2, 6), (6, 4), (2, 2, 2)])
@pytest.mark.parametrize("normalize", [True, False])
def test_transfers_vs_explicit(oracle, nel, normalize):
    """Restriction / prolongation kernels vs the explicit sparse R; P == R^T."""
    dof = len(nel)
    R = oracle.transfer_matrix(nel, dof, normalize)
    rng = _rng(1)
    v = rng.standard_normal(R.shape[1])
    w = rng.standard_normal(R.shape[0])
    assert np.abs(oracle.restrict(v, nel, dof, normalize) - R @ v).max() < 1e-13
    assert np.abs(oracle.prolong(w, nel, dof, normalize) - R.T @ w).max() < 1e-13


def test_transfer_constants(oracle):
    """Normalised R maps constants to constants (rows sum to 1); unnormalised
    P_lin interpolates constants exactly."""
    nel, dof = (4, 2, 6), 3
    R = oracle.transfer_matrix(nel, dof, True)
    assert np.allclose(np.asarray(R.sum(axis=1)).ravel(), 1.0)
    P = oracle.transfer_matrix(nel, dof, False).T
    nc = R.shape[0]
    assert np.allclose(P @ np.ones(nc), 1.0)


def test_galerkin_product(oracle):
    """R A R^T (sparse) vs a dense product."""
    nel, dof = (2, 2, 2), 3
    K0, el, E, u, con = _mesh_case(oracle, nel)
    A = oracle.assemble(K0, el, E, len(u), con)
    Ac = oracle.galerkin(A, nel, dof).toarray()
    R = oracle.transfer_matrix(nel, dof, True).toarray()
    assert np.abs(Ac - R @ A.toarray() @ R.T).max() < 1e-13


# --------------------------------------------------------------------------- A15
@pytest.mark.parametrize("nel,r", [((5, 4, 3), 1.5), ((6, 5), 1.5), ((4, 4, 4), 2.0), ((7, 3), 2.5)])
def test_filter_vs_dense(oracle, nel, r):
    offs, w = oracle.filter_taps(len(nel), r)
    H = oracle.filter_dense(nel, offs, w)
    x = _rng(2).uniform(0, 1, H.shape[0])
    y = _rng(3).standard_normal(H.shape[0])
    assert np.abs(oracle.filter_fwd(nel, offs, w, x) - H @ x).max() < 1e-14
    assert np.abs(oracle.filter_adj(nel, offs, w, y) - H.T @ y).max() < 1e-14
    assert np.allclose(H @ np.ones(H.shape[0]), 1.0)


def test_filter_tap_counts(oracle):
    """r = 1.5 -> 19 taps (3D) / 9 (2D); r = 2.0 -> 27 (3D)."""
    assert len(oracle.filter_taps(3, 1.5)[1]) == 19
    assert len(oracle.filter_taps(2, 1.5)[1]) == 9
    assert len(oracle.filter_taps(3, 2.0)[1]) == 27


# --------------------------------------------------------------------------- A16
def test_oc_hand_case(oracle):
    """4 elements: the bisection finds the multiplier that meets the volume."""
    x = np.array([0.5, 0.5, 0.5, 0.5])
    df = -np.array([4.0, 1.0, 0.25, 1.0])
    dg = np.full(4, 0.25)
    xn, lm = oracle.oc_update(x, df, dg, 0.5, 0.2)
    # ocP = x*sqrt(-df/dg) = [2, 1, 0.5, 1]; x_new = clip(ocP/l, 0.3, 0.7).
    # The volume sum(x_new)/4 = 0.5 holds only at l = 2, where element 0 is
    # capped at 0.7, element 2 floored at 0.3 and elements 1, 3 sit at 0.5.
    assert np.allclose(xn, [0.7, 0.5, 0.3, 0.5], atol=1e-7)
    assert abs(lm - 2.0) < 1e-7
    assert abs(np.dot(dg, xn) - 0.5) < 1e-7


def test_oc_bisection_invariants(oracle):
    rng = _rng(4)
    x = rng.uniform(0, 1, 1000)
    df = -rng.uniform(0.01, 5, 1000)
    dg = np.full(1000, 1e-3)
    xn, lm = oracle.oc_update(x, df, dg, 0.4, 0.2)
    assert np.all(xn >= np.maximum(0, x - 0.2) - 1e-15) and np.all(xn <= np.minimum(1, x + 0.2) + 1e-15)
    assert abs(np.dot(dg, xn) - 0.4) < 1e-6
    # monotone in the multiplier: a larger multiplier never increases a variable
    ocp = x * np.sqrt(-df / dg)
    lo = np.maximum(0, np.maximum(np.clip(x - 0.2, 0, 1), np.minimum(1, np.minimum(np.clip(x + 0.2, 0, 1), ocp / (lm * 1.01)))))
    assert np.all(lo <= xn + 1e-15)


# --------------------------------------------------------------------------- A9-A13
@pytest.mark.parametrize("nel,nlev", [((8, 4, 4), 2), ((16, 8), 3)])
def test_mgpcg_vs_direct(oracle, nel, nlev):
    """GMG-PCG to 1e-12 equals a sparse direct solve; compliance identity."""
    import scipy.sparse.linalg as spla
    dim = len(nel)
    K0 = oracle.k0_exact(dim, 0.3)
    el = oracle.elements(nel)
    n_nodes = int(np.prod([n + 1 for n in nel]))
    E = _rng(0).uniform(0.01, 1, len(el)) ** 3
    con = np.zeros(n_nodes * dim, bool)
    # clamp the x = 0 face
    if dim == 3:
        nx, nz = nel[0] + 1, nel[2] + 1
        for j in range(nel[1] + 1):
            for k in range(nz):
                n = k + nz * 0 + nz * nx * j
                con[3 * n:3 * n + 3] = True
    else:
        nx = nel[0] + 1
        for j in range(nel[1] + 1):
            con[2 * (j * nx):2 * (j * nx) + 2] = True
    f = np.zeros(n_nodes * dim)
    f[-dim + 1] = -1.0
    u, it, res = oracle.mgpcg_solve(nel, K0, E, con, f, n_level=nlev, n_smooth=2, omega=0.6, tol=1e-12)
    assert it > 0 and res < 1e-12
    K = oracle.assemble(K0, el, E, len(f), con)
    ud = spla.spsolve(K.tocsc(), f)
    assert np.abs(u - ud).max() <= 1e-9 * np.abs(ud).max()
    c1, c2 = f @ u, u @ (K @ u)
    assert abs(c1 - c2) <= 1e-10 * abs(c1)
    assert abs(c1 - f @ ud) <= 1e-10 * abs(c1)


# --------------------------------------------------------------------------- A4
def test_penalize_grad_fd(oracle):
    rho = np.linspace(0.05, 0.95, 19)
    h = 1e-6
    for hv in (True, False):
        g = oracle.penalize_grad_reference(rho, heavyside=hv)
        fd = (oracle.penalize_reference(rho + h, void=0, heavyside=hv)
              - oracle.penalize_reference(rho - h, void=0, heavyside=hv)) / (2 * h)
        assert np.abs(g - fd).max() < 1e-7
    g = oracle.penalize_grad_modified(rho)
    fd = (oracle.penalize_modified(rho + h) - oracle.penalize_modified(rho - h)) / (2 * h)
    assert np.abs(g - fd).max() < 1e-7


@pytest.mark.parametrize("dim,h", [(3, (1.0, 1.0, 1.0)), (3, (0.5, 1.0, 2.0)), (2, (1.0, 1.0)), (2, (0.5, 2.0))])
def test_b_center_and_d_iso(oracle, dim, h):
    """Patch test: a linear displacement u = G x gives the exact constant
    strain through B (Voigt, engineering shear); the energy of that field
    under the exactly integrated K0 is V eps^T D eps, which ties D's
    convention to K0's."""
    B = oracle.b_center(dim, *h)
    D = oracle.d_iso(dim, 0.3)
    corners = ([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
               if dim == 3 else [(0, 0), (1, 0), (1, 1), (0, 1)])
    X = np.array(corners, float) * np.array(h)
    G = np.random.default_rng(3).standard_normal((dim, dim))
    u = (X @ G.T).ravel()
    if dim == 3:
        eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])
    else:
        eps = np.array([G[0, 0], G[1, 1], G[0, 1] + G[1, 0]])
    assert np.allclose(B @ u, eps, rtol=0, atol=1e-13)
    K0 = oracle.k0_exact(dim, 0.3, *h)
    V = float(np.prod(h))
    assert np.isclose(u @ K0 @ u, V * eps @ D @ eps, rtol=1e-12)


def test_fea_locals_identities(oracle):
    """Element energies sum to 1/2 u^T K u of the assembled matrix; a uniaxial
    stress state has von Mises stress |sigma_xx|."""
    nel = (3, 2, 2)
    K0 = oracle.k0_exact(3, 0.3)
    el = oracle.elements(nel)
    rng = np.random.default_rng(1)
    w = rng.uniform(0.1, 1.0, len(el))
    u = rng.standard_normal(int(np.prod([n + 1 for n in nel])) * 3)
    A = oracle.assemble(K0, el, w, len(u))
    strain, stress, energy, vm = oracle.fea_locals(K0, oracle.d_iso(3, 0.3), oracle.b_center(3), el, w, u)
    assert np.isclose(energy.sum(), 0.5 * u @ (A @ u), rtol=1e-12)
    # uniaxial: sigma = (s, 0, ...) -> von Mises = |s|; set D so that stress = strain
    s_strain, s_stress, _, s_vm = oracle.fea_locals(K0, np.eye(6), oracle.b_center(3), el, w,
                                                     (oracle.nodes(nel, nel) * np.array([0.7, 0, 0])).ravel())
    assert np.allclose(s_stress[:, 0], 0.7) and np.allclose(s_stress[:, 1:], 0, atol=1e-14)
    assert np.allclose(s_vm, 0.7)


@pytest.mark.parametrize("dim,h", [(3, (1.0, 1.0, 1.0)), (3, (0.5, 1.0, 2.0)), (2, (1.0, 1.0))])
def test_product_locals_match_oracle(oracle, dim, h):
    """The host physics' constitutive matrix and averaged B (what the
    post-processing kernel is fed) equal the oracle's closed forms."""
    from pyfantom_amd.physics import LinearElasticity
    corners = ([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
               if dim == 3 else [(0, 0), (1, 0), (1, 1), (0, 1)])
    X = np.array(corners, float) * np.array(h)
    D, B = LinearElasticity(E=1.0, nu=0.3).locals(X)
    assert np.allclose(D, oracle.d_iso(dim, 0.3), rtol=1e-14, atol=0)
    assert np.allclose(B, oracle.b_center(dim, *h), rtol=0, atol=1e-14)


# ------------------------------------------------------- explicit CSR assembly
def _clamp_x0(nel):
    """Per-DOF constraint flags clamping the x = 0 node face (3D or 2D)."""
    X = np.asarray(nel)
    nodes_per_axis = X + 1
    dim = len(nel)
    n_nodes = int(np.prod(nodes_per_axis))
    con = np.zeros(n_nodes * dim, bool)
    P = None
    if dim == 3:
        # node n(i, j, k) = k + nz i + nz nx j; x index i
        nx, nz = nodes_per_axis[0], nodes_per_axis[2]
        ids = np.arange(n_nodes)
        P = (ids % (nx * nz)) // nz
    else:
        nx = nodes_per_axis[0]
        P = np.arange(n_nodes) % nx
    for d in range(dim):
        con[np.flatnonzero(P == 0) * dim + d] = True
    return con


@pytest.mark.parametrize("nel", [(1, 1, 1), (3, 2, 4), (2, 3), (4, 1)])
def test_csr_assemble_pattern_and_values(oracle, nel):
    """Nonzero counts against closed forms, the dense matrix against the scipy
    assembly, K u against the C matvec: a wrong pattern rule (dropped explicit
    zeros, constrained columns kept, a missing unit row) changes the counts; a
    wrong local index or weight changes the values."""
    dim = len(nel)
    K0 = oracle.k0_exact(dim, 0.3)
    el = oracle.elements(nel)
    rng = np.random.default_rng(3)
    E = rng.uniform(0.1, 1.0, len(el))
    n = np.asarray(nel) + 1
    n_dof = int(np.prod(n)) * dim
    # free grid: every pair of nodes within one step per axis, dim x dim DOFs
    indptr, idx, data = oracle.csr_assemble(K0, el, E, n_dof)
    assert indptr[-1] == dim * dim * np.prod(3 * n - 2)
    # clamped x = 0 face: free rows see x-neighbours in [1, nx - 1] only
    con = _clamp_x0(nel)
    indptr_c, idx_c, data_c = oracle.csr_assemble(K0, el, E, n_dof, con)
    m = np.concatenate([[3 * (n[0] - 1) - 2], 3 * n[1:] - 2])
    assert indptr_c[-1] == dim * dim * np.prod(m) + con.sum()
    import scipy.sparse as sps
    for (p, j, x), c in (((indptr, idx, data), None), ((indptr_c, idx_c, data_c), con)):
        K = sps.csr_matrix((x, j, p), shape=(n_dof, n_dof))
        for r in range(n_dof):
            assert np.all(np.diff(j[p[r]:p[r + 1]]) > 0)  # ascending, no duplicates
        Kd = K.toarray()
        A = oracle.assemble(K0, el, E, n_dof, c).toarray()
        assert np.abs(Kd - A).max() <= 1e-14 * np.abs(A).max()
        pat = sps.csr_matrix((np.ones_like(x), j, p), shape=(n_dof, n_dof)).toarray()
        assert np.all(pat[A != 0] == 1)  # every nonzero is in the pattern
        assert np.array_equal(pat, pat.T)
        u = rng.standard_normal(n_dof)
        assert np.allclose(K @ u, oracle.matvec(K0, el, E, u, c), rtol=0, atol=1e-13 * np.abs(u).max())
    # constrained rows are unit rows
    rows_c = np.flatnonzero(con)
    assert np.array_equal(np.diff(indptr_c)[rows_c], np.ones(len(rows_c), np.int64))
    assert np.array_equal(idx_c[indptr_c[rows_c]], rows_c) and np.all(data_c[indptr_c[rows_c]] == 1.0)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_chebyshev4_weights_match_reference_recurrence(oracle, n):
    """The product form the drivers run (n Jacobi sweeps with weights
    fantom_chebyshev4_weights(n) / rho) is the reference's fourth-kind
    three-term smoother: on the scalar problem a = t, D = 1, b = 0 both give
    the error p_n(t) e_0 for every t; p_n(1) = (-1)^n / (2n + 1) and |p_n| <= 1
    on [0, 1]."""
    import ctypes
    from pyfantom_amd import _lib
    w = (ctypes.c_double * n)()
    _lib.check(_lib.load().fantom_chebyshev4_weights(n, w))
    w = np.array(list(w))
    assert np.all(np.diff(w) > 0)
    t = np.linspace(0.0, 1.3, 131)
    ref = oracle.chebyshev4_smoother(np.ones_like(t), np.zeros_like(t), lambda u: t * u, np.ones_like(t), n, 1.0)
    prod = np.ones_like(t)
    for wi in w:
        prod = prod - wi * t * prod      # one sweep: e <- (1 - w t) e
    assert np.abs(prod - ref).max() <= 1e-13
    assert abs(ref[100] - (-1) ** n / (2 * n + 1)) <= 1e-14
    assert np.abs(ref[:101]).max() <= 1.0 + 1e-14


def test_estimate_rho_power_method(oracle):
    """The power method of the reference converges to the top eigenvalue of
    D^-1 A (a generalised symmetric eigenproblem, checked with scipy) on a
    small clamped hex8 block."""
    import scipy.linalg as sla
    nel = (4, 2, 2)
    K0 = oracle.k0_exact(3, 0.3)
    el = oracle.elements(nel)
    E = np.random.default_rng(2).uniform(0.01, 1.0, len(el)) ** 3
    n = int(np.prod([v + 1 for v in nel])) * 3
    con = np.zeros(n, bool)
    con[:9] = True
    A = oracle.assemble(K0, el, E, n, con).toarray()
    d = np.diag(A)
    lam = sla.eigh(A, np.diag(d), eigvals_only=True)[-1]
    est = oracle.estimate_rho(lambda u: A @ u, 1.0 / d, np.random.default_rng(0).random(n), 200)
    assert abs(est - lam) <= 1e-6 * lam
