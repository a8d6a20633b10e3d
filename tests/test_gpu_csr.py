"""Explicit CSR assembly (StructuredStiffnessKernel.construct,
fantom_csr_row_nnz + fantom_csr_assemble_f64) against the oracle's element
scatter: row pointers and column indices bit-exact, values to 1e-14 of the
largest entry; K u of the CSR matrix against the matrix-free K.u; the sparse
product K @ B; pattern reuse across designs and rebuild on new constraints."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _kernel(pf, nel):
    phys = pf.LinearElasticity(E=1.0, nu=0.3)
    if len(nel) == 3:
        mesh = pf.StructuredMesh3D(*nel, *nel, physics=phys)
    else:
        mesh = pf.StructuredMesh2D(*nel, *nel, physics=phys)
    return mesh, pf.StructuredStiffnessKernel(mesh)


def _clamp_x0(nel):
    dim = len(nel)
    n = np.asarray(nel) + 1
    ids = np.arange(int(np.prod(n)))
    xi = (ids % (n[0] * n[2])) // n[2] if dim == 3 else ids % n[0]
    nodes = np.flatnonzero(xi == 0)
    return (nodes[:, None] * dim + np.arange(dim)[None, :]).ravel()


@pytest.mark.parametrize("nel,clamp", [((5, 3, 4), False), ((5, 3, 4), True), ((9, 2, 3), True),
                                       ((1, 1, 1), True), ((12, 7), False), ((12, 7), True)])
def test_construct_matches_oracle(oracle, nel, clamp):
    import pyfantom_amd as pf
    torch = _torch()
    print(f"csr {nel} clamp={clamp}", flush=True)
    mesh, K = _kernel(pf, nel)
    rng = np.random.default_rng(7)
    E = rng.uniform(0.01, 1.0, K.n_elements)
    con = None
    if clamp:
        ids = _clamp_x0(nel)
        K.add_constraints(torch.from_numpy(ids).cuda())
        con = np.zeros(K.shape[0], bool)
        con[ids] = True
    A = K.construct(torch.from_numpy(E).cuda())
    torch.cuda.synchronize()
    K0 = np.asarray(K.K_single, np.float64)
    p_ref, j_ref, x_ref = oracle.csr_assemble(K0, oracle.elements(nel), E, K.shape[0], con)
    p = A.crow_indices().cpu().numpy().astype(np.int64)
    j = A.col_indices().cpu().numpy()
    x = A.values().cpu().numpy()
    assert np.array_equal(p, p_ref)
    assert np.array_equal(j, j_ref)
    assert np.abs(x - x_ref).max() <= 1e-14 * np.abs(x_ref).max()
    # the explicit matrix acts as the matrix-free operator
    K.set_rho(torch.from_numpy(E).cuda())
    u = torch.from_numpy(rng.standard_normal(K.shape[0])).cuda()
    y_free = K.dot(u).clone()
    y_csr = (A @ u[:, None]).squeeze(1)
    assert torch.allclose(y_csr, y_free, rtol=0, atol=1e-13 * float(y_free.abs().max()))


def test_construct_reuses_pattern_and_tracks_constraints(oracle):
    import pyfantom_amd as pf
    torch = _torch()
    nel = (6, 4, 3)
    mesh, K = _kernel(pf, nel)
    rng = np.random.default_rng(8)
    K0 = np.asarray(K.K_single, np.float64)
    el = oracle.elements(nel)
    E1, E2 = rng.uniform(0.1, 1, K.n_elements), rng.uniform(0.1, 1, K.n_elements)
    A1 = K.construct(torch.from_numpy(E1).cuda())
    x1 = A1.values().cpu().numpy().copy()
    A2 = K.construct(torch.from_numpy(E2).cuda())
    assert A2.crow_indices().data_ptr() == A1.crow_indices().data_ptr()  # pattern kept
    assert np.abs(x1 - oracle.csr_assemble(K0, el, E1, K.shape[0])[2]).max() < 1e-14 * np.abs(x1).max()
    x2 = A2.values().cpu().numpy()
    assert np.abs(x2 - oracle.csr_assemble(K0, el, E2, K.shape[0])[2]).max() < 1e-14 * np.abs(x2).max()
    ids = _clamp_x0(nel)
    K.add_constraints(torch.from_numpy(ids).cuda())
    A3 = K.construct(torch.from_numpy(E2).cuda())
    con = np.zeros(K.shape[0], bool)
    con[ids] = True
    p_ref, j_ref, _ = oracle.csr_assemble(K0, el, E2, K.shape[0], con)
    assert np.array_equal(A3.crow_indices().cpu().numpy().astype(np.int64), p_ref)
    assert np.array_equal(A3.col_indices().cpu().numpy(), j_ref)


def test_dot_sparse_matrix(oracle):
    """K @ B for a sparse B (stiffness/CUDA/_FEA.py:623-673): constrained rows
    take B's rows, free rows sum over free DOFs only."""
    import scipy.sparse as sps
    import pyfantom_amd as pf
    torch = _torch()
    nel = (4, 3, 3)
    mesh, K = _kernel(pf, nel)
    ids = _clamp_x0(nel)
    K.add_constraints(torch.from_numpy(ids).cuda())
    con = np.zeros(K.shape[0], bool)
    con[ids] = True
    rng = np.random.default_rng(9)
    E = rng.uniform(0.1, 1, K.n_elements)
    K.set_rho(torch.from_numpy(E).cuda())
    B = sps.random(K.shape[0], 17, density=0.05, random_state=3, format="csr")
    Bt = torch.sparse_csr_tensor(torch.from_numpy(B.indptr.astype(np.int64)), torch.from_numpy(B.indices.astype(np.int64)),
                                 torch.from_numpy(B.data), size=B.shape).cuda()
    C = K.dot(Bt).to_dense().cpu().numpy()
    A = oracle.assemble(np.asarray(K.K_single, np.float64), oracle.elements(nel), E, K.shape[0], con)
    ref = (A @ B).toarray()
    assert np.abs(C - ref).max() <= 1e-13 * np.abs(ref).max()
