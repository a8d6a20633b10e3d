"""Mesh arrays generated on the device (fantom_mesh_elements,
fantom_mesh_nodes_*) and the device boundary-condition / load path of
FiniteElement: connectivity bit-exact against the committed SHA-256 fixture
and the oracle, coordinates bit-exact against numpy's linspace/meshgrid,
nearest-node lookup against brute force, point forces against numpy's
buffered fancy-index add."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _mesh(pf, nel, L, dtype=np.float64):
    if len(nel) == 3:
        return pf.StructuredMesh3D(*nel, *L, dtype=dtype, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    return pf.StructuredMesh2D(*nel, *L, dtype=dtype, physics=pf.LinearElasticity(E=1.0, nu=0.3))


def test_elements_match_golden_hashes():
    import pyfantom_amd as pf
    _torch()
    with open(os.path.join(GOLDEN, "connectivity_sha256.json")) as f:
        ref = json.load(f)
    for key, h in ref.items():
        nel = tuple(int(v) for v in key.split("x"))
        print(f"elements {key}", flush=True)
        el = _mesh(pf, nel, nel).elements
        assert el.dtype.itemsize == 4 and tuple(el.shape) == (int(np.prod(nel)), 2 ** len(nel))
        assert hashlib.sha256(el.cpu().numpy().tobytes()).hexdigest() == h, key


@pytest.mark.parametrize("nel,L", [((3, 2, 4), (3.0, 2.0, 4.0)), ((7, 5, 3), (1.0, 0.7, 0.3)),
                                   ((5, 4), (5.0, 4.0)), ((9, 6), (1.3, 0.1))])
def test_elements_and_nodes_match_oracle(oracle, nel, L):
    import pyfantom_amd as pf
    _torch()
    m = _mesh(pf, nel, L)
    assert np.array_equal(m.elements.cpu().numpy(), oracle.elements(nel))
    X = m.nodes.cpu().numpy()
    assert np.array_equal(X, oracle.nodes(nel, L))   # bit-exact linspace / meshgrid
    m32 = _mesh(pf, nel, L, dtype=np.float32)
    assert m32.nodes.dtype.itemsize == 4
    assert np.array_equal(m32.nodes.cpu().numpy(), oracle.nodes(nel, L, np.float32))


@pytest.mark.parametrize("dim", [3, 2])
def test_nearest_nodes_and_point_forces(oracle, dim):
    import pyfantom_amd as pf
    torch = _torch()
    nel = (6, 4, 5) if dim == 3 else (8, 5)
    L = (3.0, 2.0, 2.5) if dim == 3 else (4.0, 2.5)
    mesh = _mesh(pf, nel, L)
    K = pf.StructuredStiffnessKernel(mesh)
    solver = pf.CG(K, maxiter=10)
    FE = pf.FiniteElement(mesh, K, solver)
    X = oracle.nodes(nel, L)
    rng = np.random.default_rng(5)
    # random points off the cell midplanes (no distance ties)
    P = rng.uniform(0, 1, (40, dim)) * np.array(L)
    h = np.array(L) / np.array(nel)
    frac = (P / h) % 1.0
    P = P[np.all(np.abs(frac - 0.5) > 0.05, axis=1)]
    brute = np.argmin(((P[:, None, :] - X[None, :, :]) ** 2).sum(-1), axis=1)
    got = FE.nearest_nodes(P).cpu().numpy()
    assert np.array_equal(got, brute)
    # point forces: a node listed twice receives its last force (numpy's
    # buffered rhs[idx] += f), Dirichlet DOFs carry their prescribed value
    ids = np.array([3, 7, 3, 11])
    F = rng.standard_normal((4, dim))
    FE.add_dirichlet_boundary_condition(node_ids=np.array([0, 1]), rhs=0.25)
    FE.add_point_forces(forces=F, node_ids=ids)
    ref = np.zeros(K.shape[0])
    for i in range(dim):
        ref[ids * dim + i] += F[:, i]
    ref[np.array([0, 1])[:, None] * dim + np.arange(dim)] = 0.25
    assert np.array_equal(FE.rhs.cpu().numpy(), ref)
    con = K.constraints.cpu().numpy()
    assert con.sum() == 2 * dim and con[:2 * dim].all()
    # positions resolve to the same nodes as their ids
    FE2 = pf.FiniteElement(mesh, pf.StructuredStiffnessKernel(mesh), solver)
    FE2.add_dirichlet_boundary_condition(positions=X[[0, 1]], rhs=0.25)
    FE2.add_point_forces(forces=F, positions=X[ids])
    assert np.array_equal(FE2.rhs.cpu().numpy(), ref)
