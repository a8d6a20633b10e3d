"""The benchmark workloads as reference-shaped problem set-ups.

Each builder returns (mesh, kernel, solver, FE, filter, problem), assembled
through the same public classes a pyFANTOM script uses
(docs/source/examples/3d_topology_optimization.rst:197-312): structured hex8
mesh with unit elements, E0 = 1, nu = 0.3, modified SIMP (Emin = 1e-9, p = 3),
density filter, Dirichlet nodes and point loads through FiniteElement.

cantilever: x = 0 face clamped in all DOFs; load 'edge' = total -1 in y over
the nodes of the bottom edge (y = 0) of the x = L face; load 'random' = a unit
downward point load at a seeded random node of the x = L face.
bridge: the four bottom corner nodes (y = 0, x in {0, L}, z in {0, W}) pinned
in all DOFs, total load -1 in y spread evenly over the nodes of the top
surface y = H.
"""
from __future__ import annotations

import numpy as np

from . import geom, physics, problem as _problem, solvers, stiffness, fe as _fe


def _common(nel, r, vf, solver_kw):
    mesh = geom.StructuredMesh3D(*nel, *nel, physics=physics.LinearElasticity(E=1.0, nu=0.3))
    K = stiffness.StructuredStiffnessKernel(mesh)
    kw = dict(maxiter=1000, tol=1e-8, n_smooth=2, omega=0.6, n_level=4, cycle="V")
    kw.update(solver_kw)
    solver = solvers.MultiGrid(mesh, K, **kw)
    FE = _fe.FiniteElement(mesh, K, solver)
    return mesh, K, solver, FE


def _finish(mesh, K, solver, FE, r, vf):
    filt = geom.StructuredFilter3D(mesh, r)
    prob = _problem.MinimumCompliance(FE, filt, E_mul=[1.0], void=1e-9, penalty=3.0, volume_fraction=[vf],
                                      heavyside=False, law="modified")
    return mesh, K, solver, FE, filt, prob


def cantilever_nodes(nel):
    """(clamped node ids on x = 0, edge-load node ids on x = L, y = 0)."""
    nx, ny, nz = (n + 1 for n in nel)
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    return (k + nz * nx * j).ravel(), np.arange(nz) + nz * (nx - 1)


def cantilever(nel, vf=0.3, r=1.5, load="edge", seed=0, **solver_kw):
    mesh, K, solver, FE = _common(nel, r, vf, solver_kw)
    nx, ny, nz = (n + 1 for n in nel)
    clamp, edge = cantilever_nodes(nel)
    FE.add_dirichlet_boundary_condition(node_ids=clamp, rhs=0)
    if load == "edge":
        FE.add_point_forces(forces=np.array([[0.0, -1.0 / nz, 0.0]]), node_ids=edge)
    elif load == "random":
        rng = np.random.default_rng(seed)
        jj, kk = int(rng.integers(0, ny)), int(rng.integers(0, nz))
        FE.add_point_forces(forces=np.array([[0.0, -1.0, 0.0]]),
                            node_ids=np.array([kk + nz * (nx - 1) + nz * nx * jj]))
    else:
        raise ValueError(f"unknown load {load!r}")
    return _finish(mesh, K, solver, FE, r, vf)


def bridge_nodes(nel):
    """(four pinned bottom corner node ids, top-surface node ids)."""
    nx, ny, nz = (n + 1 for n in nel)
    corners = np.array([k + nz * i for i in (0, nx - 1) for k in (0, nz - 1)])
    i, k = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    return corners, (k + nz * i + nz * nx * (ny - 1)).ravel()


def bridge(nel, vf=0.15, r=2.0, **solver_kw):
    mesh, K, solver, FE = _common(nel, r, vf, solver_kw)
    corners, top = bridge_nodes(nel)
    FE.add_dirichlet_boundary_condition(node_ids=corners, rhs=0)
    FE.add_point_forces(forces=np.array([[0.0, -1.0 / top.size, 0.0]]), node_ids=top)
    return _finish(mesh, K, solver, FE, r, vf)


def seeded_moduli(n_el, seed=0, Emin=1e-9):
    """The isolated-solve field: x ~ U[0.01, 1] (default_rng(seed)), E = Emin + x^3 (1 - Emin)."""
    x = np.random.default_rng(seed).uniform(0.01, 1.0, n_el)
    return Emin + x ** 3 * (1 - Emin)
