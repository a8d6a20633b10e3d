"""FiniteElement: boundary conditions, loads and ``solve(rho)``.

Mirrors FiniteElement/CPU/FiniteElement.py:89-445 (and the CUDA twin): the
same constructor, ``add_dirichlet_boundary_condition`` (node ids or
positions, optional per-DOF flags, prescribed values), ``add_point_forces``,
``reset_forces``, ``reset_dirichlet_boundary_conditions`` and
``solve(rho) -> (U, residual)``.  Positions are mapped to nodes by nearest
grid node (the reference queries a KDTree over the node coordinates; on a
structured grid the nearest node is the rounded grid index).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


class FiniteElement:
    def __init__(self, mesh, kernel, solver):
        self.mesh = mesh
        self.kernel = kernel
        self.solver = solver
        self.dtype = mesh.dtype
        n_dof = kernel.shape[0]
        self.rhs = torch.zeros(n_dof, dtype=torch.float64, device=kernel.device)
        self.d_rhs = torch.full((n_dof,), float("nan"), dtype=torch.float64, device=kernel.device)
        self.nel = mesh.n_elements
        self.is_3D = len(mesh.nel) == 3
        self.dof = mesh.dof

    def nearest_nodes(self, positions) -> np.ndarray:
        """Node ids nearest to the given coordinates (KDTree query semantics)."""
        P = np.asarray(positions.detach().cpu().numpy() if isinstance(positions, torch.Tensor) else positions,
                       dtype=np.float64).reshape(-1, len(self.mesh.nel))
        h = np.asarray(self.mesh.spacing, dtype=np.float64)
        idx = np.rint(P / h[None, :]).astype(np.int64)
        idx = np.clip(idx, 0, np.asarray(self.mesh.nel, dtype=np.int64)[None, :])
        nx = int(self.mesh.nel[0]) + 1
        if self.is_3D:
            nz = int(self.mesh.nel[2]) + 1
            return idx[:, 2] + nz * idx[:, 0] + nz * nx * idx[:, 1]
        return idx[:, 1] * nx + idx[:, 0]

    def _ids(self, node_ids):
        return np.asarray(node_ids.detach().cpu().numpy() if isinstance(node_ids, torch.Tensor) else node_ids,
                          dtype=np.int64).reshape(-1)

    def add_dirichlet_boundary_condition(self, node_ids=None, positions=None, dofs=None, rhs=0):
        if node_ids is None and positions is None:
            raise ValueError("Either node_ids or positions must be provided.")
        if node_ids is not None and positions is not None:
            raise ValueError("Only one of node_ids or positions should be provided.")
        if positions is not None:
            return self.add_dirichlet_boundary_condition(node_ids=self.nearest_nodes(positions), dofs=dofs, rhs=rhs)
        node_ids = self._ids(node_ids)
        n_con = node_ids.shape[0]
        rhs_arr = None
        if isinstance(rhs, (np.ndarray, torch.Tensor)):
            rhs_arr = np.asarray(rhs.detach().cpu().numpy() if isinstance(rhs, torch.Tensor) else rhs, dtype=np.float64)
            if rhs_arr.shape[0] != n_con:
                raise ValueError("rhs must have the same length as node_ids or positions.")
        if dofs is not None:
            dofs = np.asarray(dofs.detach().cpu().numpy() if isinstance(dofs, torch.Tensor) else dofs)
            if dofs.ndim == 1:
                dofs = dofs[None, :]
            if dofs.shape[0] != n_con and dofs.shape[0] != 1:
                raise ValueError("dofs must have the same length as node_ids.")
            if dofs.shape[0] == 1:
                dofs = np.tile(dofs, (n_con, 1))
        dev = self.kernel.device
        d_rhs = self.d_rhs.cpu().numpy()
        for i in range(self.dof):
            sel = np.ones(n_con, dtype=bool) if dofs is None else (dofs[:, i] == 1)
            cons = node_ids[sel] * self.dof + i
            self.kernel.add_constraints(torch.as_tensor(cons, device=dev))
            if rhs_arr is not None:
                d_rhs[cons] = rhs_arr[sel, i]
            else:
                d_rhs[cons] = rhs + np.nan_to_num(d_rhs[cons], nan=0)
        self.d_rhs = torch.from_numpy(d_rhs).to(dev)

    def add_neumann_boundary_condition(self, **kwargs):
        raise NotImplementedError("Neumann boundary conditions are not implemented in this version of FiniteElement.")

    def add_point_forces(self, forces, node_ids=None, positions=None):
        if node_ids is None and positions is None:
            raise ValueError("Either node_ids or positions must be provided.")
        if node_ids is not None and positions is not None:
            raise ValueError("Only one of node_ids or positions should be provided.")
        if positions is not None:
            return self.add_point_forces(forces=forces, node_ids=self.nearest_nodes(positions))
        node_ids = self._ids(node_ids)
        forces = np.asarray(forces.detach().cpu().numpy() if isinstance(forces, torch.Tensor) else forces,
                            dtype=np.float64)
        if forces.ndim == 1:
            forces = forces[None, :]
        if (forces.shape[0] != node_ids.shape[0] and forces.shape[0] != 1) or forces.shape[1] != self.dof:
            raise ValueError("forces must have shape (N_forces, mesh.dof).")
        if forces.shape[0] == 1:
            forces = np.tile(forces, (node_ids.shape[0], 1))
        rhs = self.rhs.cpu().numpy()
        for i in range(self.dof):
            # buffered fancy-index add, as the reference does: a node listed
            # twice receives its force once
            rhs[node_ids * self.dof + i] += forces[:, i]
        d_rhs = self.d_rhs.cpu().numpy()
        dir_dofs = ~np.isnan(d_rhs)
        rhs[dir_dofs] = d_rhs[dir_dofs]
        self.rhs = torch.from_numpy(rhs).to(self.kernel.device)

    def reset_forces(self):
        self.rhs.zero_()

    def reset_dirichlet_boundary_conditions(self):
        self.kernel.set_constraints([])
        self.kernel.has_con = False
        self.d_rhs.fill_(float("nan"))

    def solve(self, rho=None):
        if rho is not None and rho.shape[0] != self.nel:
            raise ValueError("rho must have the same length as the number of elements in the mesh.")
        if rho is None:
            rho = torch.ones(self.nel, dtype=torch.float64, device=self.kernel.device)
        self.kernel.set_rho(rho)
        return self.solver.solve(self.rhs, use_last=True)
