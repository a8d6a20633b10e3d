"""FiniteElement: boundary conditions, loads and ``solve(rho)``.

Mirrors FiniteElement/CPU/FiniteElement.py:89-445 (and the CUDA twin): the
same constructor, ``add_dirichlet_boundary_condition`` (node ids or
positions, optional per-DOF flags, prescribed values), ``add_point_forces``,
``reset_forces``, ``reset_dirichlet_boundary_conditions`` and
``solve(rho) -> (U, residual)``.  Positions are mapped to nodes by nearest
grid node (the reference queries a KDTree over the node coordinates; on a
structured grid the nearest node is the rounded grid index).  The constraint
flags, prescribed values and load vector are assembled with device tensor
operations: no host round trip.
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
        self.is_3D = (len(mesh.nel) if hasattr(mesh, "nel") else np.asarray(mesh.nodes).shape[1]) == 3
        self._nodes = None
        self.dof = mesh.dof

    def _dev(self, a, dtype):
        """A host array, list or tensor as a device tensor of the given dtype."""
        if isinstance(a, torch.Tensor):
            return a.to(device=self.kernel.device, dtype=dtype)
        return torch.as_tensor(np.asarray(a), dtype=dtype, device=self.kernel.device)

    def nearest_nodes(self, positions) -> torch.Tensor:
        """Node ids nearest to the given coordinates (KDTree query semantics,
        FiniteElement/CPU/FiniteElement.py:179-184): on a structured grid the
        nearest node is the rounded grid index, computed on the device; on an
        unstructured mesh it is the argmin of the distances to every node
        (lowest id on ties, as a KDTree query over the node array)."""
        if not hasattr(self.mesh, "nel"):
            if self._nodes is None:
                self._nodes = self._dev(self.mesh.nodes, torch.float64)
            X = self._nodes
            P = self._dev(positions, torch.float64).reshape(-1, X.shape[1])
            out = [torch.cdist(P[i:i + 1024], X).argmin(dim=1) for i in range(0, P.shape[0], 1024)]
            return torch.cat(out) if out else torch.zeros(0, dtype=torch.int64, device=X.device)
        d = len(self.mesh.nel)
        P = self._dev(positions, torch.float64).reshape(-1, d)
        h = torch.tensor([float(v) for v in self.mesh.spacing], dtype=torch.float64, device=P.device)
        nel = torch.tensor([int(v) for v in self.mesh.nel], dtype=torch.int64, device=P.device)
        idx = torch.minimum(torch.round(P / h).to(torch.int64).clamp_min(0), nel)
        nx = int(self.mesh.nel[0]) + 1
        if self.is_3D:
            nz = int(self.mesh.nel[2]) + 1
            return idx[:, 2] + nz * idx[:, 0] + nz * nx * idx[:, 1]
        return idx[:, 1] * nx + idx[:, 0]

    def _ids(self, node_ids) -> torch.Tensor:
        return self._dev(node_ids, torch.int64).reshape(-1)

    def add_dirichlet_boundary_condition(self, node_ids=None, positions=None, dofs=None, rhs=0):
        """FiniteElement/CPU/FiniteElement.py:108-188, on device tensors."""
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
            rhs_arr = self._dev(rhs, torch.float64)
            if rhs_arr.shape[0] != n_con:
                raise ValueError("rhs must have the same length as node_ids or positions.")
            rhs_arr = rhs_arr.reshape(n_con, -1)
        if dofs is not None:
            dofs = self._dev(dofs, torch.int64)
            if dofs.ndim == 1:
                dofs = dofs[None, :]
            if dofs.shape[0] != n_con and dofs.shape[0] != 1:
                raise ValueError("dofs must have the same length as node_ids.")
            if dofs.shape[0] == 1:
                dofs = dofs.expand(n_con, -1)
        for i in range(self.dof):
            sel = slice(None) if dofs is None else (dofs[:, i] == 1)
            cons = node_ids[sel] * self.dof + i
            self.kernel.add_constraints(cons)
            if rhs_arr is not None:
                self.d_rhs[cons] = rhs_arr[sel, i]
            else:
                self.d_rhs[cons] = float(rhs) + torch.nan_to_num(self.d_rhs[cons], nan=0.0)

    def add_neumann_boundary_condition(self, **kwargs):
        raise NotImplementedError("Neumann boundary conditions are not implemented in this version of FiniteElement.")

    def add_point_forces(self, forces, node_ids=None, positions=None):
        """FiniteElement/CPU/FiniteElement.py:190-259, on device tensors."""
        if node_ids is None and positions is None:
            raise ValueError("Either node_ids or positions must be provided.")
        if node_ids is not None and positions is not None:
            raise ValueError("Only one of node_ids or positions should be provided.")
        if positions is not None:
            return self.add_point_forces(forces=forces, node_ids=self.nearest_nodes(positions))
        node_ids = self._ids(node_ids)
        forces = self._dev(forces, torch.float64)
        if forces.ndim == 1:
            forces = forces[None, :]
        if (forces.shape[0] != node_ids.shape[0] and forces.shape[0] != 1) or forces.shape[1] != self.dof:
            raise ValueError("forces must have shape (N_forces, mesh.dof).")
        if forces.shape[0] == 1:
            forces = forces.expand(node_ids.shape[0], -1)
        # the reference's buffered fancy-index add (rhs[idx] += f): a node
        # listed twice receives one force, its last occurrence's
        uniq, inv = torch.unique(node_ids, return_inverse=True)
        pos = torch.arange(node_ids.shape[0], device=node_ids.device)
        last = torch.full((uniq.shape[0],), -1, dtype=torch.int64, device=node_ids.device)
        last.scatter_reduce_(0, inv, pos, reduce="amax")
        f = forces[last]
        for i in range(self.dof):
            idx = uniq * self.dof + i
            self.rhs[idx] = self.rhs[idx] + f[:, i]
        dir_dofs = ~torch.isnan(self.d_rhs)
        self.rhs[dir_dofs] = self.d_rhs[dir_dofs]

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
