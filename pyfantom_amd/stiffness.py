"""Matrix-free structured stiffness operator K(E) on the device.

``StructuredStiffnessKernel`` mirrors the reference's CUDA class
(stiffness/CUDA/_FEA.py:6-262 base, :265-673 structured): ``set_rho``,
``dot`` / ``@`` / ``matvec``, ``diagonal``, ``add_constraints``,
``set_constraints``, ``process_grad`` and ``reset``, with the same checks and
exception types.  ``dot`` returns a reused output buffer, as the reference's
does (``self.mat_vec``), so callers must not hold the result across calls.

Differences from the reference, by design:
  * no connectivity/incidence arrays: the kernels index the grid implicitly;
  * constraints are kept both as the reference's bool[N_dof] and as a
    one-byte-per-node bit mask, which is what the kernels read;
  * in 3D the product uses the element-tiled Walsh-Hadamard block kernel
    (``kernel='tile'``); ``kernel='node'`` selects the node-centric gather.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def hex8_blocks(K0: np.ndarray) -> np.ndarray:
    """Walsh-Hadamard block form (72 doubles) of a box hex8 element matrix."""
    K0 = np.ascontiguousarray(K0, dtype=np.float64)
    out = np.zeros(72, dtype=np.float64)
    off = ctypes.c_double(0.0)
    _lib.check(_lib.load().fantom_hex8_blocks(K0.ctypes.data, out.ctypes.data, ctypes.addressof(off)))
    return out


class StructuredStiffnessKernel:
    def __init__(self, mesh, kernel: str | None = None):
        self.mesh = mesh
        self.dof = mesh.dof
        self.dtype = mesh.dtype
        self.K_single = mesh.K_single
        self.elements_size = mesh.elements_size
        self.n_nodes = mesh.n_nodes
        self.n_elements = mesh.n_elements
        self.grid = mesh.grid
        self.dim = len(mesh.nel)
        n_dof = self.n_nodes * self.dof
        self.shape = (n_dof, n_dof)
        dev = _lib.device()
        self.device = dev
        K64 = getattr(mesh, "K_single64", None)
        K64 = np.asarray(self.K_single, dtype=np.float64) if K64 is None else K64
        self.K0 = torch.from_numpy(np.ascontiguousarray(K64)).to(dev)
        if kernel is None:
            kernel = "tile" if self.dim == 3 else "node"
        if kernel not in ("tile", "node"):
            raise ValueError("kernel must be 'tile' or 'node'")
        if kernel == "tile" and self.dim != 3:
            raise ValueError("the tiled kernel is for hex8 meshes only")
        self.kernel_kind = kernel
        self.blocks = hex8_blocks(K64) if self.dim == 3 else None
        self.constraints = torch.zeros(n_dof, dtype=torch.bool, device=dev)
        self.mask = torch.zeros(self.n_nodes, dtype=torch.uint8, device=dev)
        self.has_con = False
        self.idx_map = torch.arange(n_dof, dtype=torch.int32, device=dev)
        self.non_con_map = self.idx_map
        self.has_rho = False
        self.rho = None
        self.rho64 = None   # FP64 moduli read by the solvers (the same tensor for FP64 meshes)
        self.f32 = np.dtype(self.dtype) == np.float32
        if self.f32 and self.dim != 3:
            raise NotImplementedError("dtype=float32 is implemented for hex8 meshes (the FP32 tiled operator)")
        self.vdtype = torch.float32 if self.f32 else torch.float64
        self.rho_version = 0
        self.con_version = 0
        self.mat_vec = None
        self.diag = None
        self.dk = None
        self._csr_pattern = None
        self.CSR = None
        self.ptr = None
        self.has_been_constructed = False
        self.matvec = self.dot
        self.rmatvec = self.dot
        self.matmat = self.dot
        self.rmatmat = self.dot

    # the CPU reference spells the flag has_cons (stiffness/CPU/_FEA.py:367)
    @property
    def has_cons(self):
        return self.has_con

    @has_cons.setter
    def has_cons(self, v):
        self.has_con = bool(v)

    # ---- operator handles used by the solvers -------------------------------------
    def op_args(self, scale: float = 1.0):
        """(kernel id, K0 argument) for the C ABI, with the element matrix scaled."""
        if self.kernel_kind == "tile":
            blk = self.blocks * scale if scale != 1.0 else self.blocks
            return _lib.KU_TILE, np.ascontiguousarray(blk)
        K0 = self.K0 * scale if scale != 1.0 else self.K0
        return _lib.KU_NODE, K0

    @staticmethod
    def k0_ptr(k0):
        if isinstance(k0, np.ndarray):
            return k0.ctypes.data
        return k0.data_ptr()

    def n_partials(self):
        kid, _ = self.op_args()
        return _lib.load().fantom_ku_partials(kid, *self.grid)

    def ku_residual(self, x, rhs, out):
        """out = rhs - K(E) x (FP64 moduli)."""
        kid, k0 = self.op_args()
        _lib.call("fantom_ku_residual_f64", kid, *self.grid, self.k0_ptr(k0), _lib.ptr(self.rho64),
                  _lib.ptr(self.mask) if self.has_con else None, _lib.ptr(rhs), _lib.ptr(x), _lib.ptr(out))

    def ku_dot(self, p, q, partials):
        """q = K(E) p and the block partial sums of p.q; returns their count."""
        kid, k0 = self.op_args()
        _lib.call("fantom_ku_dot_f64", kid, *self.grid, self.k0_ptr(k0), _lib.ptr(self.rho64),
                  _lib.ptr(self.mask) if self.has_con else None, _lib.ptr(p), _lib.ptr(q), _lib.ptr(partials))
        return _lib.load().fantom_ku_partials(kid, *self.grid)

    def _update_mask(self):
        c = self.constraints.view(self.n_nodes, self.dof).to(torch.uint8)
        bits = torch.zeros(self.n_nodes, dtype=torch.uint8, device=self.device)
        for d in range(self.dof):
            bits |= c[:, d] << d
        self.mask = bits
        self.con_version += 1
        self.non_con_map = self.idx_map[~self.constraints]

    # ---- reference API ---------------------------------------------------------------
    def set_rho(self, rho):
        """Set per-element moduli E_e (stiffness/CUDA/_FEA.py:178-194)."""
        if not isinstance(rho, torch.Tensor):
            rho = torch.as_tensor(np.asarray(rho))
        rho = rho.to(device=self.device, dtype=torch.float64).contiguous()
        if rho.shape[0] != self.n_elements:
            raise ValueError("rho must have one entry per element.")
        self.rho64 = rho
        self.rho = rho.to(torch.float32) if self.f32 else rho
        self.has_rho = True
        self.rho_version += 1

    def _as_vec(self, v, dtype=None):
        if not isinstance(v, torch.Tensor):
            raise NotImplementedError("Only device tensors as vectors are supported.")
        return v.to(device=self.device, dtype=dtype or self.vdtype).contiguous()

    def _matvec(self, rho, vec):
        if self.mat_vec is None or self.mat_vec.shape != vec.shape or self.mat_vec.dtype != vec.dtype:
            self.mat_vec = torch.empty_like(vec)
        if vec.dtype == torch.float32:
            # dtype float32: the FP32 tiled operator, as the reference dispatches f4 kernels
            _lib.call("fantom_ku_f32", *self.grid, self.blocks.ctypes.data, _lib.ptr(self.rho),
                      _lib.ptr(self.mask) if self.has_con else None, _lib.ptr(vec), _lib.ptr(self.mat_vec))
            return self.mat_vec
        kid, k0 = self.op_args()
        _lib.call("fantom_ku_f64", kid, *self.grid, self.k0_ptr(k0), _lib.ptr(rho),
                  _lib.ptr(self.mask) if self.has_con else None, _lib.ptr(vec), _lib.ptr(self.mat_vec))
        return self.mat_vec

    def _rmatvec(self, rho, vec):
        return self._matvec(rho, vec)

    def dot(self, rhs):
        """K(rho) @ rhs (stiffness/CUDA/_FEA.py:623-673); returns a reused buffer."""
        if not self.has_rho:
            raise ValueError("Rho has not been set. dot works only after setting rho.")
        if isinstance(rhs, torch.Tensor) and rhs.layout in (torch.sparse_csr, torch.sparse_coo):
            # K @ B = construct(rho) @ B, constrained rows taking B's rows
            # (core/CUDA/Kernels/matmat.cu:54-126 for a general B)
            if rhs.shape[0] != self.n_nodes * self.dof:
                raise ValueError("Shape of the input matrix does not match the number of nodes and dof.")
            A = self.construct(self.rho)
            B = rhs.to(device=self.device, dtype=torch.float64).to_sparse_csr()
            it = A.crow_indices().dtype
            if B.crow_indices().dtype != it:
                B = torch.sparse_csr_tensor(B.crow_indices().to(it), B.col_indices().to(it), B.values(), size=B.shape)
            return torch.sparse.mm(A, B)
        if isinstance(rhs, torch.Tensor):
            if rhs.ndim == 1 and rhs.shape[0] == self.n_nodes * self.dof:
                return self._matvec(self.rho, self._as_vec(rhs))
            raise ValueError("Shape of the input vector does not match the number of nodes and dof.")
        raise NotImplementedError("Only device tensors as vectors and sparse matrices are supported.")

    def __matmul__(self, rhs):
        return self.dot(rhs)

    def diagonal(self, rho=None):
        """diag(K(rho)), constrained DOFs -> 1 (stiffness/CUDA/_FEA.py:368-420)."""
        if rho is None and not self.has_rho:
            raise ValueError("Rho has not been set. diagonal works only after setting rho or if rho is provided.")
        rho = self.rho64 if rho is None else rho.to(device=self.device, dtype=torch.float64).contiguous()
        if self.diag is None:
            self.diag = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
        _lib.call("fantom_diag_f64", *self.grid, _lib.ptr(self.K0), _lib.ptr(rho),
                  _lib.ptr(self.mask) if self.has_con else None, _lib.ptr(self.diag), 0)
        # float32 meshes get the FP64 diagonal rounded to their dtype
        return self.diag.to(torch.float32) if self.f32 else self.diag

    def set_constraints(self, constraints):
        """Replace the Dirichlet set (stiffness/CUDA/_FEA.py:422-439)."""
        self.constraints[:] = False
        idx = torch.as_tensor(np.asarray(constraints, dtype=np.int64) if not isinstance(constraints, torch.Tensor)
                              else constraints, device=self.device).long()
        if idx.numel():
            self.constraints[idx] = True
        self.has_con = True
        self._update_mask()

    def add_constraints(self, constraints):
        """Add Dirichlet DOFs (stiffness/CUDA/_FEA.py:441-456)."""
        idx = torch.as_tensor(np.asarray(constraints, dtype=np.int64) if not isinstance(constraints, torch.Tensor)
                              else constraints, device=self.device).long()
        if idx.numel():
            self.constraints[idx] = True
        self.has_con = True
        self._update_mask()

    def process_grad(self, U):
        """dk_e = -u_e^T K0 u_e (stiffness/CUDA/_FEA.py:458-482)."""
        U = self._as_vec(U, torch.float64)
        if self.dk is None:
            self.dk = torch.empty(self.n_elements, dtype=torch.float64, device=self.device)
        _lib.call("fantom_dc_f64", *self.grid, _lib.ptr(self.K0), _lib.ptr(U), _lib.ptr(self.dk))
        # float32 meshes: the sensitivity of the FP64 contraction in their dtype
        return self.dk.to(torch.float32) if self.f32 else self.dk

    def construct(self, rho):
        """Explicit K(rho) as a torch sparse CSR tensor on the device
        (stiffness/CUDA/_FEA.py:484-516): the product K(rho) @ I of the
        reference, with the Dirichlet convention of ``dot`` (unit rows for
        constrained DOFs, their columns dropped elsewhere).  The pattern is
        kept while the constraints stay the same; columns are ascending within
        each row.  Index dtype int32 while the nonzeros fit, else int64."""
        if not isinstance(rho, torch.Tensor):
            rho = torch.as_tensor(np.asarray(rho))
        rho = rho.to(device=self.device, dtype=torch.float64).contiguous()
        if rho.shape[0] != self.n_elements:
            raise ValueError("rho must have one entry per element.")
        n_dof = self.shape[0]
        mask = _lib.ptr(self.mask) if self.has_con else None
        if self._csr_pattern is None or self._csr_pattern[0] != (self.con_version, self.has_con):
            counts = torch.empty(n_dof, dtype=torch.int32, device=self.device)
            _lib.call("fantom_csr_row_nnz", *self.grid, mask, _lib.ptr(counts))
            indptr = torch.zeros(n_dof + 1, dtype=torch.int64, device=self.device)
            torch.cumsum(counts, 0, out=indptr[1:])
            nnz = int(indptr[-1])
            indices = torch.empty(nnz, dtype=torch.int32, device=self.device)
            self._csr_pattern = ((self.con_version, self.has_con), indptr, indices, None)
        _, indptr, indices, crow = self._csr_pattern
        data = torch.empty(indices.shape[0], dtype=torch.float64, device=self.device)
        _lib.call("fantom_csr_assemble_f64", *self.grid, _lib.ptr(self.K0), _lib.ptr(rho), mask,
                  _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(data))
        if crow is None:
            # torch wants one index dtype for both arrays
            crow = (indptr.to(torch.int32), indices) if indices.shape[0] < 2 ** 31 else (indptr, indices.long())
            self._csr_pattern = self._csr_pattern[:3] + (crow,)
        self.CSR = torch.sparse_csr_tensor(crow[0], crow[1], data, size=self.shape)
        self.ptr = crow[0]
        self.has_been_constructed = True
        return self.CSR

    def reset(self):
        self.has_rho = False
        self.rho = None
        self.rho64 = None
        self.constraints[:] = False
        self.has_con = False
        self._update_mask()
        self.mat_vec = None
        self.diag = None
        self.dk = None


class _NodeBasisStiffnessKernel:
    """Shared body of ``UniformStiffnessKernel`` and ``GeneralStiffnessKernel``
    (stiffness/CUDA/_FEA.py:675-1105 and :1107-1630): unstructured meshes with
    per-element matrices, applied matrix-free.

    Device layout: the elements are grouped by node count; each group keeps
    its element ids, its connectivity (npe x n_el) and its matrices
    structure-of-arrays (m*m x n_el), so that the per-element product reads
    the matrices coalesced.  A product is one pass per group computing
    E_e K_e u_e into a workspace, then a per-DOF gather over the node-sorted
    incidence (the reference's ``sorter`` / ``node_ids`` pair, here a stable
    argsort of ``elements_flat``: each node visits its elements in ascending
    id order, so every row is summed in a fixed order).  FP64 only."""

    def __init__(self, mesh):
        if np.dtype(mesh.dtype) != np.float64:
            raise NotImplementedError("unstructured meshes are implemented in float64")
        self.mesh = mesh
        self.nodes, self.elements = mesh.nodes, mesh.elements
        self.dof = int(mesh.dof)
        self.dtype = mesh.dtype
        self.n_nodes = int(mesh.n_nodes)
        self.n_elements = int(mesh.n_elements)
        n_dof = self.n_nodes * self.dof
        self.shape = (n_dof, n_dof)
        flat = np.asarray(mesh.elements_flat, dtype=np.int64)
        eptr = np.asarray(mesh.elements_ptr, dtype=np.int64)
        sizes = np.diff(eptr)
        kp = np.asarray(mesh.K_ptr, dtype=np.int64)
        K_flat = np.asarray(mesh.K_flat, dtype=np.float64).reshape(-1)
        if kp[-1] != K_flat.size or np.any(np.diff(kp) != (sizes * self.dof) ** 2):
            raise ValueError("K_ptr does not match the element sizes and dof")
        npes = np.unique(sizes)
        if npes.size > 8 or npes.max() > 255 or npes.max() * self.dof > 96:
            raise ValueError("at most 8 element sizes of at most 96 DOF each are supported")
        dev = _lib.device()
        self.device = dev
        up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
        # element groups
        group = np.empty(self.n_elements, dtype=np.int64)
        local = np.empty(self.n_elements, dtype=np.int64)
        self.groups = []
        for g, npe in enumerate(npes.tolist()):
            ids = np.flatnonzero(sizes == npe)
            group[ids] = g
            local[ids] = np.arange(ids.size)
            m = npe * self.dof
            nodes = flat[eptr[ids][:, None] + np.arange(npe)[None, :]]            # n_g x npe
            if ids.size == self.n_elements:
                Kg = K_flat.reshape(ids.size, m * m)                                 # one size: a view
            else:
                Kg = K_flat[kp[ids][:, None] + np.arange(m * m)[None, :]]          # n_g x m*m
            Kd = up(Kg, np.float64).t().contiguous()                               # m*m x n_g
            self.groups.append({"npe": npe, "m": m, "n_el": int(ids.size), "el_ids": up(ids, np.int32),
                                "nodes": up(nodes.T, np.int32), "K": Kd})
        self._gtab = (_lib.ElemGroup * len(self.groups))(*[
            _lib.ElemGroup(g["n_el"], g["npe"], g["el_ids"].data_ptr(), g["nodes"].data_ptr(), g["K"].data_ptr())
            for g in self.groups])
        # node-sorted incidence: (element index in its group, group * 256 + local node)
        el_ids = np.repeat(np.arange(self.n_elements, dtype=np.int64), sizes)
        order = np.argsort(flat, kind="stable")
        inc_el = el_ids[order]
        node_ptr = np.searchsorted(flat[order], np.arange(self.n_nodes + 1), side="left")
        self.node_ptr = up(node_ptr, np.int32)
        self.inc_l = up(local[inc_el], np.int32)
        self.inc_ga = up(group[inc_el] * 256 + (order - eptr[inc_el]), np.int32)
        nY = _lib.load().fantom_ku_general_workspace(self.dof, len(self.groups), ctypes.addressof(self._gtab))
        if nY < 0:
            raise ValueError("invalid element groups")
        self._Y = torch.empty(max(int(nY), 1), dtype=torch.float64, device=dev)
        self.elements_size = int(sizes.max())
        self.max_con_count = int(sizes.max()) * self.dof * int(np.diff(node_ptr).max())
        self.constraints = torch.zeros(n_dof, dtype=torch.bool, device=dev)
        self._con_u8 = torch.zeros(n_dof, dtype=torch.uint8, device=dev)
        self.has_con = False
        self.idx_map = torch.arange(n_dof, dtype=torch.int32, device=dev)
        self.non_con_map = self.idx_map
        self.has_rho = False
        self.rho = None
        self.rho64 = None
        self.f32 = False
        self.vdtype = torch.float64
        self.mat_vec = None
        self.diag = None
        self.dk = None
        self.CSR = None
        self.ptr = None
        self._coo = None
        self.has_been_constructed = False
        self.matvec = self.dot
        self.rmatvec = self.dot
        self.matmat = self.dot
        self.rmatmat = self.dot

    @property
    def has_cons(self):
        return self.has_con

    @has_cons.setter
    def has_cons(self, v):
        self.has_con = bool(v)

    def _con(self):
        return _lib.ptr(self._con_u8) if self.has_con else None

    def _inc(self):
        return (len(self.groups), ctypes.addressof(self._gtab), _lib.ptr(self.node_ptr), _lib.ptr(self.inc_l),
                _lib.ptr(self.inc_ga))

    def _ku(self, mode, E, u, y, b=None, partials=None):
        _lib.call("fantom_ku_general_f64", mode, self.n_nodes, self.dof, *self._inc(), _lib.ptr(E), self._con(),
                  _lib.ptr(u), _lib.ptr(y), _lib.ptr(b), _lib.ptr(self._Y), _lib.ptr(partials))

    # ---- operator calls used by the solvers -----------------------------------------
    def n_partials(self):
        return _lib.load().fantom_ku_general_partials(self.n_nodes, self.dof)

    def ku_residual(self, x, rhs, out):
        """out = rhs - K(E) x."""
        self._ku(2, self.rho64, x, out, b=rhs)

    def ku_dot(self, p, q, partials):
        """q = K(E) p and the block partial sums of p.q; returns their count."""
        self._ku(3, self.rho64, p, q, partials=partials)
        return self.n_partials()

    # ---- reference API ---------------------------------------------------------------
    def set_rho(self, rho):
        """Set per-element moduli (stiffness/CUDA/_FEA.py:172-187)."""
        if not isinstance(rho, torch.Tensor):
            rho = torch.as_tensor(np.asarray(rho))
        rho = rho.to(device=self.device, dtype=torch.float64).contiguous()
        if rho.shape[0] != self.n_elements:
            raise ValueError("rho must have one entry per element.")
        self.rho = self.rho64 = rho
        self.has_rho = True

    def _as_vec(self, v, dtype=None):
        if not isinstance(v, torch.Tensor):
            raise NotImplementedError("Only device tensors as vectors are supported.")
        return v.to(device=self.device, dtype=dtype or torch.float64).contiguous()

    def _matvec(self, rho, vec):
        if self.mat_vec is None or self.mat_vec.shape != vec.shape:
            self.mat_vec = torch.empty_like(vec)
        rho = rho.to(device=self.device, dtype=torch.float64).contiguous()
        self._ku(0, rho, vec, self.mat_vec)
        return self.mat_vec

    def _rmatvec(self, rho, vec):
        return self._matvec(rho, vec)

    def dot(self, rhs):
        """K(rho) @ rhs (stiffness/CUDA/_FEA.py:1051-1105, :1574-1628); returns a
        reused buffer.  A sparse right-hand side goes through ``construct``."""
        if not self.has_rho:
            raise ValueError("Rho has not been set. dot works only after setting rho.")
        if isinstance(rhs, torch.Tensor) and rhs.layout in (torch.sparse_csr, torch.sparse_coo):
            if rhs.shape[0] != self.shape[0]:
                raise ValueError("Shape of the input matrix does not match the number of nodes and dof.")
            A = self.construct(self.rho)
            B = rhs.to(device=self.device, dtype=torch.float64).to_sparse_csr()
            it = A.crow_indices().dtype
            if B.crow_indices().dtype != it:
                B = torch.sparse_csr_tensor(B.crow_indices().to(it), B.col_indices().to(it), B.values(), size=B.shape)
            return torch.sparse.mm(A, B)
        if isinstance(rhs, torch.Tensor):
            if rhs.ndim == 1 and rhs.shape[0] == self.shape[0]:
                return self._matvec(self.rho, self._as_vec(rhs))
            raise ValueError("Shape of the input vector does not match the number of nodes and dof.")
        raise NotImplementedError("Only device tensors as vectors and sparse matrices are supported.")

    def __matmul__(self, rhs):
        return self.dot(rhs)

    def diagonal(self, rho=None):
        """diag(K(rho)), constrained DOFs -> 1 (stiffness/CUDA/_FEA.py:772-828)."""
        if rho is None and not self.has_rho:
            raise ValueError("Rho has not been set. diagonal works only after setting rho or if rho is provided.")
        rho = self.rho64 if rho is None else rho.to(device=self.device, dtype=torch.float64).contiguous()
        if self.diag is None:
            self.diag = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
        _lib.call("fantom_diag_general_f64", self.n_nodes, self.dof, *self._inc(), _lib.ptr(rho), self._con(),
                  _lib.ptr(self.diag))
        return self.diag

    def _idx(self, constraints):
        if isinstance(constraints, torch.Tensor):
            return constraints.to(self.device).long()
        return torch.as_tensor(np.asarray(constraints, dtype=np.int64), device=self.device)

    def set_constraints(self, constraints):
        """Replace the Dirichlet set (stiffness/CUDA/_FEA.py:830-847)."""
        self.constraints[:] = False
        self.add_constraints(constraints)

    def add_constraints(self, constraints):
        """Add Dirichlet DOFs (stiffness/CUDA/_FEA.py:849-864)."""
        idx = self._idx(constraints)
        if idx.numel():
            self.constraints[idx] = True
        self.has_con = True
        self._con_u8 = self.constraints.to(torch.uint8)
        self.non_con_map = self.idx_map[~self.constraints]

    def process_grad(self, U):
        """dk_e = -u_e^T K_e u_e (stiffness/CUDA/_FEA.py:866-889)."""
        U = self._as_vec(U, torch.float64)
        if self.dk is None:
            self.dk = torch.empty(self.n_elements, dtype=torch.float64, device=self.device)
        _lib.call("fantom_dc_general_f64", self.dof, len(self.groups), ctypes.addressof(self._gtab), _lib.ptr(U),
                  _lib.ptr(self.dk))
        return self.dk

    def construct(self, rho):
        """Explicit K(rho) as a torch sparse CSR tensor on the device
        (stiffness/CUDA/_FEA.py:891-923): the element matrices scattered by
        their global DOFs and summed, with the Dirichlet convention of ``dot``
        (unit rows for constrained DOFs, their columns dropped elsewhere)."""
        if not isinstance(rho, torch.Tensor):
            rho = torch.as_tensor(np.asarray(rho))
        rho = rho.to(device=self.device, dtype=torch.float64).contiguous()
        if rho.shape[0] != self.n_elements:
            raise ValueError("rho must have one entry per element.")
        if self._coo is None:
            # global (row, col) of every stored entry, in the groups' K order
            parts = []
            for g in self.groups:
                m, n_g, dof = g["m"], g["n_el"], self.dof
                i = torch.arange(m, device=self.device)
                gdof = g["nodes"].long()[i // dof] * dof + (i % dof)[:, None]           # m x n_g
                row = gdof[:, None, :].expand(m, m, n_g).reshape(-1)
                col = gdof[None, :, :].expand(m, m, n_g).reshape(-1)
                el = g["el_ids"].long()[None, :].expand(m * m, n_g).reshape(-1)
                parts.append((el, row, col))
            self._coo = tuple(torch.cat(t) for t in zip(*parts))
        el, row, col = self._coo
        val = torch.cat([g["K"].reshape(-1) for g in self.groups]) * rho[el]
        if self.has_con:
            keep = ~(self.constraints[row] | self.constraints[col])
            cdofs = torch.nonzero(self.constraints).reshape(-1)
            row = torch.cat([row[keep], cdofs])
            col = torch.cat([col[keep], cdofs])
            val = torch.cat([val[keep], torch.ones(cdofs.numel(), dtype=torch.float64, device=self.device)])
        A = torch.sparse_coo_tensor(torch.stack([row, col]), val, size=self.shape).coalesce()
        self.CSR = A.to_sparse_csr()
        self.ptr = self.CSR.crow_indices()
        self.has_been_constructed = True
        return self.CSR

    def reset(self):
        self.has_rho = False
        self.rho = None
        self.rho64 = None
        self.constraints[:] = False
        self._con_u8.zero_()
        self.has_con = False
        self.non_con_map = self.idx_map
        self.mat_vec = None
        self.diag = None
        self.dk = None


class UniformStiffnessKernel(_NodeBasisStiffnessKernel):
    """One element type (stiffness/CUDA/_FEA.py:675-1105): ``mesh.Ks`` holds
    the n_el x m x m element matrices."""

    def __init__(self, mesh):
        if not mesh.is_uniform:
            raise ValueError("Mesh is not uniform, you should use GeneralStiffnessKernel instead.")
        super().__init__(mesh)


class GeneralStiffnessKernel(_NodeBasisStiffnessKernel):
    """Mixed element types (stiffness/CUDA/_FEA.py:1107-1630): element matrices
    of varying size in ``mesh.K_flat`` at offsets ``mesh.K_ptr``."""

    def __init__(self, mesh):
        if mesh.is_uniform:
            raise ValueError("Mesh is uniform, you should use UniformStiffnessKernel instead.")
        super().__init__(mesh)
