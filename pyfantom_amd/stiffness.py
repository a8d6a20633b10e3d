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
