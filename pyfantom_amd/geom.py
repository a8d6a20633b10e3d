"""Structured and unstructured meshes, structured density filters.

``StructuredMesh2D/3D`` mirror the reference's structured meshes
(geom/CPU/_mesh.py:10-210, geom/CUDA/_mesh.py:10-170): same constructor,
attributes and node/element numbering (core/CPU/_geom.py:4-98).  The kernels
never read the connectivity -- they index the grid implicitly -- so
``elements`` / ``nodes`` are generated on first access only, on the device
(fantom_mesh_elements / fantom_mesh_nodes_*, the counterpart of
generate_structured_mesh_cuda, core/CUDA/Kernels/_geom.py:54-118).

``StructuredFilter2D/3D`` mirror geom/CPU/_filters.py:14-194 and
geom/commons/_filters.py:19-157: cone weights over integer offsets,
normalised by their sum, forward ``dot`` and adjoint ``_rmatvec`` / ``T``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .physics import LinearElasticity


_HEX_BITS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                     dtype=np.float64)
_QUAD_BITS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64)


class _StructuredMesh:
    """Common parts of the 2D / 3D structured meshes."""

    def _finish(self, physics, dtype):
        self.dtype = dtype
        self.physics = physics
        # first element's corners (the reference evaluates physics on
        # nodes[elements[0]]): local corner pattern times the mesh spacing
        h = np.array(self.spacing, dtype=np.float64)
        bits = _HEX_BITS if len(self.nel) == 3 else _QUAD_BITS
        ref = bits * h[None, :]
        # K_single in the mesh dtype (the reference's attribute); the kernels
        # take their element matrix from the FP64 one, so a float32 mesh
        # differs from a float64 one only in its vector storage
        self.K_single64 = np.asarray(physics.K(ref), dtype=np.float64)
        self.K_single = self.K_single64.astype(dtype)
        self.locals = [np.asarray(a, dtype=dtype) for a in physics.locals(ref)]
        self.A_single = np.array([physics.volume(ref)], dtype=dtype)
        self.As = self.A_single
        self.volume = self.A_single[0] * int(np.prod(self.nel))
        self.elements_size = 2 ** len(self.nel)
        self.dof = int(self.K_single.shape[0] // self.elements_size)
        self._elements = None
        self._nodes = None

    @property
    def n_elements(self) -> int:
        return int(np.prod(self.nel))

    @property
    def n_nodes(self) -> int:
        return int(np.prod(self.nel + 1))

    @property
    def elements(self) -> torch.Tensor:
        """Connectivity int32 [n_el, 2^d] (core/CPU/_geom.py:4-59)."""
        if self._elements is None:
            el = torch.empty((self.n_elements, self.elements_size), dtype=torch.int32, device=_lib.device())
            _lib.call("fantom_mesh_elements", *self.grid, _lib.ptr(el))
            self._elements = el
        return self._elements

    @property
    def nodes(self) -> torch.Tensor:
        """Node coordinates [n_nodes, d] in meshgrid('xy') flatten order
        (core/CPU/_geom.py:74-91)."""
        if self._nodes is None:
            f32 = np.dtype(self.dtype) == np.float32
            X = torch.empty((self.n_nodes, len(self.nel)), dtype=torch.float32 if f32 else torch.float64,
                            device=_lib.device())
            L = [float(v) for v in self.dim] + [0.0] * (3 - len(self.nel))
            _lib.call("fantom_mesh_nodes_f32" if f32 else "fantom_mesh_nodes_f64", *self.grid, *L, _lib.ptr(X))
            self._nodes = X
        return self._nodes

    @property
    def grid(self):
        """(nelx, nely, nelz) as the kernels take it (nelz = 0 in 2D)."""
        return (int(self.nel[0]), int(self.nel[1]), int(self.nel[2]) if len(self.nel) == 3 else 0)


class StructuredMesh3D(_StructuredMesh):
    """Box of nx x ny x nz hex8 elements (geom/CUDA/_mesh.py:89-170)."""

    def __init__(self, nx, ny, nz, lx, ly, lz, dtype=np.float64, physics=None):
        physics = physics if physics is not None else LinearElasticity(E=1.0, nu=1 / 3)
        self.nelx, self.nely, self.nelz = int(nx), int(ny), int(nz)
        self.lx, self.ly, self.lz = lx, ly, lz
        self.nel = np.array([nx, ny, nz], dtype=np.int32)
        self.dim = np.array([lx, ly, lz], dtype=dtype)
        self.dx, self.dy, self.dz = lx / nx, ly / ny, lz / nz
        self.spacing = (self.dx, self.dy, self.dz)
        self._finish(physics, dtype)

    @property
    def centroids(self) -> torch.Tensor:
        c = np.meshgrid(np.linspace(self.dx / 2, self.lx - self.dx / 2, self.nelx),
                        np.linspace(self.dy / 2, self.ly - self.dy / 2, self.nely),
                        np.linspace(self.dz / 2, self.lz - self.dz / 2, self.nelz), indexing="ij")
        return torch.from_numpy(np.stack(c, -1).reshape(-1, 3).astype(self.dtype)).to(_lib.device())


class StructuredMesh2D(_StructuredMesh):
    """Rectangle of nx x ny bilinear quads (geom/CPU/_mesh.py:10-104)."""

    def __init__(self, nx, ny, lx, ly, dtype=np.float64, physics=None):
        physics = physics if physics is not None else LinearElasticity(E=1.0, nu=1 / 3)
        self.nelx, self.nely = int(nx), int(ny)
        self.lx, self.ly = lx, ly
        self.nel = np.array([nx, ny], dtype=np.int32)
        self.dim = np.array([lx, ly], dtype=dtype)
        self.dx, self.dy = lx / nx, ly / ny
        self.spacing = (self.dx, self.dy)
        self._finish(physics, dtype)

    @property
    def centroids(self) -> torch.Tensor:
        c = np.meshgrid(np.linspace(self.dx / 2, self.lx - self.dx / 2, self.nelx),
                        np.linspace(self.dy / 2, self.ly - self.dy / 2, self.nely), indexing="ij")
        return torch.from_numpy(np.stack(c, -1).reshape(-1, 2).astype(self.dtype)).to(_lib.device())


class GeneralMesh:
    """Unstructured mesh of triangles / quads (2D) or tetrahedra / hexes (3D)
    (geom/CPU/_mesh.py GeneralMesh, geom/CUDA/_mesh.py:195-240).

    ``elements`` is an (n_el x npe) array (uniform mesh) or a list of node
    lists of varying length (mixed element types).  Attributes as the
    reference: ``is_uniform``, ``elements_flat``, ``element_sizes``,
    ``elements_ptr``, ``Ks`` (n_el x m x m, uniform) or ``K_flat`` / ``K_ptr``
    (mixed), ``As``, ``volume``, ``dof``, ``centeroids`` (and ``centroids``),
    ``locals`` / ``locals_flat``.  Nodes that no element uses are dropped and
    the connectivity renumbered (the reference drops the nodes but computes
    the element matrices from the old numbering; here both use the new one).
    Host arrays; the stiffness kernels upload what they need.
    """

    def __init__(self, nodes, elements, dtype=np.float64, physics=None):
        physics = physics or LinearElasticity(E=1.0, nu=1.0 / 3.0)
        nodes = np.asarray(nodes, dtype=np.float64)
        if isinstance(elements, np.ndarray) and elements.ndim == 2:
            # one element type: no per-element Python work
            if elements.shape[0] == 0:
                raise ValueError("the mesh has no elements")
            sizes = np.full(elements.shape[0], elements.shape[1], dtype=np.int32)
            flat = np.asarray(elements, dtype=np.int64).ravel()
        else:
            elems = [np.asarray(e, dtype=np.int64).ravel() for e in elements]
            if not elems:
                raise ValueError("the mesh has no elements")
            sizes = np.array([e.size for e in elems], dtype=np.int32)
            flat = np.concatenate(elems)
        if flat.min() < 0 or flat.max() >= nodes.shape[0]:
            raise ValueError("element connectivity refers to missing nodes")
        used = np.unique(flat)
        remap = np.full(nodes.shape[0], -1, dtype=np.int64)
        remap[used] = np.arange(used.size)
        flat = remap[flat]
        self.dtype = dtype
        self.physics = physics
        self.nodes = nodes[used]
        self.element_sizes = sizes
        self.elements_ptr = np.concatenate(([0], np.cumsum(sizes))).astype(np.int32)
        self.elements_flat = flat.astype(np.int32)
        self.is_uniform = bool(np.all(sizes == sizes[0]))
        n_el = sizes.shape[0]
        if self.is_uniform:
            self.elements = flat.reshape(n_el, sizes[0]).astype(np.int32)
            X = self.nodes[self.elements]
            self.centeroids = X.mean(axis=1).astype(dtype)
            self.Ks = np.asarray(physics.K(X), dtype=dtype)
            self.locals = [np.asarray(a, dtype=dtype) for a in physics.locals(X)]
            self.As = np.asarray(physics.volume(X), dtype=dtype).reshape(-1)
            m = self.Ks.shape[1]
            self.K_flat = self.Ks.reshape(-1)
            self.K_ptr = (np.arange(n_el + 1, dtype=np.int64) * m * m)
        else:
            per = [flat[a:b] for a, b in zip(self.elements_ptr[:-1], self.elements_ptr[1:])]
            self.centeroids = np.stack([self.nodes[e].mean(axis=0) for e in per]).astype(dtype)
            self.elements = [e.astype(np.int32) for e in per]
            Ks = [np.asarray(physics.K(self.nodes[e]), dtype=dtype) for e in per]
            self.K_flat = np.concatenate([k.ravel() for k in Ks])
            self.K_ptr = np.concatenate(([0], np.cumsum([k.size for k in Ks]))).astype(np.int64)
            loc = [physics.locals(self.nodes[e]) for e in per]
            self.locals_flat = [np.concatenate([np.asarray(l[j], dtype=dtype).ravel() for l in loc])
                                for j in range(len(loc[0]))]
            self.locals_ptr = [np.concatenate(([0], np.cumsum([np.asarray(l[j]).size for l in loc])))
                               .astype(np.int32) for j in range(len(loc[0]))]
            self.As = np.array([physics.volume(self.nodes[e]) for e in per], dtype=dtype)
        self.centroids = self.centeroids
        m0 = int(np.sqrt(self.K_ptr[1] - self.K_ptr[0]))
        self.dof = m0 // int(sizes[0])
        self.volume = float(np.sum(self.As))
        self.n_elements = n_el
        self.n_nodes = self.nodes.shape[0]


def filter_taps(dim: int, r_min: float, spacing):
    """Offsets and normalised cone weights (geom/CPU/_filters.py:74-88)."""
    scales = np.asarray(spacing, dtype=np.float64)
    scales = scales / scales.min()
    n = int(np.ceil(r_min))
    rng = np.arange(-n, n + 1, dtype=np.int32)
    grids = np.meshgrid(*([rng] * dim), indexing="ij")
    offsets = np.vstack([g.ravel() for g in grids]).T.astype(np.int32)
    dist = np.linalg.norm(offsets * scales[None], axis=1)
    w = (r_min - dist) / r_min
    keep = w > 0
    offsets, w = offsets[keep], w[keep]
    w = w / w.sum()
    return np.ascontiguousarray(offsets), np.ascontiguousarray(w)


class _TransposeView:
    def __init__(self, f):
        self._f = f

    def dot(self, y):
        return self._f._rmatvec(y)

    def __matmul__(self, y):
        return self._f._rmatvec(y)


class _StructuredFilter:
    def _setup(self, mesh, r_min):
        self.mesh = mesh
        self.r_min = r_min
        self.dtype = mesh.dtype
        self.grid = mesh.grid
        n = mesh.n_elements
        self.shape = (n, n)
        offsets, w = filter_taps(len(mesh.nel), r_min, mesh.spacing)
        dev = _lib.device()
        self.offsets = offsets
        self.weights = torch.from_numpy(w).to(dev)
        self._offsets_d = torch.from_numpy(offsets.reshape(-1)).to(dev)
        self.normalizer = torch.empty(n, dtype=torch.float64, device=dev)
        _lib.call("fantom_filter_normalizer_f64", *self.grid, len(w), _lib.ptr(self._offsets_d),
                  _lib.ptr(self.weights), _lib.ptr(self.normalizer))

    def _matvec(self, rho):
        out = torch.empty_like(rho)
        _lib.call("fantom_filter_fwd_f64", *self.grid, self.weights.numel(), _lib.ptr(self._offsets_d),
                  _lib.ptr(self.weights), _lib.ptr(rho), _lib.ptr(out))
        return out

    def _rmatvec(self, y):
        y = y.contiguous()
        out = torch.empty_like(y)
        _lib.call("fantom_filter_adj_f64", *self.grid, self.weights.numel(), _lib.ptr(self._offsets_d),
                  _lib.ptr(self.weights), _lib.ptr(self.normalizer), _lib.ptr(y), _lib.ptr(out))
        return out

    def dot(self, rho):
        """Forward filter (geom/commons/_filters.py:106-143)."""
        if isinstance(rho, torch.Tensor):
            if rho.ndim == 1:
                if rho.shape[0] == self.shape[1]:
                    return self._matvec(rho.to(torch.float64).contiguous())
                raise ValueError("Input vector size does not match the filter kernel size.")
            raise NotImplementedError("Only vector inputs are supported.")
        raise ValueError(f"Input must be a {torch.Tensor} array vector.")

    def __matmul__(self, rhs):
        return self.dot(rhs)

    @property
    def T(self):
        return _TransposeView(self)


class StructuredFilter3D(_StructuredFilter):
    """Cone density filter on a 3D structured mesh (geom/CUDA/_filters.py:65-108)."""

    def __init__(self, mesh: StructuredMesh3D, r_min):
        self.nelx, self.nely, self.nelz = mesh.nelx, mesh.nely, mesh.nelz
        self._setup(mesh, r_min)


class StructuredFilter2D(_StructuredFilter):
    """Cone density filter on a 2D structured mesh (geom/CPU/_filters.py:106-194)."""

    def __init__(self, mesh: StructuredMesh2D, r_min):
        self.nelx, self.nely = mesh.nelx, mesh.nely
        self._setup(mesh, r_min)
