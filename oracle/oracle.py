"""FP64 CPU oracle for the structured SIMP finite-element solve path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only as the checker
(or as the timed CPU baseline).  The product package ``pyfantom_amd`` never
imports it.

Two layers:

* ``liboracle.so`` (``oracle/oracle.c``): a plain-C/OpenMP restatement of the
  reference's Numba kernels -- connectivity, node-basis K(E)u, diagonal,
  sensitivity, restriction/prolongation, density filter, OC update and an FP64
  GMG-PCG solve.  Each C function cites the reference file:line it follows.
* numpy/scipy/sympy helpers below that pin those functions independently:
  exact symbolic element stiffness matrices, brute-force sparse assembly,
  explicit transfer matrices and the Galerkin product.

Parity status of every function is listed in ``PINNED`` at the bottom.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from functools import lru_cache

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
SRC_PATH = os.path.join(_HERE, "oracle.c")


def build(force: bool = False) -> str:
    """Compile ``oracle.c`` into ``liboracle.so`` with gcc/OpenMP."""
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC_PATH):
        tmp = LIB_PATH + ".tmp"
        subprocess.check_call(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c99", "-o", tmp, SRC_PATH, "-lm"])
        os.replace(tmp, LIB_PATH)
    return LIB_PATH


_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_U8 = ctypes.POINTER(ctypes.c_uint8)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.or_elements_3d.argtypes = [c_int, c_int, c_int, _I]
        L.or_elements_2d.argtypes = [c_int, c_int, _I]
        L.or_matvec.argtypes = [_D, _I, c_long, c_int, c_int, c_int, _D, _D, _U8, _D]
        L.or_diag.argtypes = [_D, _I, c_long, c_int, c_int, c_int, _D, _U8, _D]
        L.or_process_dk.argtypes = [_D, _I, c_long, c_int, c_int, _D, _D]
        _L64 = ctypes.POINTER(ctypes.c_int64)
        L.or_matvec_flat.argtypes = [_D, _L64, _I, _I, c_long, c_int, c_int, _D, _D, _U8, _D]
        L.or_diag_flat.argtypes = [_D, _L64, _I, _I, c_long, c_int, c_int, _D, _U8, _D]
        L.or_process_dk_flat.argtypes = [_D, _L64, _I, _I, c_long, c_int, _D, _D]
        L.or_restrict_3d.argtypes = [c_int, c_int, c_int, c_int, c_int, _D, _D]
        L.or_restrict_2d.argtypes = [c_int, c_int, c_int, c_int, _D, _D]
        L.or_prolong_3d.argtypes = [c_int, c_int, c_int, c_int, c_int, _D, _D]
        L.or_prolong_2d.argtypes = [c_int, c_int, c_int, c_int, _D, _D]
        L.or_filter_normalizer.argtypes = [c_int, _I, c_int, _I, _D, _D]
        L.or_filter_fwd.argtypes = [c_int, _I, c_int, _I, _D, _D, _D]
        L.or_filter_adj.argtypes = [c_int, _I, c_int, _I, _D, _D, _D, _D]
        L.or_oc_update.argtypes = [c_long, _D, _D, _D, c_double, c_double, _D]
        L.or_oc_update.restype = c_double
        L.or_mgpcg_solve.argtypes = [c_int, _I, _D, _D, _U8, _D, _D, c_int, c_int, c_double, c_double, c_int, _D,
                                     c_int]
        L.or_mgpcg_solve.restype = c_int
        L.or_num_threads.restype = c_int
        L.or_set_num_threads.argtypes = [c_int]
        L.or_set_num_threads.restype = None
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _mask(con, n):
    if con is None:
        return None
    m = np.ascontiguousarray(np.asarray(con).astype(np.uint8))
    assert m.shape[0] == n
    return m


# --------------------------------------------------------------------------
# A1 connectivity and node coordinates (core/CPU/_geom.py:4-98)
# --------------------------------------------------------------------------
def elements(nel) -> np.ndarray:
    nel = [int(v) for v in nel]
    if len(nel) == 3:
        out = np.empty((nel[0] * nel[1] * nel[2], 8), dtype=np.int32)
        lib().or_elements_3d(nel[0], nel[1], nel[2], _p(out, _I))
    else:
        out = np.empty((nel[0] * nel[1], 4), dtype=np.int32)
        lib().or_elements_2d(nel[0], nel[1], _p(out, _I))
    return out


def nodes(nel, dim_len, dtype=np.float64) -> np.ndarray:
    """Node coordinates, meshgrid 'xy' flatten order (core/CPU/_geom.py:74-91):
    numpy's linspace of each length taken in the mesh's dtype."""
    axes = [np.linspace(0, dtype(L), n + 1, dtype=dtype) for L, n in zip(dim_len, nel)]
    grids = np.meshgrid(*axes)
    return np.stack([g.flatten() for g in grids], axis=-1)


def elements_bruteforce(nel) -> np.ndarray:
    """Connectivity by explicit enumeration of element corners from coordinates.

    Independent of the index formulas: it builds node coordinates in the
    reference's meshgrid order, hashes them, and looks up each element's
    corners in the local-node order of the reference (counter-clockwise bottom
    face then top face).
    """
    nel = [int(v) for v in nel]
    X = nodes(nel, nel)
    lut = {tuple(np.round(p).astype(int)): idx for idx, p in enumerate(X)}
    out = []
    if len(nel) == 3:
        corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        for i in range(nel[0]):
            for j in range(nel[1]):
                for k in range(nel[2]):
                    out.append([lut[(i + a, j + b, k + c)] for a, b, c in corners])
    else:
        corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
        for i in range(nel[0]):
            for j in range(nel[1]):
                out.append([lut[(i + a, j + b)] for a, b in corners])
    return np.asarray(out, dtype=np.int32)


# --------------------------------------------------------------------------
# A3 element stiffness: exact symbolic integration (independent of the
# reference's Gauss-quadrature code physics/LinearElasticity.py:700-834 and
# :305-443, which it must equal because 2-point Gauss is exact for the
# bilinear/trilinear integrands).
# --------------------------------------------------------------------------
@lru_cache(maxsize=None)
def k0_exact(dim: int, nu: float, hx: float = 1.0, hy: float = 1.0, hz: float = 1.0, E: float = 1.0,
             t: float = 1.0) -> np.ndarray:
    """Element stiffness by exact rational integration of B^T C B.

    Each bilinear/trilinear shape function is a product of 1D linear factors
    f0 = 1 - s/h, f1 = s/h, so every entry of the integrand is a product of 1D
    integrals whose closed forms are tabulated below with ``fractions``.  This
    is independent of the Gauss-quadrature code the product uses."""
    from fractions import Fraction as Fr

    nu_f = Fr(nu).limit_denominator(10 ** 9)
    if abs(float(nu_f) - nu) > 1e-15:
        nu_f = Fr(nu)

    def tab(h):
        h = Fr(h).limit_denominator(10 ** 9)
        # integrals over [0, h] of products of f_a and f_b or their derivatives
        ff = [[h / 3, h / 6], [h / 6, h / 3]]
        df = [[Fr(-1, 2), Fr(-1, 2)], [Fr(1, 2), Fr(1, 2)]]      # int f_a' f_b
        dd = [[1 / h, -1 / h], [-1 / h, 1 / h]]                  # int f_a' f_b'
        return ff, df, dd

    if dim == 3:
        corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        tabs = [tab(hx), tab(hy), tab(hz)]
        lam = nu_f / ((1 + nu_f) * (1 - 2 * nu_f))
        mu = 1 / (2 * (1 + nu_f))
        ncomp = 3
    else:
        corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
        tabs = [tab(hx), tab(hy)]
        # plane stress written as an isotropic law with the reduced lambda
        lam = nu_f / (1 - nu_f ** 2)
        mu = 1 / (2 * (1 + nu_f))
        ncomp = 2

    def G(a, b, p, q):
        """int dN_a/dx_p * dN_b/dx_q over the element."""
        val = Fr(1)
        for ax in range(dim):
            ff, df, dd = tabs[ax]
            ia, ib = corners[a][ax], corners[b][ax]
            if ax == p and ax == q:
                val *= dd[ia][ib]
            elif ax == p:
                val *= df[ia][ib]
            elif ax == q:
                val *= df[ib][ia]
            else:
                val *= ff[ia][ib]
        return val

    n = len(corners)
    K = np.zeros((n * ncomp, n * ncomp))
    for a in range(n):
        for b in range(n):
            for i in range(ncomp):
                for j in range(ncomp):
                    # isotropic: k_ij = lam G(i,j) + mu G(j,i) + mu delta_ij sum_k G(k,k)
                    v = lam * G(a, b, i, j) + mu * G(a, b, j, i)
                    if i == j:
                        v += mu * sum(G(a, b, k, k) for k in range(dim))
                    K[a * ncomp + i, b * ncomp + j] = float(v)
    return K * E * t


def k0_top88(nu: float) -> np.ndarray:
    """The 8x8 plane-stress Q4 matrix of Andreassen et al. (2011), 'Efficient
    topology optimization in MATLAB using 88 lines of code', in its own node
    order (upper-left, upper-right, lower-right, lower-left for a y-down grid)."""
    A11 = np.array([[12, 3, -6, -3], [3, 12, 3, 0], [-6, 3, 12, -3], [-3, 0, -3, 12]], float)
    A12 = np.array([[-6, -3, 0, 3], [-3, -6, -3, -6], [0, -3, -6, 3], [3, -6, 3, -6]], float)
    B11 = np.array([[-4, 3, -2, 9], [3, -4, -9, 4], [-2, -9, -4, -3], [9, 4, -3, -4]], float)
    B12 = np.array([[2, -3, 4, -9], [-3, 2, 9, -2], [4, 9, 2, 3], [-9, -2, 3, 2]], float)
    return 1 / (1 - nu ** 2) / 24 * (np.block([[A11, A12], [A12.T, A11]]) + nu * np.block([[B11, B12], [B12.T, B11]]))


# --------------------------------------------------------------------------
# Brute-force assembly (independent of the node-basis gather)
# --------------------------------------------------------------------------
def assemble(K0, elems, E, n_dof, con=None):
    """Sparse K(E) with the reference's Dirichlet convention: constrained
    rows and columns removed and unit diagonal put back
    (core/CPU/_ops.py:330-368 semantics)."""
    import scipy.sparse as sps

    npe = elems.shape[1]
    dof = K0.shape[0] // npe
    edofs = (elems[:, :, None] * dof + np.arange(dof)[None, None, :]).reshape(len(elems), -1)
    rows = np.repeat(edofs, edofs.shape[1], axis=1).ravel()
    cols = np.tile(edofs, (1, edofs.shape[1])).ravel()
    vals = (E[:, None, None] * K0[None]).reshape(len(elems), -1).ravel()
    K = sps.coo_matrix((vals, (rows, cols)), shape=(n_dof, n_dof)).tocsr()
    if con is not None:
        con = np.asarray(con, bool)
        free = sps.diags((~con).astype(float))
        K = (free @ K @ free + sps.diags(con.astype(float))).tocsr()
    return K


def csr_assemble(K0, elems, E, n_dof, con=None):
    """K(E) @ I in the CSR form of the reference's construct
    (stiffness/CUDA/_FEA.py:484-516; pattern from
    core/CUDA/Kernels/_matmat.py:258-300, values from
    core/CUDA/Kernels/matmat.cu:54-126): the pattern holds every pair of DOFs
    of one element with both DOFs free, explicit zeros of K0 included; a
    constrained DOF's row is the unit row.  Element scatter with duplicates
    summed; columns ascending within a row.  Returns (indptr int64, indices
    int32, data float64)."""
    npe = elems.shape[1]
    dof = K0.shape[0] // npe
    nl = npe * dof
    edofs = (elems[:, :, None].astype(np.int64) * dof + np.arange(dof)[None, None, :]).reshape(len(elems), -1)
    rows = np.repeat(edofs, nl, axis=1).ravel()
    cols = np.tile(edofs, (1, nl)).ravel()
    vals = (np.asarray(E, np.float64)[:, None, None] * K0[None]).reshape(len(elems), -1).ravel()
    if con is not None:
        con = np.asarray(con, bool)
        keep = ~con[rows] & ~con[cols]
        c = np.flatnonzero(con)
        rows = np.concatenate([rows[keep], c])
        cols = np.concatenate([cols[keep], c])
        vals = np.concatenate([vals[keep], np.ones(len(c))])
    key = rows * n_dof + cols
    order = np.argsort(key, kind="stable")
    key, vals = key[order], vals[order]
    uniq, start = np.unique(key, return_index=True)
    data = np.add.reduceat(vals, start) if len(vals) else vals
    indptr = np.zeros(n_dof + 1, np.int64)
    np.cumsum(np.bincount(uniq // n_dof, minlength=n_dof), out=indptr[1:])
    return indptr, (uniq % n_dof).astype(np.int32), data


# --------------------------------------------------------------------------
# C-library wrappers
# --------------------------------------------------------------------------
def matvec(K0, elems, E, u, con=None):
    K0 = _f64(K0)
    elems = np.ascontiguousarray(elems, np.int32)
    npe = elems.shape[1]
    dof = K0.shape[0] // npe
    u = _f64(u)
    n_nodes = u.shape[0] // dof
    m = _mask(con, u.shape[0])
    out = np.empty_like(u)
    lib().or_matvec(_p(K0, _D), _p(elems, _I), len(elems), npe, dof, n_nodes, _p(_f64(E), _D), _p(u, _D),
                    _p(m, _U8) if m is not None else None, _p(out, _D))
    return out


def diag(K0, elems, E, n_dof, con=None):
    K0 = _f64(K0)
    elems = np.ascontiguousarray(elems, np.int32)
    npe = elems.shape[1]
    dof = K0.shape[0] // npe
    m = _mask(con, n_dof)
    out = np.empty(n_dof)
    lib().or_diag(_p(K0, _D), _p(elems, _I), len(elems), npe, dof, n_dof // dof, _p(_f64(E), _D),
                  _p(m, _U8) if m is not None else None, _p(out, _D))
    return out


def process_dk(K0, elems, u):
    K0 = _f64(K0)
    elems = np.ascontiguousarray(elems, np.int32)
    npe = elems.shape[1]
    dof = K0.shape[0] // npe
    out = np.empty(len(elems))
    lib().or_process_dk(_p(K0, _D), _p(elems, _I), len(elems), npe, dof, _p(_f64(u), _D), _p(out, _D))
    return out


# --------------------------------------------------------------------------
# Unstructured ("flat") meshes: per-element matrices of varying size
# --------------------------------------------------------------------------
def _flat_args(K_flat, K_ptr, elements_flat, elements_ptr):
    return (_f64(np.asarray(K_flat).reshape(-1)), np.ascontiguousarray(K_ptr, np.int64),
            np.ascontiguousarray(elements_flat, np.int32), np.ascontiguousarray(elements_ptr, np.int32))


def matvec_flat(K_flat, K_ptr, elements_flat, elements_ptr, dof, E, u, con=None):
    """Node-basis K(E)u for mixed element sizes (core/CPU/_ops.py:145-176,
    :247-286)."""
    Kf, Kp, ef, ep = _flat_args(K_flat, K_ptr, elements_flat, elements_ptr)
    u = _f64(u)
    m = _mask(con, u.shape[0])
    out = np.empty_like(u)
    _L64 = ctypes.POINTER(ctypes.c_int64)
    lib().or_matvec_flat(_p(Kf, _D), _p(Kp, _L64), _p(ef, _I), _p(ep, _I), len(ep) - 1, dof, u.shape[0] // dof,
                         _p(_f64(E), _D), _p(u, _D), _p(m, _U8) if m is not None else None, _p(out, _D))
    return out


def diag_flat(K_flat, K_ptr, elements_flat, elements_ptr, dof, E, n_dof, con=None):
    """diag K(E), constrained DOFs 1 (core/CPU/_ops.py:110-143)."""
    Kf, Kp, ef, ep = _flat_args(K_flat, K_ptr, elements_flat, elements_ptr)
    m = _mask(con, n_dof)
    out = np.empty(n_dof)
    _L64 = ctypes.POINTER(ctypes.c_int64)
    lib().or_diag_flat(_p(Kf, _D), _p(Kp, _L64), _p(ef, _I), _p(ep, _I), len(ep) - 1, dof, n_dof // dof,
                       _p(_f64(E), _D), _p(m, _U8) if m is not None else None, _p(out, _D))
    return out


def process_dk_flat(K_flat, K_ptr, elements_flat, elements_ptr, dof, u):
    """-u_e^T K_e u_e per element (core/CPU/_ops.py:25-43)."""
    Kf, Kp, ef, ep = _flat_args(K_flat, K_ptr, elements_flat, elements_ptr)
    out = np.empty(len(ep) - 1)
    _L64 = ctypes.POINTER(ctypes.c_int64)
    lib().or_process_dk_flat(_p(Kf, _D), _p(Kp, _L64), _p(ef, _I), _p(ep, _I), len(ep) - 1, dof, _p(_f64(u), _D),
                             _p(out, _D))
    return out


def assemble_flat(K_flat, K_ptr, elements_flat, elements_ptr, dof, E, n_dof, con=None):
    """Brute-force scipy assembly of sum_e E_e P_e^T K_e P_e with the
    Dirichlet convention of ``assemble`` (independent of the node gather)."""
    import scipy.sparse as sps

    Kf = np.asarray(K_flat, np.float64).reshape(-1)
    rows, cols, vals = [], [], []
    for e in range(len(elements_ptr) - 1):
        nodes = np.asarray(elements_flat[elements_ptr[e]:elements_ptr[e + 1]], np.int64)
        g = (nodes[:, None] * dof + np.arange(dof)[None, :]).ravel()
        m = g.size
        Ke = Kf[K_ptr[e]:K_ptr[e] + m * m].reshape(m, m)
        rows.append(np.repeat(g, m))
        cols.append(np.tile(g, m))
        vals.append((E[e] * Ke).ravel())
    K = sps.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(n_dof, n_dof)).tocsr()
    if con is not None:
        con = np.asarray(con, bool)
        free = sps.diags((~con).astype(float))
        K = (free @ K @ free + sps.diags(con.astype(float))).tocsr()
    return K


def _coarse_nodes(nel):
    return int(np.prod([n // 2 + 1 for n in nel]))


def restrict(v, nel, dof, normalize=True):
    v = _f64(v)
    out = np.empty(_coarse_nodes(nel) * dof)
    if len(nel) == 3:
        lib().or_restrict_3d(int(nel[0]), int(nel[1]), int(nel[2]), dof, int(normalize), _p(v, _D), _p(out, _D))
    else:
        lib().or_restrict_2d(int(nel[0]), int(nel[1]), dof, int(normalize), _p(v, _D), _p(out, _D))
    return out


def prolong(v, nel, dof, normalize=True):
    v = _f64(v)
    out = np.empty(int(np.prod([n + 1 for n in nel])) * dof)
    if len(nel) == 3:
        lib().or_prolong_3d(int(nel[0]), int(nel[1]), int(nel[2]), dof, int(normalize), _p(v, _D), _p(out, _D))
    else:
        lib().or_prolong_2d(int(nel[0]), int(nel[1]), dof, int(normalize), _p(v, _D), _p(out, _D))
    return out


def transfer_matrix(nel, dof, normalize=True):
    """Explicit sparse restriction R (coarse x fine) from the weight formulas
    of core/CPU/_mgm.py:10-84 / :86-148, built without the C kernels."""
    import scipy.sparse as sps

    nel = [int(v) for v in nel]
    dim = len(nel)
    nf = [n + 1 for n in nel]
    nc = [n // 2 + 1 for n in nel]
    rows, cols, vals = [], [], []

    def fid(p):
        return p[2] + nf[2] * p[0] + nf[2] * nf[0] * p[1] if dim == 3 else p[1] * nf[0] + p[0]

    def cid(p):
        return p[2] + nc[2] * p[0] + nc[2] * nc[0] * p[1] if dim == 3 else p[1] * nc[0] + p[0]

    for cp in np.ndindex(*nc):
        b = sum(1 for a in range(dim) if cp[a] == 0 or cp[a] == nc[a] - 1)
        tw = ({3: 3.375, 2: 4.5, 1: 6.0, 0: 8.0} if dim == 3 else {2: 2.25, 1: 3.0, 0: 4.0})[b]
        if not normalize:
            tw = 1.0
        for off in np.ndindex(*([3] * dim)):
            d = [o - 1 for o in off]
            fp = [2 * cp[a] + d[a] for a in range(dim)]
            if any(fp[a] < 0 or fp[a] >= nf[a] for a in range(dim)):
                continue
            w = 0.5 ** sum(abs(x) for x in d)
            for c in range(dof):
                rows.append(cid(cp) * dof + c)
                cols.append(fid(fp) * dof + c)
                vals.append(w / tw)
    return sps.csr_matrix((vals, (rows, cols)), shape=(int(np.prod(nc)) * dof, int(np.prod(nf)) * dof))


def galerkin(A, nel, dof):
    """A_c = R A R^T with the reference's normalised R
    (core/CPU/_mgm.py:308-475 computes the same product structurally)."""
    R = transfer_matrix(nel, dof, True)
    return (R @ A @ R.T).tocsr()


# --------------------------------------------------------------------------
# A15 density filter (geom/CPU/_filters.py:61-93 / :152-183)
# --------------------------------------------------------------------------
def filter_taps(dim, r_min, scales=None):
    n = int(np.ceil(r_min))
    rng = np.arange(-n, n + 1, dtype=np.int32)
    grids = np.meshgrid(*([rng] * dim), indexing="ij")
    offsets = np.vstack([g.ravel() for g in grids]).T.astype(np.int32)
    sc = np.ones(dim) if scales is None else np.asarray(scales, float) / np.min(scales)
    dist = np.linalg.norm(offsets * sc[None], axis=1)
    w = (r_min - dist) / r_min
    keep = w > 0
    offsets, w = offsets[keep], w[keep]
    w = w / w.sum()
    return np.ascontiguousarray(offsets), np.ascontiguousarray(w)


def filter_normalizer(nel, offsets, weights):
    nel_a = np.ascontiguousarray(nel, np.int32)
    out = np.empty(int(np.prod(nel)))
    lib().or_filter_normalizer(len(nel), _p(nel_a, _I), len(weights), _p(np.ascontiguousarray(offsets, np.int32), _I),
                               _p(_f64(weights), _D), _p(out, _D))
    return out


def filter_fwd(nel, offsets, weights, x):
    nel_a = np.ascontiguousarray(nel, np.int32)
    out = np.empty(int(np.prod(nel)))
    lib().or_filter_fwd(len(nel), _p(nel_a, _I), len(weights), _p(np.ascontiguousarray(offsets, np.int32), _I),
                        _p(_f64(weights), _D), _p(_f64(x), _D), _p(out, _D))
    return out


def filter_adj(nel, offsets, weights, y):
    nel_a = np.ascontiguousarray(nel, np.int32)
    norm = filter_normalizer(nel, offsets, weights)
    out = np.empty(int(np.prod(nel)))
    lib().or_filter_adj(len(nel), _p(nel_a, _I), len(weights), _p(np.ascontiguousarray(offsets, np.int32), _I),
                        _p(_f64(weights), _D), _p(norm, _D), _p(_f64(y), _D), _p(out, _D))
    return out


def filter_dense(nel, offsets, weights):
    """Dense H by brute force over element pairs (tiny meshes only)."""
    dim = len(nel)
    n = int(np.prod(nel))
    H = np.zeros((n, n))
    pos = list(np.ndindex(*nel))
    index = {p: i for i, p in enumerate(pos)}
    for i, p in enumerate(pos):
        for o, w in zip(offsets, weights):
            q = tuple(p[a] + int(o[a]) for a in range(dim))
            if q in index:
                H[i, index[q]] += w
        H[i] /= H[i].sum()
    return H


# --------------------------------------------------------------------------
# A16 OC (Optimizers/CPU/OC.py:82-140)
# --------------------------------------------------------------------------
def oc_update(x, df, dg, vf, move=0.2):
    x = _f64(x)
    out = np.empty_like(x)
    lm = lib().or_oc_update(len(x), _p(x, _D), _p(_f64(df), _D), _p(_f64(dg), _D), float(vf), float(move), _p(out, _D))
    return out, lm


# --------------------------------------------------------------------------
# A4 SIMP laws (Problem/CPU/MinimumCompliance.py:358-464)
# --------------------------------------------------------------------------
def penalize_reference(rho, p=3.0, void=1e-6, E_mul=1.0, heavyside=True, beta=2.0, eta=0.5):
    if heavyside:
        h = (np.tanh(beta * eta) + np.tanh(beta * (rho - eta))) / (np.tanh(beta * eta) + np.tanh(beta * (1 - eta)))
    else:
        h = rho
    e = np.clip(h ** p, void, 1.0) * E_mul
    return np.clip(e, void, None)


def penalize_grad_reference(rho, p=3.0, E_mul=1.0, heavyside=True, beta=2.0, eta=0.5):
    if heavyside:
        den = np.tanh(beta * eta) + np.tanh(beta * (1 - eta))
        h = (np.tanh(beta * eta) + np.tanh(beta * (rho - eta))) / den
        return p * h ** (p - 1) * beta * (1 - np.tanh(beta * (rho - eta)) ** 2) / den * E_mul
    return p * rho ** (p - 1) * E_mul


def penalize_modified(rho, p=3.0, E0=1.0, Emin=1e-9):
    return Emin + rho ** p * (E0 - Emin)


def penalize_grad_modified(rho, p=3.0, E0=1.0, Emin=1e-9):
    return p * rho ** (p - 1) * (E0 - Emin)


# --------------------------------------------------------------------------
# FP64 GMG-PCG (the timed CPU baseline)
# --------------------------------------------------------------------------
def mgpcg_solve(nel, K0, E, con, f, n_level=4, n_smooth=2, omega=0.6, tol=1e-8, maxiter=1000, x0=None):
    nel_a = np.ascontiguousarray(nel, np.int32)
    f = _f64(f)
    u = np.zeros_like(f) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    m = _mask(con, f.shape[0])
    res = ctypes.c_double(0.0)
    it = lib().or_mgpcg_solve(len(nel), _p(nel_a, _I), _p(_f64(K0), _D), _p(_f64(E), _D),
                              _p(m, _U8) if m is not None else None, _p(f, _D), _p(u, _D), int(n_level),
                              int(n_smooth), float(omega), float(tol), int(maxiter), ctypes.byref(res),
                              int(x0 is not None))
    return u, it, res.value


def num_threads():
    return lib().or_num_threads()


def set_num_threads(n):
    lib().or_set_num_threads(int(n))


# Which check pins each oracle function (see tests/test_oracle.py).
# --------------------------------------------------------------------------
# Element post-processing (core/CPU/_ops.py:1030-1058,
# Problem/CPU/MinimumCompliance.py:659-736)
# --------------------------------------------------------------------------
def b_center(dim: int, hx: float = 1.0, hy: float = 1.0, hz: float = 1.0) -> np.ndarray:
    """The reference's Gauss-point-averaged B operator of a box element
    (physics/LinearElasticity.py:780-829 for hex8, :405-424 for quad4).  The
    derivatives of multilinear shape functions are multilinear in the other
    coordinates, so their average over the symmetric Gauss points is their
    value at the centre: dN_a/dx_p = (2 b_p - 1) / (2^(dim-1) h_p) for corner
    bits b.  Rows (Voigt, engineering shear): 3D xx, yy, zz, xy, yz, xz;
    2D xx, yy, xy."""
    if dim == 3:
        corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
        h = (hx, hy, hz)
        B = np.zeros((6, 24))
        for a, bits in enumerate(corners):
            dn = [(2 * bits[p] - 1) / (4.0 * h[p]) for p in range(3)]
            for i in range(3):
                B[i, 3 * a + i] = dn[i]
            B[3, 3 * a + 0], B[3, 3 * a + 1] = dn[1], dn[0]     # du/dy + dv/dx
            B[4, 3 * a + 1], B[4, 3 * a + 2] = dn[2], dn[1]     # dv/dz + dw/dy
            B[5, 3 * a + 0], B[5, 3 * a + 2] = dn[2], dn[0]     # du/dz + dw/dx
        return B
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
    h = (hx, hy)
    B = np.zeros((3, 8))
    for a, bits in enumerate(corners):
        dn = [(2 * bits[p] - 1) / (2.0 * h[p]) for p in range(2)]
        B[0, 2 * a] = dn[0]
        B[1, 2 * a + 1] = dn[1]
        B[2, 2 * a], B[2, 2 * a + 1] = dn[1], dn[0]
    return B


def d_iso(dim: int, nu: float, E: float = 1.0) -> np.ndarray:
    """Constitutive matrix: 3D isotropic (physics/LinearElasticity.py:744-765),
    2D plane stress (:358-364), Voigt order with engineering shear."""
    if dim == 3:
        f = E / ((1 + nu) * (1 - 2 * nu))
        D = np.zeros((6, 6))
        D[:3, :3] = f * nu
        D[[0, 1, 2], [0, 1, 2]] = f * (1 - nu)
        D[[3, 4, 5], [3, 4, 5]] = f * (1 - 2 * nu) / 2
        return D
    c = E / (1 - nu * nu)
    return np.array([[c, c * nu, 0.0], [c * nu, c, 0.0], [0.0, 0.0, c * (1 - nu) / 2]])


def fea_locals(K0, D, B, elems, w, u):
    """strain_e = B u_e, stress_e = D strain_e, energy_e = 1/2 u_e^T w_e K0 u_e
    (core/CPU/_ops.py:1030-1058) and the von Mises stress of
    MinimumCompliance.FEA (Problem/CPU/MinimumCompliance.py:719-722)."""
    npe = elems.shape[1]
    dof = K0.shape[0] // npe
    edofs = (elems[:, :, None] * dof + np.arange(dof)[None, None, :]).reshape(len(elems), -1)
    ue = u[edofs]
    strain = ue @ B.T
    stress = strain @ D.T
    energy = 0.5 * w * np.einsum("ei,ij,ej->e", ue, K0, ue)
    s = stress
    if dof == 3:
        vm = np.sqrt(0.5 * ((s[:, 0] - s[:, 1]) ** 2 + (s[:, 1] - s[:, 2]) ** 2 + (s[:, 2] - s[:, 0]) ** 2
                            + 6 * (s[:, 3] ** 2 + s[:, 4] ** 2 + s[:, 5] ** 2)))
    else:
        vm = np.sqrt(s[:, 0] ** 2 + s[:, 1] ** 2 - s[:, 0] * s[:, 1] + 3 * s[:, 2] ** 2)
    return strain, stress, energy, vm


PINNED = {
    "elements": "brute-force coordinate lookup (elements_bruteforce) + SHA-256 fixture",
    "k0 (product K0)": "exact sympy integration; top88 closed form up to node permutation; rigid-body nullspace",
    "matvec": "scipy assembly (assemble) on tiny meshes; symmetry <Ku,v>=<u,Kv>",
    "diag": "diagonal of scipy assembly",
    "process_dk": "u_e^T K0 u_e via dense einsum; finite difference of u^T K(E) u in E",
    "restrict/prolong": "explicit transfer_matrix; P == R^T; constants restrict to constants",
    "galerkin": "R A R^T product vs explicit dense product",
    "filter": "dense brute-force H; adjoint identity; H 1 = 1",
    "oc_update": "hand-checked 4-element case; volume feasibility and bisection invariant",
    "mgpcg_solve": "scipy spsolve of the same system",
    "b_center / d_iso": "patch test (linear field -> exact constant strain); constant-strain energy u^T K0 u = "
                        "V eps^T D eps against the exactly integrated K0",
    "fea_locals": "energy sum equals 1/2 u^T K u of the scipy assembly; von Mises of a uniaxial stress is |sigma|",
    "matvec_flat / diag_flat": "scipy assembly (assemble_flat) on tri/quad/tet/hex and mixed meshes; equals the "
                               "shared-K0 matvec/diag on a structured grid; symmetry",
    "process_dk_flat": "equals the shared-K0 process_dk on a structured grid; dense u_e^T K_e u_e",
    "csr_assemble": "closed-form nonzero counts (free grid 9 prod(3n-2), clamped face); dense matrix equals the "
                    "scipy assembly; K u equals the C matvec with constraints; symmetric pattern",
}


# --------------------------------------------------------------------------
# Chebyshev smoothing (solvers/CPU/_solvers.py:884-954, the CPU MultiGrid's
# _estimate_rho and _chebyshev_smoother, whose calls the reference leaves
# commented out at :1045/:1066/:1088)
# --------------------------------------------------------------------------
def chebyshev4_smoother(x, b, A, dinv, n_steps, rho):
    """Fourth-kind Chebyshev smoother, the reference's three-term form
    (:913-954): z_1 = 4/3 rho^-1 D^-1 r_0, z_k = (2k-3)/(2k+1) z_{k-1} +
    (8k-4)/(2k+1) rho^-1 D^-1 r_{k-1}, x_k = x_{k-1} + z_k.  A: callable u -> A u."""
    x = np.array(x, dtype=np.float64, copy=True)
    z = np.zeros_like(x)
    for k in range(1, n_steps + 1):
        r = b - A(x)
        f2 = (8 * k - 4) / (2 * k + 1)
        if k == 1:
            z = f2 / rho * (dinv * r)
        else:
            z = (2 * k - 3) / (2 * k + 1) * z + f2 / rho * (dinv * r)
        x = x + z
    return x


def estimate_rho(A, dinv, x0, num_iterations=10):
    """Power method on D^-1 A (:884-911): x <- D^-1 A x / |x|, then the
    quotient x . D^-1 A x.  The reference starts from np.random.rand; here the
    start vector is an argument (its body returns 0 before running, :885)."""
    x = np.asarray(x0, dtype=np.float64) / np.linalg.norm(x0)
    for _ in range(num_iterations):
        x = dinv * A(x)
        x /= np.linalg.norm(x)
    return float(x @ (dinv * A(x)))
