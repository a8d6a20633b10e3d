"""Element physics: isotropic linear elasticity on 3-node triangles, bilinear
quads, 4-node tetrahedra and trilinear hexes.

Mirrors the constructor and the ``K`` / ``locals`` / ``volume`` methods of the
reference's ``LinearElasticity`` (physics/LinearElasticity.py:74-118), which
take one element's node coordinates (npe x dim) or a batch (n x npe x dim).
Quads and hexes are integrated with 2-point Gauss quadrature per axis, which
is exact for the polynomial integrand of an affine (box) element; triangles
and tetrahedra have constant strain (one-point rule, exact).  Node order:
counter-clockwise triangles and quads, right-handed tetrahedra, hexes with a
counter-clockwise bottom face then the top face; a non-positive Jacobian is
an error, as in the reference.  This runs at mesh setup on the host.
"""
from __future__ import annotations

import numpy as np

_G = 1.0 / np.sqrt(3.0)

# natural coordinates of the corners in the reference's local order
_HEX_XI = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
_QUAD_XI = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=np.float64)


def _iso_3d(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] = lam + 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def _plane(E, nu, plane_stress):
    if plane_stress:
        c = E / (1 - nu * nu)
        return np.array([[c, c * nu, 0], [c * nu, c, 0], [0, 0, c * (1 - nu) / 2]])
    c = E / ((1 + nu) * (1 - 2 * nu))
    return np.array([[c * (1 - nu), c * nu, 0], [c * nu, c * (1 - nu), 0], [0, 0, c * (1 - 2 * nu) / 2]])


def _hex_B(X, xi):
    """Strain-displacement matrix (6x24) and det J at natural point xi."""
    # dN/dxi (3 x 8): d/dxi_a of prod_b (1 + xi_b * s_b)/8
    s = _HEX_XI
    f = 1 + s * xi[None, :]
    dN = np.empty((3, 8))
    for a in range(3):
        others = [b for b in range(3) if b != a]
        dN[a] = s[:, a] * f[:, others[0]] * f[:, others[1]] / 8.0
    J = dN @ X
    detJ = np.linalg.det(J)
    G = np.linalg.solve(J, dN)  # dN/dx (3 x 8)
    B = np.zeros((6, 24))
    for n in range(8):
        gx, gy, gz = G[:, n]
        B[0, 3 * n] = gx
        B[1, 3 * n + 1] = gy
        B[2, 3 * n + 2] = gz
        B[3, 3 * n], B[3, 3 * n + 1] = gy, gx          # gamma_xy
        B[4, 3 * n + 1], B[4, 3 * n + 2] = gz, gy      # gamma_yz
        B[5, 3 * n], B[5, 3 * n + 2] = gz, gx          # gamma_xz
    return B, detJ


def _quad_B(X, xi):
    s = _QUAD_XI
    f = 1 + s * xi[None, :]
    dN = np.empty((2, 4))
    dN[0] = s[:, 0] * f[:, 1] / 4.0
    dN[1] = s[:, 1] * f[:, 0] / 4.0
    J = dN @ X
    detJ = np.linalg.det(J)
    G = np.linalg.solve(J, dN)
    B = np.zeros((3, 8))
    for n in range(4):
        gx, gy = G[:, n]
        B[0, 2 * n] = gx
        B[1, 2 * n + 1] = gy
        B[2, 2 * n], B[2, 2 * n + 1] = gy, gx
    return B, detJ


def _tensor_B_batch(X, xi):
    """Batched strain-displacement matrices and det J of quads (n x 4 x 2) or
    hexes (n x 8 x 3) at one natural point xi (the same shape functions as
    ``_quad_B`` / ``_hex_B``)."""
    n, npe, dim = X.shape
    s = _HEX_XI if dim == 3 else _QUAD_XI
    f = 1 + s * xi[None, :]
    dN = np.empty((dim, npe))
    for a in range(dim):
        prod = np.ones(npe)
        for b in range(dim):
            if b != a:
                prod = prod * f[:, b]
        dN[a] = s[:, a] * prod / float(npe)
    J = np.matmul(dN[None], X)                       # n x dim x dim
    det = np.linalg.det(J)
    G = np.linalg.solve(J, np.broadcast_to(dN, (n, dim, npe)))
    if dim == 2:
        B = np.zeros((n, 3, 8))
        B[:, 0, 0::2] = G[:, 0]
        B[:, 1, 1::2] = G[:, 1]
        B[:, 2, 0::2], B[:, 2, 1::2] = G[:, 1], G[:, 0]
        return B, det
    B = np.zeros((n, 6, 24))
    B[:, 0, 0::3] = G[:, 0]
    B[:, 1, 1::3] = G[:, 1]
    B[:, 2, 2::3] = G[:, 2]
    B[:, 3, 0::3], B[:, 3, 1::3] = G[:, 1], G[:, 0]     # gamma_xy
    B[:, 4, 1::3], B[:, 4, 2::3] = G[:, 2], G[:, 1]     # gamma_yz
    B[:, 5, 0::3], B[:, 5, 2::3] = G[:, 2], G[:, 0]     # gamma_xz
    return B, det


def _gauss_points(dim):
    g = (-_G, _G)
    if dim == 3:
        return [np.array([a, b, c]) for a in g for b in g for c in g]
    return [np.array([a, b]) for a in g for b in g]


def _simplex_B(X):
    """Constant strain-displacement matrices and measures of a batch of
    triangles (n x 3 x 2) or tetrahedra (n x 4 x 3): N_0 = 1 - sum xi, N_i =
    xi_i, J = dN/dxi X, dN/dx = J^-1 dN/dxi, measure = det J / dim!."""
    n, npe, dim = X.shape
    dNr = np.zeros((dim, npe))
    dNr[:, 0] = -1.0
    dNr[np.arange(dim), np.arange(1, npe)] = 1.0
    J = np.einsum("an,enb->eab", dNr, X)
    det = np.linalg.det(J)
    if np.any(det <= 0):
        bad = np.flatnonzero(det <= 0)
        raise Exception(f"Node Order Is Not Correct for elements: {bad.tolist()[:10]}")
    G = np.linalg.solve(J, np.broadcast_to(dNr, (n, dim, npe)))   # dN/dx, n x dim x npe
    if dim == 2:
        B = np.zeros((n, 3, 2 * npe))
        B[:, 0, 0::2] = G[:, 0]
        B[:, 1, 1::2] = G[:, 1]
        B[:, 2, 0::2] = G[:, 1]
        B[:, 2, 1::2] = G[:, 0]
        return B, det / 2.0
    B = np.zeros((n, 6, 3 * npe))
    B[:, 0, 0::3] = G[:, 0]
    B[:, 1, 1::3] = G[:, 1]
    B[:, 2, 2::3] = G[:, 2]
    B[:, 3, 0::3], B[:, 3, 1::3] = G[:, 1], G[:, 0]     # gamma_xy
    B[:, 4, 1::3], B[:, 4, 2::3] = G[:, 2], G[:, 1]     # gamma_yz
    B[:, 5, 0::3], B[:, 5, 2::3] = G[:, 2], G[:, 0]     # gamma_xz
    return B, det / 6.0


class LinearElasticity:
    """Isotropic linear elasticity (physics/LinearElasticity.py:74-118).

    ``type`` selects plane stress or plane strain for 2D elements."""

    def __init__(self, E=1.0, nu=1.0 / 3.0, thickness=1.0, type="PlaneStress"):
        self.E = E
        self.nu = nu
        self.thickness = thickness
        self.type = 0 if type == "PlaneStress" else 1

    def _check(self, x0s):
        x0s = np.asarray(x0s, dtype=np.float64)
        if x0s.shape == (8, 3):
            return 3, x0s
        if x0s.shape == (4, 2):
            return 2, x0s
        raise ValueError("Invalid input shape")

    def _simplex(self, x0s):
        """(B batch, measure batch, D, batched?) for triangles / tetrahedra, or None."""
        X = np.asarray(x0s, dtype=np.float64)
        batched = X.ndim == 3
        Xb = X if batched else X[None]
        if Xb.ndim != 3 or Xb.shape[1:] not in ((3, 2), (4, 3)):
            return None
        B, meas = _simplex_B(Xb)
        if Xb.shape[2] == 2:
            D = _plane(self.E, self.nu, self.type == 0)
            meas = meas * self.thickness
        else:
            D = _iso_3d(self.E, self.nu)
        return B, meas, D, batched

    def _tensor_batch(self, x0s):
        """(dim, X, D, per-Gauss-point (B, det J) list) of a batch of quads or hexes."""
        X = np.asarray(x0s, dtype=np.float64)
        if X.shape[1:] not in ((8, 3), (4, 2)):
            raise ValueError("Invalid input shape")
        dim = X.shape[2]
        D = _iso_3d(self.E, self.nu) if dim == 3 else _plane(self.E, self.nu, self.type == 0)
        pts = [_tensor_B_batch(X, xi) for xi in _gauss_points(dim)]
        for _, det in pts:
            if np.any(det <= 0):
                bad = np.flatnonzero(det <= 0)
                raise Exception(f"Node Order Is Not Correct for elements: {bad.tolist()[:10]}")
        return dim, X, D, pts

    def K(self, x0s):
        sx = self._simplex(x0s)
        if sx is not None:
            B, meas, D, batched = sx
            K = np.matmul(B.transpose(0, 2, 1), np.matmul(D, B)) * meas[:, None, None]
            return K if batched else K[0]
        if np.asarray(x0s).ndim == 3:
            dim, X, D, pts = self._tensor_batch(x0s)
            K = 0.0
            for B, det in pts:
                K = K + np.matmul(B.transpose(0, 2, 1), np.matmul(D, B)) * det[:, None, None]
            return K * (self.thickness if dim == 2 else 1.0)
        dim, X = self._check(x0s)
        if dim == 3:
            C = _iso_3d(self.E, self.nu)
            K = np.zeros((24, 24))
            for a in (-_G, _G):
                for b in (-_G, _G):
                    for c in (-_G, _G):
                        B, detJ = _hex_B(X, np.array([a, b, c]))
                        if detJ <= 0:
                            raise Exception("Node Order Is Not Correct for elements: [0]")
                        K += B.T @ C @ B * detJ
            return K
        D = _plane(self.E, self.nu, self.type == 0)
        K = np.zeros((8, 8))
        for a in (-_G, _G):
            for b in (-_G, _G):
                B, detJ = _quad_B(X, np.array([a, b]))
                if detJ <= 0:
                    raise Exception("Node Order Is Not Correct for elements: [0]")
                K += B.T @ D @ B * detJ * self.thickness
        return K

    def locals(self, x0s):
        """[constitutive matrix, centroid-averaged B] as the reference returns."""
        sx = self._simplex(x0s)
        if sx is not None:
            B, meas, D, batched = sx
            return [D, B] if batched else [D, B[0]]
        if np.asarray(x0s).ndim == 3:
            dim, X, D, pts = self._tensor_batch(x0s)
            return [D, sum(B for B, _ in pts) / len(pts)]
        dim, X = self._check(x0s)
        if dim == 3:
            C = _iso_3d(self.E, self.nu)
            Bm = np.zeros((6, 24))
            for a in (-_G, _G):
                for b in (-_G, _G):
                    for c in (-_G, _G):
                        Bm += _hex_B(X, np.array([a, b, c]))[0] / 8
            return [C, Bm]
        D = _plane(self.E, self.nu, self.type == 0)
        Bm = np.zeros((3, 8))
        for a in (-_G, _G):
            for b in (-_G, _G):
                Bm += _quad_B(X, np.array([a, b]))[0] / 4
        return [D, Bm]

    def volume(self, x0s):
        sx = self._simplex(x0s)
        if sx is not None:
            B, meas, D, batched = sx
            if np.asarray(x0s).shape[-1] == 2:
                meas = meas / self.thickness   # the area, as for quads
            return meas if batched else float(meas[0])
        if np.asarray(x0s).ndim == 3:
            dim, X, D, pts = self._tensor_batch(x0s)
            return sum(det for _, det in pts)
        dim, X = self._check(x0s)
        if dim == 3:
            v = 0.0
            for a in (-_G, _G):
                for b in (-_G, _G):
                    for c in (-_G, _G):
                        v += _hex_B(X, np.array([a, b, c]))[1]
            return v
        v = 0.0
        for a in (-_G, _G):
            for b in (-_G, _G):
                v += _quad_B(X, np.array([a, b]))[1]
        return v
