"""Element physics: isotropic linear elasticity on bilinear quads / trilinear hexes.

Mirrors the constructor and the ``K`` / ``locals`` / ``volume`` methods of the
reference's ``LinearElasticity`` (physics/LinearElasticity.py:74-118).  The
element matrix is integrated with 2-point Gauss quadrature per axis, which is
exact for the polynomial integrand of an affine (box) element; the node order
is the reference's (counter-clockwise bottom face, then top face).  This runs
once at mesh setup on the host.
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

    def K(self, x0s):
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
