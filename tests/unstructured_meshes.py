"""Unstructured test meshes: jittered triangle / tetrahedron / hex grids and
mixed-element meshes, each element with positive orientation.  Shared by the
CPU and GPU tests of the unstructured operators."""
from __future__ import annotations

import numpy as np


def _grid2(nx, ny, lx=1.0, ly=1.0):
    x, y = np.meshgrid(np.linspace(0, lx, nx + 1), np.linspace(0, ly, ny + 1))
    return np.stack([x.ravel(), y.ravel()], axis=1)


def _jitter(X, h, amount, rng):
    """Move interior nodes by up to amount*h in each coordinate."""
    lo, hi = X.min(axis=0), X.max(axis=0)
    interior = np.all((X > lo + 1e-12) & (X < hi - 1e-12), axis=1)
    X = X.copy()
    X[interior] += rng.uniform(-amount * h, amount * h, size=(interior.sum(), X.shape[1]))
    return X


def tri_mesh(nx, ny, jitter=0.2, seed=0):
    """Rectangle [0,2] x [0,1] cut into 2 nx ny counter-clockwise triangles."""
    rng = np.random.default_rng(seed)
    X = _jitter(_grid2(nx, ny, 2.0, 1.0), min(2.0 / nx, 1.0 / ny), jitter, rng)
    nid = lambda i, j: j * (nx + 1) + i
    el = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if (i + j) % 2:
                el += [[a, b, c], [a, c, d]]
            else:
                el += [[a, b, d], [b, c, d]]
    return X, np.array(el, dtype=np.int64)


def quad_tri_mesh(nx, ny, jitter=0.2, seed=0):
    """Mixed mesh: quads on the left half, triangle pairs on the right."""
    rng = np.random.default_rng(seed)
    X = _jitter(_grid2(nx, ny, 2.0, 1.0), min(2.0 / nx, 1.0 / ny), jitter, rng)
    nid = lambda i, j: j * (nx + 1) + i
    el = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if i < nx // 2:
                el.append([a, b, c, d])
            else:
                el += [[a, b, c], [a, c, d]]
    return X, el


def _grid3(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    g = np.stack(np.meshgrid(np.linspace(0, lx, nx + 1), np.linspace(0, ly, ny + 1),
                             np.linspace(0, lz, nz + 1), indexing="ij"), axis=-1)
    return g.reshape(-1, 3)


def _nid3(ny, nz):
    return lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k


def _hex_nodes(n, i, j, k):
    """Corners of cell (i, j, k) in the hex8 order of the element matrices:
    bottom face (z = k) counter-clockwise, then the top face."""
    return [n(i, j, k), n(i + 1, j, k), n(i + 1, j + 1, k), n(i, j + 1, k),
            n(i, j, k + 1), n(i + 1, j, k + 1), n(i + 1, j + 1, k + 1), n(i, j + 1, k + 1)]


# the six tetrahedra of a cube along its 0-6 diagonal (Kuhn subdivision),
# as hex-corner indices
_KUHN = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]


def orient(X, el):
    """Swap the last two nodes of every negatively oriented simplex."""
    el = np.array(el, dtype=np.int64)
    P = X[el]
    d = P[:, 1:] - P[:, :1]
    det = np.linalg.det(d) if el.shape[1] == 4 else d[:, 0, 0] * d[:, 1, 1] - d[:, 0, 1] * d[:, 1, 0]
    neg = det < 0
    el[neg, -2], el[neg, -1] = el[neg, -1].copy(), el[neg, -2].copy()
    return el


def tet_mesh(nx, ny, nz, jitter=0.15, seed=0):
    """Box [0,2] x [0,1] x [0,1] cut into 6 nx ny nz tetrahedra."""
    rng = np.random.default_rng(seed)
    X = _jitter(_grid3(nx, ny, nz, 2.0, 1.0, 1.0), min(2.0 / nx, 1.0 / ny, 1.0 / nz), jitter, rng)
    n = _nid3(ny, nz)
    el = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                h = _hex_nodes(n, i, j, k)
                el += [[h[t] for t in tet] for tet in _KUHN]
    return X, orient(X, el)


def hex_mesh(nx, ny, nz, jitter=0.15, seed=0):
    """Jittered hex8 mesh of the box [0,2] x [0,1] x [0,1]."""
    rng = np.random.default_rng(seed)
    X = _jitter(_grid3(nx, ny, nz, 2.0, 1.0, 1.0), min(2.0 / nx, 1.0 / ny, 1.0 / nz), jitter, rng)
    n = _nid3(ny, nz)
    el = [_hex_nodes(n, i, j, k) for i in range(nx) for j in range(ny) for k in range(nz)]
    return X, np.array(el, dtype=np.int64)


def hex_tet_mesh(nx, ny, nz, jitter=0.1, seed=0):
    """Mixed 3D mesh: hexes for x < half, the rest cut into tetrahedra."""
    rng = np.random.default_rng(seed)
    X = _jitter(_grid3(nx, ny, nz, 2.0, 1.0, 1.0), min(2.0 / nx, 1.0 / ny, 1.0 / nz), jitter, rng)
    n = _nid3(ny, nz)
    el = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                h = _hex_nodes(n, i, j, k)
                if i < nx // 2:
                    el.append(h)
                else:
                    el += [list(r) for r in orient(X, [[h[t] for t in tet] for tet in _KUHN])]
    return X, el
