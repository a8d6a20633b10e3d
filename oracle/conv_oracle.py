"""Float64 numpy reference for the learned preconditioner's building blocks.

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker).

The learned preconditioner has no counterpart in the reference, so this file
restates the mathematics directly: 3x3x3 convolution with zero padding on a
channels-last volume (D, H, W, C), its adjoints with respect to the input
and the weights, the bias gradient, and the vertex-centred two-grid
transfers on volumes.  Each function is pinned in tests/test_conv_oracle.py
by finite differences and adjoint identities (<A x, y> = <x, A^T y>).

Weight layout: W[tap, c_out, c_in] with tap = (dy+1)*9 + (dx+1)*3 + (dz+1),
offsets (dy, dx, dz) in {-1, 0, 1}^3 along (D, H, W).
"""
import numpy as np

OFFSETS = [(dy, dx, dz) for dy in (-1, 0, 1) for dx in (-1, 0, 1) for dz in (-1, 0, 1)]


def _shift(x, off):
    """y[v] = x[v + off] with zeros outside the volume."""
    D, H, W = x.shape[:3]
    y = np.zeros_like(x)
    dy, dx, dz = off
    src = [slice(max(0, o), min(n, n + o)) for o, n in zip(off, (D, H, W))]
    dst = [slice(max(0, -o), min(n, n - o)) for o, n in zip(off, (D, H, W))]
    y[dst[0], dst[1], dst[2]] = x[src[0], src[1], src[2]]
    return y


def conv3d(x, w, b=None):
    """out[v, co] = sum_tap sum_ci w[tap, co, ci] x[v + off(tap), ci] + b[co]."""
    D, H, W, Ci = x.shape
    Co = w.shape[1]
    out = np.zeros((D, H, W, Co))
    for t, off in enumerate(OFFSETS):
        out += _shift(x, off) @ w[t].T
    if b is not None:
        out += b
    return out


def conv3d_bwd_data(dy, w):
    """Adjoint of conv3d in x: dx[v, ci] = sum_tap sum_co w[tap, co, ci] dy[v - off(tap), co]."""
    D, H, W, Co = dy.shape
    Ci = w.shape[2]
    dx = np.zeros((D, H, W, Ci))
    for t, off in enumerate(OFFSETS):
        dx += _shift(dy, tuple(-o for o in off)) @ w[t]
    return dx


def conv3d_bwd_weight(x, dy):
    """dw[tap, co, ci] = sum_v dy[v, co] x[v + off(tap), ci]."""
    Co, Ci = dy.shape[3], x.shape[3]
    dw = np.zeros((27, Co, Ci))
    for t, off in enumerate(OFFSETS):
        xs = _shift(x, off).reshape(-1, Ci)
        dw[t] = dy.reshape(-1, Co).T @ xs
    return dw


def restrict_vol(x, scale=1.0 / 8.0):
    """Vertex-centred full weighting (coarse node c <- fine node 2c, weights
    2^-dist over the 3^3 neighbourhood), times scale."""
    D, H, W, C = x.shape
    Dc, Hc, Wc = (D - 1) // 2 + 1, (H - 1) // 2 + 1, (W - 1) // 2 + 1
    out = np.zeros((Dc, Hc, Wc, C))
    for jc in range(Dc):
        for ic in range(Hc):
            for kc in range(Wc):
                s = np.zeros(C)
                for dj in (-1, 0, 1):
                    for di in (-1, 0, 1):
                        for dk in (-1, 0, 1):
                            j, i, k = 2 * jc + dj, 2 * ic + di, 2 * kc + dk
                            if 0 <= j < D and 0 <= i < H and 0 <= k < W:
                                s += 0.5 ** (abs(dj) + abs(di) + abs(dk)) * x[j, i, k]
                out[jc, ic, kc] = s * scale
    return out


def prolong_vol(xc, fine_shape, scale=1.0):
    """Trilinear interpolation from the coarse vertex grid, times scale (the
    transpose of restrict_vol up to the scale factor)."""
    D, H, W = fine_shape
    C = xc.shape[3]
    out = np.zeros((D, H, W, C))
    for j in range(D):
        for i in range(H):
            for k in range(W):
                s = np.zeros(C)
                for cj in {j // 2, (j + 1) // 2}:
                    for ci in {i // 2, (i + 1) // 2}:
                        for ck in {k // 2, (k + 1) // 2}:
                            w = 0.5 ** ((j % 2) + (i % 2) + (k % 2))
                            s += w * xc[cj, ci, ck]
                out[j, i, k] = s * scale
    return out


def bf16_round(a):
    """Round float64 values to the nearest bfloat16 (ties to even)."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
