"""CPU checks that pin the float64 conv oracle (oracle/conv_oracle.py):
adjoint identities for the input and weight gradients, a brute-force
convolution on a tiny volume, and the transfer pair R = P^T / 8."""
import numpy as np

from oracle import conv_oracle as C


def _r(shape, s):
    return np.random.default_rng(s).standard_normal(shape)


def test_conv_bruteforce():
    x = _r((3, 4, 5, 2), 0)
    w = _r((27, 3, 2), 1)
    b = _r(3, 2)
    out = C.conv3d(x, w, b)
    D, H, W, _ = x.shape
    ref = np.zeros((D, H, W, 3))
    for j in range(D):
        for i in range(H):
            for k in range(W):
                s = b.copy()
                for t, (dy, dx, dz) in enumerate(C.OFFSETS):
                    jj, ii, kk = j + dy, i + dx, k + dz
                    if 0 <= jj < D and 0 <= ii < H and 0 <= kk < W:
                        s += w[t] @ x[jj, ii, kk]
                ref[j, i, k] = s
    assert np.abs(out - ref).max() < 1e-12


def test_conv_adjoints():
    x = _r((5, 3, 4, 3), 3)
    w = _r((27, 2, 3), 4)
    y = _r((5, 3, 4, 2), 5)
    lhs = np.sum(C.conv3d(x, w) * y)
    assert abs(lhs - np.sum(x * C.conv3d_bwd_data(y, w))) < 1e-10 * abs(lhs)
    assert abs(lhs - np.sum(w * C.conv3d_bwd_weight(x, y))) < 1e-10 * abs(lhs)


def test_transfer_pair():
    x = _r((5, 7, 3, 2), 6)
    xc = _r((3, 4, 2, 2), 7)
    lhs = np.sum(C.restrict_vol(x, 1.0) * xc)
    assert abs(lhs - np.sum(x * C.prolong_vol(xc, x.shape[:3], 1.0))) < 1e-10 * abs(lhs)
    ones = np.ones((3, 4, 2, 1))
    assert np.allclose(C.prolong_vol(ones, (5, 7, 3)), 1.0)


def test_bf16_round():
    v = np.array([1.0, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 1e-3])
    r = C.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == -2.5
    assert abs(r[4] - 1e-3) <= 2 ** -9 * 1e-3


def _net_params(seed=0):
    from oracle import unet_oracle as U
    rng = np.random.default_rng(seed)
    W, B = {}, {}
    for name, ci, co, lev, relu in U.LAYERS:
        W[name] = rng.standard_normal((27, co, ci)) * np.sqrt(2.0 / (27 * ci))
        B[name] = rng.standard_normal(co) * 0.1
    return W, B


def test_unet_oracle_backward_fd():
    """The hand-written network backward equals finite differences (float64)."""
    from oracle import unet_oracle as U
    W, B = _net_params(0)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((9, 9, 9, 16))
    t = rng.standard_normal((9, 9, 9, 16))
    y, cache = U.forward(x, W, B, rounding=False)
    gW, gB = U.backward(t, W, cache, rounding=False)
    h = 1e-6
    for name in ("e0", "e3", "d1", "out"):
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in W[name].shape)
            W[name][idx] += h
            lp = np.sum(U.forward(x, W, B, rounding=False)[0] * t)
            W[name][idx] -= 2 * h
            lm = np.sum(U.forward(x, W, B, rounding=False)[0] * t)
            W[name][idx] += h
            fd = (lp - lm) / (2 * h)
            assert abs(fd - gW[name][idx]) <= 1e-5 * max(1.0, abs(fd)), (name, fd, gW[name][idx])
        c = rng.integers(0, B[name].shape[0])
        B[name][c] += h
        lp = np.sum(U.forward(x, W, B, rounding=False)[0] * t)
        B[name][c] -= 2 * h
        lm = np.sum(U.forward(x, W, B, rounding=False)[0] * t)
        B[name][c] += h
        assert abs((lp - lm) / (2 * h) - gB[name][c]) <= 1e-5 * max(1.0, abs(gB[name][c]))
