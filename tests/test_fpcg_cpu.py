"""Flexible PCG control logic (pyfantom_amd/learned.py:fpcg_solve) on the CPU
with a small SPD matrix and a vector backend in torch: convergence with a good
preconditioner, restarts when the preconditioner breaks down (z = 0), and the
hand-over to the fallback solver."""
import math

import numpy as np
import pytest
import torch

from pyfantom_amd import _lib
from pyfantom_amd.learned import fpcg_solve


class _Vec:
    """The device-vector contract of learned._Vec in torch on the CPU: the
    scalars live in scal and the finalize ops (store, alpha = r.z / p.Kp or 0
    on a breakdown) act as the device finalize kernel does."""
    S_A, S_B, S_C = 8, 9, 10

    def __init__(self, A):
        self.A = A
        self.scal = torch.zeros(16, dtype=torch.float64)

    def _fin(self, v, op, slot):
        if op == _lib.FIN_STORE:
            self.scal[slot] = v
        elif op == _lib.FIN_ALPHA:
            self.scal[_lib.S_PQ] = v
            self.scal[_lib.S_ALPHA] = float(self.scal[_lib.S_RZ]) / v if (v > 0.0 and math.isfinite(v)) else 0.0
        else:
            raise AssertionError(f"unexpected finalize op {op}")

    def read(self, *slots):
        return [float(self.scal[s]) for s in slots]

    def residual(self, f, x, r):
        r.copy_(f - self.A @ x)

    def dot(self, a, b):
        return float(torch.dot(a, b))

    def ku(self, u, y):
        y.copy_(self.A @ u)

    def ku_dot(self, u, y, op=0, slot=8):
        y.copy_(self.A @ u)
        self._fin(float(torch.dot(u, y)), op, slot)

    def dot_to(self, a, b, op=0, slot=8):
        self._fin(float(torch.dot(a, b)), op, slot)

    def dot2_to(self, a, b, c, op_b, slot_b, op_c, slot_c):
        self._fin(float(torch.dot(a, b)), op_b, slot_b)
        self._fin(float(torch.dot(a, c)), op_c, slot_c)

    def xr(self, x, r, p, q):
        alpha = float(self.scal[_lib.S_ALPHA])
        x.add_(alpha * p)
        r.sub_(alpha * q)
        self.scal[_lib.S_RR] = float(torch.dot(r, r))

    def axpby(self, a, x, b, y):
        y.mul_(b).add_(a * x)


class _Precond:
    def __init__(self, A, kind):
        self.vec = _Vec(A)
        self.d = torch.diag(A)
        self.kind = kind
        self.calls = 0

    def __call__(self, r, z):
        self.calls += 1
        if self.kind == "jacobi":
            z.copy_(r / self.d)
        elif self.kind == "identity":
            z.copy_(r)
        else:
            z.zero_()


class _Fallback:
    def __init__(self, A):
        self.A = A
        self.last_iterations = 0
        self.tol = None
        self.last_x0 = None

    def solve(self, f, use_last=True):
        x = torch.linalg.solve(self.A, f)
        self.last_iterations = 1
        return x, 0.0


def _spd(n=60, seed=0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.linspace(1.0, 30.0, n)) @ Q.T
    return torch.from_numpy(A), torch.from_numpy(rng.standard_normal(n))


def test_fpcg_converges_with_jacobi():
    A, f = _spd()
    x, info = fpcg_solve(None, f, _Precond(A, "jacobi"), tol=1e-10, maxiter=200)
    assert info["residual"] < 1e-10 and not info["fallback"] and info["restarts"] == 0
    assert torch.allclose(x, torch.linalg.solve(A, f), atol=1e-8)


def test_fpcg_breakdown_restarts_without_fallback():
    A, f = _spd(seed=1)
    x, info = fpcg_solve(None, f, _Precond(A, "zero"), tol=1e-8, maxiter=2000)
    assert info["restarts"] > 0
    assert info["residual"] < 1e-8
    assert np.isfinite(x.numpy()).all()


def test_fpcg_breakdown_hands_over_to_fallback():
    A, f = _spd(seed=2)
    fb = _Fallback(A)
    x, info = fpcg_solve(None, f, _Precond(A, "zero"), tol=1e-8, maxiter=50, fallback=fb)
    assert info["fallback"]
    assert torch.allclose(x, torch.linalg.solve(A, f), atol=1e-10)


def test_fpcg_pkp_breakdown_keeps_history_aligned():
    """An indefinite operator with the right-hand side on its negative
    eigenvector makes p.Kp < 0 on every iteration; without a fallback each
    one restarts, and the residual history keeps one entry per iteration so
    the stall test (which looks stall_window entries back) stays in range."""
    n = 30
    A = torch.diag(torch.cat([torch.tensor([-10.0]), torch.linspace(1.0, 5.0, n - 1, dtype=torch.float64)]))
    f = torch.full((n,), 0.01, dtype=torch.float64)
    f[0] = 1.0
    x, info = fpcg_solve(None, f, _Precond(A, "identity"), tol=1e-12, maxiter=12, stall_window=5)
    assert info["iterations"] == 12 and info["restarts"] == 12
    assert len(info["history"]) == info["iterations"] + 1
    assert np.isfinite(x.numpy()).all()


def test_energy_loss_gradient_fd():
    """The training loss of the learned correction: L(c) = 1 - (t.c)^2 /
    ((c.Kc)(t.e)) with t = K e is the energy fraction left by the best
    multiple of c (checked against the explicit minimisation), and its
    gradient matches central finite differences."""
    from pyfantom_amd.learned import energy_loss_grad
    A, _ = _spd(n=40, seed=5)
    rng = np.random.default_rng(6)
    e = torch.from_numpy(rng.standard_normal(40))
    t = A @ e
    c = torch.from_numpy(rng.standard_normal(40))
    L, g = energy_loss_grad(t, e, c, A @ c)
    gam = float(t @ c) / float(c @ (A @ c))
    r = e - gam * c
    assert abs(L - float(r @ (A @ r)) / float(e @ (A @ e))) <= 1e-12
    h = 1e-6
    fd = torch.zeros(40, dtype=torch.float64)
    for i in range(40):
        cp, cm = c.clone(), c.clone()
        cp[i] += h
        cm[i] -= h
        fd[i] = (energy_loss_grad(t, e, cp, A @ cp)[0] - energy_loss_grad(t, e, cm, A @ cm)[0]) / (2 * h)
    assert torch.allclose(g, fd, rtol=1e-5, atol=1e-8)
    # exact direction: nothing left
    L0, _ = energy_loss_grad(t, e, 3.0 * e, A @ (3.0 * e))
    assert abs(L0) <= 1e-12
