"""Optimisers: OC on the device.

``OC`` mirrors Optimizers/CPU/OC.py:6-207 (CUDA twin Optimizers/CUDA/OC.py):
``OC(problem, move=0.2, change_tol=1e-4, fun_tol=1e-6, timer=False)`` with
``iter()``, ``converged()`` and ``logs()``.  The bisection for the Lagrange
multiplier runs on the device (fantom_oc_update_f64) and reproduces the
reference's sequence of midpoints; one host read per iteration returns the
design change.
"""
from __future__ import annotations

import math
import time

import torch

from . import _lib


class Optimizer:
    """Optimizers/_optimizer.py:4-29: binding to a problem initialises it."""

    def __init__(self, problem, *args, **kwargs):
        self.problem = problem
        self.desvars = problem.init_desvars()

    def iter(self, *args, **kwargs):
        raise NotImplementedError

    def converged(self, *args, **kwargs):
        raise NotImplementedError

    def logs(self, *args, **kwargs):
        raise NotImplementedError


class OC(Optimizer):
    def __init__(self, problem, move: float = 0.2, change_tol=1e-4, fun_tol=1e-6, timer=False):
        super().__init__(problem)
        if not problem.is_independent():
            raise ValueError("OC optimizer requires a problem with independent constraints.")
        self.last_desvars = problem.get_desvars().clone()
        self.last_f = problem.f()
        self.m = problem.m()
        self.ocP = None
        self.lambda_map = problem.constraint_map()
        self.move = move
        self.bounds = problem.bounds()
        if self.bounds != (0.0, 1.0):
            raise NotImplementedError("the device OC update assumes design bounds [0, 1]")
        self.change = math.inf
        self.change_f = math.inf
        self.change_tol = change_tol
        self.fun_tol = fun_tol
        self.timer = timer
        self.iteration = 0
        n = problem.N()
        dev = problem.device
        self._ocp = torch.empty(n, dtype=torch.float64, device=dev)
        self._state = torch.zeros(16, dtype=torch.float64, device=dev)
        self._partials = torch.zeros(16 * 1024, dtype=torch.float64, device=dev)
        self._buf = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(2)]
        self.last_lambda = None

    def iter(self):
        """One OC step (Optimizers/CPU/OC.py:97-158)."""
        if self.timer:
            torch.cuda.synchronize()
            t0 = time.time()
        p = self.problem
        x = p.get_desvars()
        xnew = self._buf[0] if self._buf[0].data_ptr() != x.data_ptr() else self._buf[1]
        n = x.numel()
        _lib.call("fantom_oc_update_f64", n, _lib.ptr(x), _lib.ptr(p.nabla_f()), _lib.ptr(p.nabla_g()),
                  float(p.volume_fraction), float(self.move), _lib.ptr(self._ocp), _lib.ptr(self._state),
                  _lib.ptr(self._partials), _lib.ptr(xnew))
        self.ocP = self._ocp
        self.iteration += 1
        if self.timer:
            torch.cuda.synchronize()
            t1 = time.time()
        # change = ||x_old - x_new|| from the device partials
        np_ = _lib.load().fantom_oc_partials(n)
        _lib.call("fantom_finalize_f64", _lib.ptr(self._partials), np_, _lib.ptr(self._state), _lib.FIN_STORE, 8)
        p.set_desvars(xnew)
        st = self._state.cpu()
        self.last_lambda = float(st[2])
        self.change = math.sqrt(max(float(st[8]), 0.0))
        fnew = p.f()
        self.change_f = abs((fnew - self.last_f) / fnew)
        self.last_f = fnew
        self.last_desvars = xnew
        if self.timer:
            return t1 - t0

    def converged(self, *args, **kwargs):
        if not self.problem.is_terminal():
            return False
        return self.change <= self.change_tol and self.change_f <= self.fun_tol

    def logs(self):
        return {"objective": float(self.last_f), "variable change": float(self.change),
                "function change": float(self.change_f), **self.problem.logs()}
