"""Optimisers: OC and MMA on the device.

``OC`` mirrors Optimizers/CPU/OC.py:6-207 (CUDA twin Optimizers/CUDA/OC.py):
``OC(problem, move=0.2, change_tol=1e-4, fun_tol=1e-6, timer=False)`` with
``iter()``, ``converged()`` and ``logs()``.  The bisection for the Lagrange
multiplier runs on the device (fantom_oc_update_f64) and reproduces the
reference's sequence of midpoints; one host read per iteration returns the
design change.
"""
from __future__ import annotations

import ctypes
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


class MMA(Optimizer):
    """Method of Moving Asymptotes, Optimizers/CPU/MMA.py:7-229 (CUDA twin
    Optimizers/CUDA/MMA.py): ``MMA(problem, sub_tol=1e-7, sub_maxiter=100,
    change_tol=1e-4, fun_tol=1e-6, move=0.5, timer=False)`` with ``iter()``,
    ``converged()`` and ``logs()``.  The subproblem (mmasub/subsolv,
    mma/CPU/_mma.py) runs in fantom_mma_update_f64: per-variable work in
    HIP kernels, the m-dimensional dual system on the host.  a0 = 1, a = 0,
    c = 1e5, d = 0 as the reference passes them."""

    def __init__(self, problem, sub_tol=1e-7, sub_maxiter=100, change_tol=1e-4, fun_tol=1e-6, move=0.5,
                 timer=False):
        super().__init__(problem)
        self.last_desvars = problem.get_desvars().clone()
        self.last_f = problem.f()
        self.m = problem.m()
        self.ocP = None
        self.lambda_map = problem.constraint_map()
        self.bounds = problem.bounds()
        self.change = math.inf
        self.change_f = math.inf
        self.change_tol = change_tol
        self.fun_tol = fun_tol
        self.iteration = 0
        self.move = move
        self.sub_tol = sub_tol
        self.sub_maxiter = sub_maxiter
        x = problem.get_desvars()
        n = x.numel()
        self.N = problem.N()
        self.x_1 = x.clone()
        self.x_2 = x.clone()
        self.low = torch.full_like(x, float(self.bounds[0]))
        self.upp = torch.full_like(x, float(self.bounds[1]))
        self.comp_base = problem.f()
        self.timer = timer
        self.last_newton_steps = 0
        nbytes = _lib.load().fantom_mma_workspace_size(n, self.m)
        if nbytes < 0:
            raise ValueError(f"the device MMA supports 1 <= m <= 8 constraints, got {self.m}")
        self._work = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=x.device)
        self._buf = [torch.empty_like(x) for _ in range(2)]
        self._a = (ctypes.c_double * self.m)(*([0.0] * self.m))
        self._c = (ctypes.c_double * self.m)(*([1e5] * self.m))
        self._d = (ctypes.c_double * self.m)(*([0.0] * self.m))

    def iter(self):
        """One MMA step (Optimizers/CPU/MMA.py:103-180)."""
        if self.timer:
            torch.cuda.synchronize()
            t0 = time.time()
        p = self.problem
        x = p.get_desvars()
        n = x.numel()
        df = p.nabla_f()
        dg = p.nabla_g().reshape(self.m, n).contiguous()
        fval = p.g().reshape(-1).double().cpu().numpy()
        fv = (ctypes.c_double * self.m)(*[float(v) for v in fval])
        xnew = self._buf[0] if self._buf[0].data_ptr() != x.data_ptr() else self._buf[1]
        steps = ctypes.c_int(0)
        _lib.call("fantom_mma_update_f64", n, self.m, self.iteration + 1, _lib.ptr(x), _lib.ptr(self.x_1),
                  _lib.ptr(self.x_2), float(self.bounds[0]), float(self.bounds[1]), _lib.ptr(df), fv, _lib.ptr(dg),
                  _lib.ptr(self.low), _lib.ptr(self.upp), float(self.move), 1.0, self._a, self._c, self._d,
                  float(self.sub_tol), int(self.sub_maxiter), _lib.ptr(xnew), _lib.ptr(self._work),
                  self._work.numel() * 8, ctypes.byref(steps))
        self.last_newton_steps = steps.value
        self.x_2.copy_(self.x_1)
        self.x_1.copy_(x)
        self.iteration += 1
        if self.timer:
            torch.cuda.synchronize()
            t1 = time.time()
        if bool(torch.isnan(xnew).any()):
            xnew = torch.clamp(x + self.sub_tol, float(self.bounds[0]), float(self.bounds[1]))
        old = self.last_desvars
        p.set_desvars(xnew)
        self.change = float(torch.linalg.norm(old - xnew))
        fnew = p.f()
        self.change_f = abs((fnew - self.last_f) / fnew)
        self.last_f = fnew
        self.last_desvars = xnew.clone()
        if self.timer:
            return t1 - t0

    def converged(self, *args, **kwargs):
        if not self.problem.is_terminal():
            return False
        return self.change <= self.change_tol and self.change_f <= self.fun_tol

    def logs(self):
        return {"objective": float(self.last_f), "variable change": float(self.change),
                "function change": float(self.change_f), **self.problem.logs()}
