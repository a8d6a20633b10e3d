"""MinimumCompliance: the SIMP compliance problem on the device.

Mirrors Problem/CPU/MinimumCompliance.py:9-657 (CUDA twin
Problem/CUDA/MinimumCompliance.py) for a single material: constructor
arguments, ``init_desvars``, ``set_desvars``, ``get_desvars``, ``penalize``,
``penalize_grad``, ``f``, ``nabla_f``, ``g``, ``nabla_g``, ``bounds``,
``constraint_map``, ``is_independent``, ``ill_conditioned``, ``is_terminal``
and ``logs``.  The chain of ``_compute`` -- filter, interpolation, solve,
compliance, sensitivity, chain rule, filter adjoint -- runs as gfx950 kernels.

``law`` selects the material interpolation: ``'reference'`` is the
reference's clip(H(rho)^p, void, 1) * E_mul; ``'modified'`` is
Emin + H(rho)^p (E0 - Emin) with Emin = ``void`` and E0 = ``E_mul``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


class MinimumCompliance:
    def __init__(self, FE, filter, E_mul=[1.0], void=1e-6, penalty=3.0, volume_fraction=[0.25],
                 penalty_schedule=None, heavyside=True, beta=2, eta=0.5, law="reference"):
        if len(E_mul) != len(volume_fraction):
            raise ValueError("E and volume_fraction must have the same length.")
        if len(E_mul) != 1:
            raise NotImplementedError("multi-material interpolation is not part of the device path")
        if law not in ("reference", "modified"):
            raise ValueError("law must be 'reference' or 'modified'")
        self.is_single_material = True
        self.E_mul = float(E_mul[0])
        self.volume_fraction = float(volume_fraction[0])
        self.n_material = 1
        self.void = float(void)
        self.penalty = penalty
        self.penalty_schedule = penalty_schedule
        self.heavyside = heavyside
        self.beta = beta
        self.eta = eta
        self.law = law
        self.filter = filter
        self.FE = FE
        self.dtype = FE.dtype
        self.iteration = 0
        self.desvars = None
        self._f = None
        self._g = None
        self._nabla_f = None
        self._residual = None
        self.nel = FE.mesh.n_elements
        self.num_vars = self.nel
        dev = FE.kernel.device
        self.device = dev
        self._nabla_g = torch.full((self.num_vars,), float(FE.mesh.As[0] / FE.mesh.volume), dtype=torch.float64,
                                   device=dev)
        self.is_3D = FE.is_3D
        self._E = torch.empty(self.nel, dtype=torch.float64, device=dev)
        self._dE = torch.empty(self.nel, dtype=torch.float64, device=dev)
        self._dr = torch.empty(self.nel, dtype=torch.float64, device=dev)
        self._scal = torch.zeros(_lib.N_SCALARS, dtype=torch.float64, device=dev)
        self._partials = torch.zeros(max(_lib.load().fantom_vec_partials(self.nel),
                                         _lib.load().fantom_vec_partials(FE.kernel.shape[0])),
                                     dtype=torch.float64, device=dev)

    # ---- problem description -----------------------------------------------------------
    def N(self):
        return self.num_vars

    def m(self):
        return 1

    def is_independent(self):
        return True

    def constraint_map(self):
        return 1

    def bounds(self):
        return (0.0, 1.0)

    # ---- design variables --------------------------------------------------------------
    def init_desvars(self):
        self.desvars = torch.full((self.num_vars,), self.volume_fraction, dtype=torch.float64, device=self.device)
        self.iteration = 0
        self._compute()

    def set_desvars(self, desvars):
        if desvars.shape[0] != self.num_vars:
            raise ValueError(f"Expected {self.num_vars} design variables, got {desvars.shape[0]}.")
        self.desvars = desvars
        self._compute()
        self.iteration += 1

    def get_desvars(self):
        return self.desvars

    # ---- interpolation (A4) ------------------------------------------------------------
    def _pen_args(self):
        pen = self.penalty
        if self.penalty_schedule is not None:
            pen = self.penalty_schedule(self.penalty, self.iteration)
        beta_val = self.beta(self.iteration) if callable(self.beta) else self.beta
        law = 0 if self.law == "reference" else 1
        return law, float(pen), self.void, self.E_mul, int(bool(self.heavyside)), float(beta_val), float(self.eta)

    def penalize(self, rho):
        out = torch.empty_like(rho)
        law, p, lo, hi, hv, b, eta = self._pen_args()
        _lib.call("fantom_simp_f64", rho.numel(), _lib.ptr(rho), law, p, lo, hi, hv, b, eta, _lib.ptr(out), None)
        return out

    def penalize_grad(self, rho):
        out = torch.empty_like(rho)
        law, p, lo, hi, hv, b, eta = self._pen_args()
        _lib.call("fantom_simp_f64", rho.numel(), _lib.ptr(rho), law, p, lo, hi, hv, b, eta, None, _lib.ptr(out))
        return out

    def _dot(self, a, b):
        n = a.numel()
        _lib.call("fantom_dot_partials_f64", n, _lib.ptr(a), _lib.ptr(b), _lib.ptr(self._partials))
        _lib.call("fantom_finalize_f64", _lib.ptr(self._partials), _lib.load().fantom_vec_partials(n),
                  _lib.ptr(self._scal), _lib.FIN_STORE, 7)
        return self._scal[7]

    def _compute(self):
        """filter -> interpolate -> solve -> c = f.u -> dc -> chain rule -> filter^T
        (Problem/CPU/MinimumCompliance.py:466-498)."""
        rho = self.filter.dot(self.desvars)
        law, p, lo, hi, hv, b, eta = self._pen_args()
        _lib.call("fantom_simp_f64", rho.numel(), _lib.ptr(rho), law, p, lo, hi, hv, b, eta, _lib.ptr(self._E),
                  _lib.ptr(self._dE))
        U, residual = self.FE.solve(self._E)
        compliance = self._dot(self.FE.rhs, U)
        df = self.FE.kernel.process_grad(U)
        if df.dtype != torch.float64:   # float32 meshes: the chain rule stays FP64
            df = df.double()
        _lib.call("fantom_mul_f64", self.nel, _lib.ptr(self._dE), _lib.ptr(df), _lib.ptr(self._dr))
        self._nabla_f = self.filter._rmatvec(self._dr)
        self._f = float(compliance.item())
        self._residual = residual
        self._U = U
        self._rho = rho

    # ---- objective / constraint ----------------------------------------------------------
    def f(self, rho=None):
        if rho is None:
            return self._f
        return self._f + float(self._dot(rho.contiguous(), self._nabla_f).item())

    def nabla_f(self, rho=None):
        return self._nabla_f

    def g(self, rho=None):
        x = self.desvars if rho is None else rho
        vf = self._dot(self._nabla_g, x)
        return (vf - self.volume_fraction).reshape(1)

    def nabla_g(self):
        return self._nabla_g

    def ill_conditioned(self):
        return self._residual >= 1e-2

    def is_terminal(self):
        if self.penalty_schedule is not None:
            return self.penalty_schedule(self.penalty, self.iteration) == self.penalty
        return True

    # ---- post-processing ----------------------------------------------------------------
    def FEA(self, thresshold: bool = True):
        """Solve the thresholded design and evaluate per-element strain,
        stress, strain energy and von Mises stress
        (Problem/CPU/MinimumCompliance.py:659-736).  As the reference: the
        element moduli are (x > 0.5) + void (times E_mul), the solve gets four
        times the solver's iteration limit, stress = D strain with the unit
        modulus D and B averaged over the Gauss points, strain energy scaled by
        the modulus.  ``thresshold=False`` uses the raw design variables (the
        reference leaves that case undefined).  Returns a dict with 'strain',
        'stress', 'strain_energy', 'von_mises' (device tensors), 'compliance'
        and 'Displacements'."""
        if self.desvars is None:
            raise ValueError("Design variables are not initialized. Call init_desvars() or set_desvars() first.")
        x = self.get_desvars()
        rho = (x > 0.5).to(torch.float64) + self.void if thresshold else x.to(torch.float64).clone()
        rho = rho * self.E_mul
        solver = self.FE.solver
        maxiter = getattr(solver, "maxiter", None)
        if maxiter is not None:
            solver.maxiter = maxiter * 4
        try:
            U, residual = self.FE.solve(rho)
        finally:
            if maxiter is not None:
                solver.maxiter = maxiter
        compliance = self._dot(self.FE.rhs, U)
        if residual > 1e-5:
            print(f"Solver residual is above 1e-5 ({residual:.4e}). Consider higher iterations (rerun this function "
                  "and more iteration from prior solve will be applied).")
        mesh, k = self.FE.mesh, self.FE.kernel
        D, B = (torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.device) for a in mesh.locals)
        K0 = k.K0
        ns = D.shape[0]
        strain = torch.empty((self.nel, ns), dtype=torch.float64, device=self.device)
        stress = torch.empty_like(strain)
        energy = torch.empty(self.nel, dtype=torch.float64, device=self.device)
        vm = torch.empty_like(energy)
        U = k._as_vec(U, torch.float64)
        _lib.call("fantom_fea_locals_f64", *k.grid, _lib.ptr(K0), _lib.ptr(D), _lib.ptr(B), _lib.ptr(rho),
                  _lib.ptr(U), _lib.ptr(strain), _lib.ptr(stress), _lib.ptr(energy), _lib.ptr(vm))
        return {"strain": strain, "stress": stress, "strain_energy": energy, "von_mises": vm,
                "compliance": float(compliance.item()), "Displacements": U}

    def logs(self):
        """The reference's keys (Problem/CPU/MinimumCompliance.py:639-657) plus
        the solve's PCG iteration count, precision and FP64 fallback."""
        out = {"iteration": int(self.iteration), "residual": float(self._residual)}
        solver = self.FE.solver
        out["cg_iterations"] = int(getattr(solver, "last_iterations", 0))
        out["precision"] = getattr(solver, "precision", "fp64")
        if getattr(solver, "last_fallback", False):
            out["fp64_fallback_iterations"] = int(solver.last_fallback_iterations)
        return out
