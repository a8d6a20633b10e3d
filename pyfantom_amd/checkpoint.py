"""Checkpoint / resume of a SIMP run in one ``.npz`` (plain arrays, no
pickled objects).

The reference has no save/load; the state a SIMP run carries between
iterations is exactly what is stored here:

* the design variables and the problem's last evaluation (objective,
  sensitivities, displacement, filtered densities, iteration count and solve
  residual -- Problem/CPU/MinimumCompliance.py:466-498);
* the optimiser: OC's last design, objective and change measures
  (Optimizers/CPU/OC.py:82-158); MMA's x_1, x_2, low, upp as well
  (Optimizers/CPU/MMA.py:94-99);
* the solver's warm start ``last_x0`` (solvers/CUDA/_solvers.py:654), which
  is the stored displacement.

``load`` puts that state back without solving, so that ``optimizer.iter()``
continues the run from where ``save`` left it; on the same device and build
the continued run is bit-identical to the uninterrupted one (the solver
hierarchy is rebuilt from the design on the next solve, as every solve
does).
"""
from __future__ import annotations

import numpy as np
import torch

__all__ = ["save", "load"]

_PROBLEM_TENSORS = ("desvars", "_nabla_f", "_U", "_rho")
_PROBLEM_SCALARS = ("iteration", "_f", "_residual")
_OPTIMIZERS = {
    "OC": (("last_desvars",), ("last_f", "change", "change_f", "iteration", "last_lambda")),
    "MMA": (("last_desvars", "x_1", "x_2", "low", "upp"),
            ("last_f", "change", "change_f", "iteration", "comp_base")),
}


def _kind(optimizer):
    if optimizer is None:
        return ""
    kind = type(optimizer).__name__
    if kind not in _OPTIMIZERS:
        raise ValueError(f"checkpointing supports OC and MMA optimisers, not {kind}")
    return kind


def save(path, problem, optimizer=None):
    """Write the run's state to ``path`` (numpy ``.npz``)."""
    kind = _kind(optimizer)
    out = {"format": np.array("pyfantom_amd-simp-checkpoint-1"), "optimizer": np.array(kind),
           "num_vars": np.array(int(problem.desvars.numel()))}
    for name in _PROBLEM_TENSORS:
        out["problem." + name] = getattr(problem, name).detach().cpu().numpy()
    for name in _PROBLEM_SCALARS:
        out["problem." + name] = np.array(float(getattr(problem, name)))
    if kind:
        tensors, scalars = _OPTIMIZERS[kind]
        for name in tensors:
            out["optimizer." + name] = getattr(optimizer, name).detach().cpu().numpy()
        for name in scalars:
            v = getattr(optimizer, name)
            out["optimizer." + name] = np.array(float("nan") if v is None else float(v))
    np.savez(path, **out)


def load(path, problem, optimizer=None):
    """Restore the state written by :func:`save` into ``problem`` (and
    ``optimizer``, which must be of the kind that was saved).  Raises
    ValueError when the file does not match the objects."""
    kind = _kind(optimizer)
    with np.load(path, allow_pickle=False) as z:
        if "format" not in z or str(z["format"]) != "pyfantom_amd-simp-checkpoint-1":
            raise ValueError(f"{path} is not a SIMP checkpoint of this package")
        if int(z["num_vars"]) != problem.desvars.numel():
            raise ValueError(f"checkpoint has {int(z['num_vars'])} design variables, the problem "
                             f"{problem.desvars.numel()}")
        if kind and str(z["optimizer"]) != kind:
            raise ValueError(f"checkpoint holds {str(z['optimizer']) or 'no'} optimiser state, not {kind}")
        data = {k: z[k] for k in z.files}
    dev = problem.desvars.device
    for name in _PROBLEM_TENSORS:
        ref = getattr(problem, name)
        arr = data["problem." + name]
        if arr.shape != tuple(ref.shape):
            raise ValueError(f"checkpoint {name} has shape {arr.shape}, the problem {tuple(ref.shape)}")
        # fresh tensors: the optimiser's work buffers must not alias them
        setattr(problem, name, torch.from_numpy(arr.copy()).to(device=dev, dtype=ref.dtype))
    problem.iteration = int(data["problem.iteration"])
    problem._f = float(data["problem._f"])
    problem._residual = float(data["problem._residual"])
    solver = getattr(getattr(problem, "FE", None), "solver", None)
    if solver is not None:
        solver.last_x0 = problem._U
    if kind:
        tensors, scalars = _OPTIMIZERS[kind]
        for name in tensors:
            ref = getattr(optimizer, name)
            setattr(optimizer, name, torch.from_numpy(data["optimizer." + name].copy()).to(device=dev,
                                                                                          dtype=ref.dtype))
        for name in scalars:
            v = float(data["optimizer." + name])
            if name == "iteration":
                v = int(v)
            elif name == "last_lambda" and np.isnan(v):
                v = None
            setattr(optimizer, name, v)
