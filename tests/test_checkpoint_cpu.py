"""CPU checks of the checkpoint file logic (pyfantom_amd.checkpoint) with
stand-in objects holding CPU tensors: round trip of every field, fresh
tensors on load, plain-array files (no pickles), and the mismatch errors."""
import types

import numpy as np
import pytest
import torch

from pyfantom_amd import checkpoint


class MMA:   # the checkpoint keys optimiser state by class name
    pass


class OC:
    pass


def _problem(n=10, ndof=33, seed=0):
    g = torch.Generator().manual_seed(seed)
    r = lambda m: torch.rand(m, generator=g, dtype=torch.float64)
    p = types.SimpleNamespace(desvars=r(n), _nabla_f=r(n), _U=r(ndof), _rho=r(n), iteration=7, _f=1.25,
                              _residual=3e-9)
    p.FE = types.SimpleNamespace(solver=types.SimpleNamespace(last_x0=None))
    return p


def _mma(n=10, seed=1):
    g = torch.Generator().manual_seed(seed)
    o = MMA()
    for name in ("last_desvars", "x_1", "x_2", "low", "upp"):
        setattr(o, name, torch.rand(n, generator=g, dtype=torch.float64))
    o.last_f, o.change, o.change_f, o.iteration, o.comp_base = 1.5, 0.25, 1e-3, 4, 9.0
    return o


def test_round_trip_mma(tmp_path):
    p, o = _problem(), _mma()
    path = tmp_path / "ck.npz"
    checkpoint.save(path, p, o)
    with np.load(path, allow_pickle=False) as z:     # plain arrays only
        assert "optimizer.low" in z.files
    q, o2 = _problem(seed=5), _mma(seed=6)
    q_desvars_before = q.desvars
    checkpoint.load(path, q, o2)
    for name in ("desvars", "_nabla_f", "_U", "_rho"):
        assert torch.equal(getattr(q, name), getattr(p, name))
    assert q.desvars is not q_desvars_before
    assert (q.iteration, q._f, q._residual) == (7, 1.25, 3e-9)
    assert q.FE.solver.last_x0 is q._U
    for name in ("last_desvars", "x_1", "x_2", "low", "upp"):
        assert torch.equal(getattr(o2, name), getattr(o, name))
    assert (o2.last_f, o2.change, o2.change_f, o2.iteration, o2.comp_base) == (1.5, 0.25, 1e-3, 4, 9.0)
    assert isinstance(o2.iteration, int)


def test_round_trip_oc_without_lambda(tmp_path):
    p = _problem()
    o = OC()
    o.last_desvars = p.desvars.clone()
    o.last_f, o.change, o.change_f, o.iteration, o.last_lambda = 2.0, 0.1, 0.01, 3, None
    checkpoint.save(tmp_path / "oc.npz", p, o)
    q, o2 = _problem(seed=2), OC()
    o2.last_desvars = torch.zeros(10, dtype=torch.float64)
    checkpoint.load(tmp_path / "oc.npz", q, o2)
    assert o2.last_lambda is None and o2.iteration == 3
    assert torch.equal(o2.last_desvars, p.desvars)


def test_mismatches_raise(tmp_path):
    p, o = _problem(), _mma()
    checkpoint.save(tmp_path / "ck.npz", p, o)
    with pytest.raises(ValueError):
        checkpoint.load(tmp_path / "ck.npz", _problem(n=12), _mma(n=12))
    oc = OC()
    oc.last_desvars = torch.zeros(10, dtype=torch.float64)
    with pytest.raises(ValueError):
        checkpoint.load(tmp_path / "ck.npz", _problem(), oc)
    np.savez(tmp_path / "other.npz", a=np.zeros(3))
    with pytest.raises(ValueError):
        checkpoint.load(tmp_path / "other.npz", _problem())
    with pytest.raises(ValueError):
        checkpoint.save(tmp_path / "x.npz", p, object())
