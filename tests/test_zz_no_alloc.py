"""Allocation-free OC update (fantom_oc_update_noalloc_f64) against the
allocating entry point (fantom_oc_update_f64) through the C ABI: the same
x_new bit for bit, the same Lagrange multiplier and the same change, on two
calls in a row (the second on a smaller design), plus the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 1 << 22


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _inputs(torch, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, n)
    df = -rng.uniform(1e-3, 2.0, n)
    df[rng.integers(0, n, max(1, n // 50))] = 0.5  # a few positive sensitivities (ocP = 0 there)
    dg = np.full(n, 1.0 / n)
    dev = lambda v: torch.from_numpy(v).cuda()
    return dev(x), dev(df), dev(dg)


def _allocating(torch, L, x, df, dg, vf, move):
    from pyfantom_amd import _lib
    n = x.numel()
    ocp = torch.empty(n, dtype=torch.float64, device="cuda")
    state = torch.zeros(16, dtype=torch.float64, device="cuda")
    partials = torch.zeros(16 * 1024, dtype=torch.float64, device="cuda")
    xn = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.call("fantom_oc_update_f64", n, _lib.ptr(x), _lib.ptr(df), _lib.ptr(dg), vf, move, _lib.ptr(ocp),
              _lib.ptr(state), _lib.ptr(partials), _lib.ptr(xn))
    _lib.call("fantom_finalize_f64", _lib.ptr(partials), L.fantom_oc_partials(n), _lib.ptr(state),
              _lib.FIN_STORE, 8)
    st = state.cpu().numpy()
    return xn, st[2], st[8]


def _noalloc(torch, x, df, dg, vf, move):
    from pyfantom_amd import _lib
    n = x.numel()
    xn = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    res = torch.full((2,), np.nan, dtype=torch.float64, device="cuda")
    _lib.call("fantom_oc_update_noalloc_f64", n, _lib.ptr(x), _lib.ptr(df), _lib.ptr(dg), vf, move,
              _lib.ptr(xn), _lib.ptr(res))
    r = res.cpu().numpy()
    return xn, r[0], r[1]


def test_oc_noalloc_matches_allocating_twice():
    from pyfantom_amd import _lib
    torch = _torch()
    L = _lib.load()
    for n, seed, vf, move in [(300_000, 0, 0.45, 0.2), (20_000, 1, 0.5, 0.1)]:
        print(f"oc noalloc n={n} seed={seed}", flush=True)
        x, df, dg = _inputs(torch, n, seed)
        xa, lam_a, ch_a = _allocating(torch, L, x, df, dg, vf, move)
        xb, lam_b, ch_b = _noalloc(torch, x, df, dg, vf, move)
        torch.cuda.synchronize()
        assert torch.equal(xa, xb)
        assert lam_a == lam_b and np.isfinite(lam_a)
        assert ch_a == ch_b and ch_a > 0
        # the update honours the move limit and the volume target
        assert float((xb - x).abs().max()) <= move + 1e-15
        assert abs(float((xb * dg).sum()) - vf) < 1e-5


def test_oc_noalloc_rejects_oversize():
    from pyfantom_amd import _lib
    torch = _torch()
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="scratch capacity"):
        # rejected by the size check before any pointer is used
        _lib.call("fantom_oc_update_noalloc_f64", MAX_N + 1, _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), 0.3, 0.2,
                  _lib.ptr(x) + 8, _lib.ptr(x))
