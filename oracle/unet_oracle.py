"""Float64 numpy restatement of the learned preconditioner's network
(pyfantom_amd/learned.py) with its hand-written backward pass.

TEST INFRASTRUCTURE ONLY.  ``rounding=True`` rounds every stored activation
and activation gradient to BF16 at the same points as the GPU kernels, so the
GPU gradients can be compared tightly; ``rounding=False`` is exact float64 and
is pinned by finite differences in tests/test_conv_oracle.py.
"""
import numpy as np

from . import conv_oracle as C

LAYERS = [("e0", 16, 16, 0, True), ("e1", 16, 32, 1, True), ("e2", 32, 64, 2, True), ("e3", 64, 64, 3, True),
          ("d2", 64, 64, 2, True), ("d1", 64, 32, 1, True), ("d0", 32, 16, 0, True), ("out", 16, 16, 0, False)]


def _r(a, rounding):
    return C.bf16_round(a) if rounding else a


def forward(x, W, B, rounding=True):
    """x: level-0 input volume (D, H, W, 16).  Returns (y, cache)."""
    a = {"x": x}

    def conv(name, inp, skip=None, relu=True):
        v = C.conv3d(inp, W[name], B[name])
        if skip is not None:
            v = v + skip
        if relu:
            v = np.maximum(v, 0)
        return v

    a["e0"] = _r(conv("e0", x), rounding)
    a["r1"] = _r(C.restrict_vol(a["e0"], 0.125), rounding)
    a["e1"] = _r(conv("e1", a["r1"]), rounding)
    a["r2"] = _r(C.restrict_vol(a["e1"], 0.125), rounding)
    a["e2"] = _r(conv("e2", a["r2"]), rounding)
    a["r3"] = _r(C.restrict_vol(a["e2"], 0.125), rounding)
    a["e3"] = _r(conv("e3", a["r3"]), rounding)
    a["u2"] = _r(C.prolong_vol(a["e3"], a["e2"].shape[:3]), rounding)
    a["d2"] = _r(conv("d2", a["u2"], a["e2"]), rounding)
    a["u1"] = _r(C.prolong_vol(a["d2"], a["e1"].shape[:3]), rounding)
    a["d1"] = _r(conv("d1", a["u1"], a["e1"]), rounding)
    a["u0"] = _r(C.prolong_vol(a["d1"], a["e0"].shape[:3]), rounding)
    a["d0"] = _r(conv("d0", a["u0"], a["e0"]), rounding)
    y = conv("out", a["d0"], relu=False)
    return y, a


def backward(dy, W, a, rounding=True):
    """Parameter gradients for the output gradient dy (level 0, 16 ch)."""
    gW, gB = {}, {}

    def layer(name, inp, g):
        gW[name] = C.conv3d_bwd_weight(inp, g)
        gB[name] = g.reshape(-1, g.shape[-1]).sum(0)
        return _r(C.conv3d_bwd_data(g, W[name]), rounding)

    dy = _r(dy, rounding)
    dd0 = layer("out", a["d0"], dy)
    g0 = dd0 * (a["d0"] > 0)
    du0 = layer("d0", a["u0"], g0)
    de0 = g0.copy()
    dd1 = _r(C.restrict_vol(du0, 1.0), rounding)
    g1 = dd1 * (a["d1"] > 0)
    du1 = layer("d1", a["u1"], g1)
    de1 = g1.copy()
    dd2 = _r(C.restrict_vol(du1, 1.0), rounding)
    g2 = dd2 * (a["d2"] > 0)
    du2 = layer("d2", a["u2"], g2)
    de2 = g2.copy()
    de3 = _r(C.restrict_vol(du2, 1.0), rounding)
    g3 = de3 * (a["e3"] > 0)
    dr3 = layer("e3", a["r3"], g3)
    de2 = _r(de2 + C.prolong_vol(dr3, de2.shape[:3], 0.125), rounding)
    gg2 = de2 * (a["e2"] > 0)
    dr2 = layer("e2", a["r2"], gg2)
    de1 = _r(de1 + C.prolong_vol(dr2, de1.shape[:3], 0.125), rounding)
    gg1 = de1 * (a["e1"] > 0)
    dr1 = layer("e1", a["r1"], gg1)
    de0 = _r(de0 + C.prolong_vol(dr1, de0.shape[:3], 0.125), rounding)
    gg0 = de0 * (a["e0"] > 0)
    layer("e0", a["x"], gg0)
    return gW, gB
