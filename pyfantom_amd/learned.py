"""Learned preconditioner: a small 3D convolutional encoder/decoder on the node
grid, trained on the GPU with the project's own BF16 MFMA kernels, used inside
flexible PCG with plain GMG-PCG as fallback.

No reference counterpart: pyFANTOM has no learned component.  Everything here
runs through libfantom_hip.so (include/fantom.h, A20 entry points).

Network (node grid (D, H, W) = (ny+1, nx+1, nz+1), channels-last BF16):
  input  16 ch: r/rms(r) (3), nodal density (1), zeros (12)
  e0 = relu(conv(x, 16->16))                level 0
  e1 = relu(conv(R e0, 16->32))             level 1   R: full weighting / 8
  e2 = relu(conv(R e1, 32->64))             level 2
  e3 = relu(conv(R e2, 64->64))             level 3
  d2 = relu(conv(P e3, 64->64) + e2)        P: trilinear interpolation
  d1 = relu(conv(P d2, 64->32) + e1)
  d0 = relu(conv(P d1, 32->16) + e0)
  y  = conv(d0, 16->16), channels 0..2 -> z = rms(r) * y
The map r -> z is positively homogeneous of degree one, so flexible PCG sees
a scale-free preconditioner.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

LAYERS = [  # name, c_in, c_out, level, relu
    ("e0", 16, 16, 0, True),
    ("e1", 16, 32, 1, True),
    ("e2", 32, 64, 2, True),
    ("e3", 64, 64, 3, True),
    ("d2", 64, 64, 2, True),
    ("d1", 64, 32, 1, True),
    ("d0", 32, 16, 0, True),
    ("out", 16, 16, 0, False),
]


def _bf16(t):
    return t.to(torch.bfloat16).view(torch.int16)


class ConvNet:
    """Parameters (FP32 master + BF16 copy) and Adam state of the network."""

    def __init__(self, seed=0, device=None):
        dev = device or _lib.device()
        self.device = dev
        g = torch.Generator().manual_seed(seed)
        self.w, self.b, self.wbf, self.m_w, self.v_w, self.m_b, self.v_b = {}, {}, {}, {}, {}, {}, {}
        for name, ci, co, lev, relu in LAYERS:
            fan_in = 27 * (4 if name == "e0" else ci)
            std = math.sqrt(2.0 / fan_in) if relu else 0.1 * math.sqrt(1.0 / fan_in)
            w = torch.randn(27, co, ci, generator=g) * std
            if name == "e0":
                w[:, :, 4:] = 0.0
            if name == "out":
                w[:, 3:, :] = 0.0
            self.w[name] = w.to(dev).contiguous()
            self.b[name] = torch.zeros(co, device=dev)
            self.wbf[name] = _bf16(self.w[name]).contiguous()
            for d in (self.m_w, self.v_w):
                d[name] = torch.zeros_like(self.w[name])
            for d in (self.m_b, self.v_b):
                d[name] = torch.zeros_like(self.b[name])
        self.step = 0

    def state_dict(self):
        return {**{f"w_{k}": v.cpu() for k, v in self.w.items()}, **{f"b_{k}": v.cpu() for k, v in self.b.items()}}

    def load_state_dict(self, sd):
        for name, *_ in LAYERS:
            self.w[name] = sd[f"w_{name}"].to(self.device).contiguous()
            self.b[name] = sd[f"b_{name}"].to(self.device).contiguous()
            self.wbf[name] = _bf16(self.w[name]).contiguous()

    def save(self, path, **extra):
        """Weights and biases as an .npz archive (numpy arrays only)."""
        np.savez(path, **{k: v.numpy() for k, v in self.state_dict().items()}, **extra)

    @classmethod
    def from_file(cls, path, device=None):
        """Network with the weights of an archive written by save()."""
        net = cls(device=device)
        with np.load(path, allow_pickle=False) as z:
            net.load_state_dict({k: torch.from_numpy(z[k].astype(np.float32)) for k in z.files
                                 if k.startswith(("w_", "b_"))})
        return net

    def n_params(self):
        return sum(v.numel() for v in self.w.values()) + sum(v.numel() for v in self.b.values())


class NetRunner:
    """Activation buffers for one grid and the forward / backward passes."""

    def __init__(self, net: ConvNet, grid):
        nelx, nely, nelz = grid
        if nelx % 8 or nely % 8 or nelz % 8:
            raise ValueError("the learned preconditioner needs element counts divisible by 8")
        self.net = net
        self.grid = grid
        dev = net.device
        self.shape = [((nely >> l) + 1, (nelx >> l) + 1, (nelz >> l) + 1) for l in range(4)]
        self.nv = [int(np.prod(s)) for s in self.shape]
        z16 = lambda l, c: torch.zeros(self.nv[l] * c, dtype=torch.int16, device=dev)
        self.act = {"x": z16(0, 16)}
        for name, ci, co, lev, relu in LAYERS:
            self.act[name] = z16(lev, co)
        # inputs of the convolutions after a transfer
        self.act["r1"], self.act["r2"], self.act["r3"] = z16(1, 16), z16(2, 32), z16(3, 64)
        self.act["u2"], self.act["u1"], self.act["u0"] = z16(2, 64), z16(1, 64), z16(0, 32)
        self.y32 = torch.zeros(self.nv[0] * 16, dtype=torch.float32, device=dev)
        self.rho_node = torch.zeros(self.nv[0], dtype=torch.float32, device=dev)
        self.scal = torch.zeros(16, dtype=torch.float64, device=dev)
        self.partials = torch.zeros(4096, dtype=torch.float64, device=dev)
        self._grad_bufs = None

    # --- helpers ---------------------------------------------------------------------------
    def _conv(self, name, x, lev, out, skip=None, relu=True, out_f32=None):
        _, ci, co, _, _ = next(l for l in LAYERS if l[0] == name)
        _lib.call("fantom_conv3d_fwd_bf16", *self.shape[lev], ci, co, _lib.ptr(x), _lib.ptr(self.net.wbf[name]),
                  _lib.ptr(self.net.b[name]), _lib.ptr(skip), _lib.ptr(out), _lib.ptr(out_f32), int(relu))

    def _restrict(self, lev, c, x, out, scale):
        _lib.call("fantom_vol_restrict_bf16", *self.shape[lev], c, _lib.ptr(x), _lib.ptr(out), float(scale))

    def _prolong(self, lev, c, xc, out, scale, accumulate):
        _lib.call("fantom_vol_prolong_bf16", *self.shape[lev], c, _lib.ptr(xc), _lib.ptr(out), float(scale),
                  int(accumulate))

    def set_density(self, x_el):
        """Nodal density channel from an element field (design variables)."""
        _lib.call("fantom_node_density", *self.grid, _lib.ptr(x_el), _lib.ptr(self.rho_node))

    def _norm2(self, r, slot):
        n = r.numel()
        _lib.call("fantom_dot_partials_f64", n, _lib.ptr(r), _lib.ptr(r), _lib.ptr(self.partials))
        _lib.call("fantom_finalize_f64", _lib.ptr(self.partials), _lib.load().fantom_vec_partials(n),
                  _lib.ptr(self.scal), _lib.FIN_STORE, slot)

    # --- forward ------------------------------------------------------------------------------
    def forward(self, r):
        """Network output y32 (level-0 volume, 16 ch FP32) for residual r (FP64 nodal vector)."""
        a = self.act
        self._norm2(r, 0)
        _lib.call("fantom_net_input", self.nv[0], _lib.ptr(r), _lib.ptr(self.rho_node), _lib.ptr(self.scal), 0,
                  _lib.ptr(a["x"]))
        self._conv("e0", a["x"], 0, a["e0"])
        self._restrict(0, 16, a["e0"], a["r1"], 0.125)
        self._conv("e1", a["r1"], 1, a["e1"])
        self._restrict(1, 32, a["e1"], a["r2"], 0.125)
        self._conv("e2", a["r2"], 2, a["e2"])
        self._restrict(2, 64, a["e2"], a["r3"], 0.125)
        self._conv("e3", a["r3"], 3, a["e3"])
        self._prolong(2, 64, a["e3"], a["u2"], 1.0, 0)
        self._conv("d2", a["u2"], 2, a["d2"], skip=a["e2"])
        self._prolong(1, 64, a["d2"], a["u1"], 1.0, 0)
        self._conv("d1", a["u1"], 1, a["d1"], skip=a["e1"])
        self._prolong(0, 32, a["d1"], a["u0"], 1.0, 0)
        self._conv("d0", a["u0"], 0, a["d0"], skip=a["e0"])
        self._conv("out", a["d0"], 0, a["out"], relu=False, out_f32=self.y32)
        return self.y32

    def apply(self, r, z, accumulate=False):
        """z (+)= rms(r) * N(r / rms(r))."""
        self.forward(r)
        _lib.call("fantom_net_output", self.nv[0], _lib.ptr(self.y32), _lib.ptr(self.scal), 0, _lib.ptr(z),
                  int(accumulate))
        return z

    # --- backward -----------------------------------------------------------------------------
    def _bufs(self):
        if self._grad_bufs is None:
            dev = self.net.device
            g = {}
            for lev in range(4):
                for c in (16, 32, 64):
                    g[(lev, c, 0)] = torch.zeros(self.nv[lev] * c, dtype=torch.int16, device=dev)
                    g[(lev, c, 1)] = torch.zeros(self.nv[lev] * c, dtype=torch.int16, device=dev)
                    g[(lev, c, 2)] = torch.zeros(self.nv[lev] * c, dtype=torch.int16, device=dev)
            nparts = max(_lib.load().fantom_conv3d_wgrad_parts(*self.shape[lv], ci, co, None)
                         for (_, ci, co, lv, _) in LAYERS)
            g["wpart"] = torch.zeros(nparts * 27 * 64 * 64, dtype=torch.float32, device=dev)
            g["bpart"] = torch.zeros(256 * 64, dtype=torch.float32, device=dev)
            g["wt"] = torch.zeros(27 * 64 * 64, dtype=torch.int16, device=dev)
            self.gw = {n: torch.zeros_like(self.net.w[n]) for n in self.net.w}
            self.gb = {n: torch.zeros_like(self.net.b[n]) for n in self.net.b}
            self._grad_bufs = g
        return self._grad_bufs

    def _layer_grads(self, name, x_in, g_out):
        """dW, db for layer `name` from its input and the gradient at its output
        (after the activation mask); returns the input gradient buffer filled."""
        _, ci, co, lev, _ = next(l for l in LAYERS if l[0] == name)
        B = self._bufs()
        _lib.call("fantom_conv3d_wgrad_bf16", *self.shape[lev], ci, co, _lib.ptr(x_in), _lib.ptr(g_out),
                  _lib.ptr(B["wpart"]), _lib.ptr(self.gw[name]))
        _lib.call("fantom_bias_grad_bf16", self.nv[lev], co, _lib.ptr(g_out), _lib.ptr(B["bpart"]),
                  _lib.ptr(self.gb[name]))

    def _input_grad(self, name, g_out, dx):
        _, ci, co, lev, _ = next(l for l in LAYERS if l[0] == name)
        B = self._bufs()
        _lib.call("fantom_flip_transpose_bf16", co, ci, _lib.ptr(self.net.wbf[name]), _lib.ptr(B["wt"]))
        _lib.call("fantom_conv3d_fwd_bf16", *self.shape[lev], co, ci, _lib.ptr(g_out), _lib.ptr(B["wt"]), None, None,
                  _lib.ptr(dx), None, 0)

    def _mask(self, y, dy, g):
        _lib.call("fantom_relu_mask_bf16", y.numel(), _lib.ptr(y), _lib.ptr(dy), _lib.ptr(g))

    def backward(self, dout):
        """Gradients of all parameters for the output gradient dout (level-0, 16 ch BF16)."""
        a, B = self.act, self._bufs()
        # out layer (linear)
        self._layer_grads("out", a["d0"], dout)
        dd0 = B[(0, 16, 0)]
        self._input_grad("out", dout, dd0)
        # d0 = relu(conv(u0) + e0)
        g0 = B[(0, 16, 1)]
        self._mask(a["d0"], dd0, g0)
        self._layer_grads("d0", a["u0"], g0)
        de0 = B[(0, 16, 2)]
        de0.copy_(g0)
        du0 = B[(0, 32, 0)]
        self._input_grad("d0", g0, du0)
        dd1 = B[(1, 32, 0)]
        self._restrict(0, 32, du0, dd1, 1.0)          # P^T
        # d1 = relu(conv(u1) + e1)
        g1 = B[(1, 32, 1)]
        self._mask(a["d1"], dd1, g1)
        self._layer_grads("d1", a["u1"], g1)
        de1 = B[(1, 32, 2)]
        de1.copy_(g1)
        du1 = B[(1, 64, 0)]
        self._input_grad("d1", g1, du1)
        dd2 = B[(2, 64, 0)]
        self._restrict(1, 64, du1, dd2, 1.0)
        # d2 = relu(conv(u2) + e2)
        g2 = B[(2, 64, 1)]
        self._mask(a["d2"], dd2, g2)
        self._layer_grads("d2", a["u2"], g2)
        de2 = B[(2, 64, 2)]
        de2.copy_(g2)
        du2 = B[(2, 64, 0)]
        self._input_grad("d2", g2, du2)
        de3 = B[(3, 64, 0)]
        self._restrict(2, 64, du2, de3, 1.0)
        # e3 = relu(conv(r3)),  r3 = R e2 (R = restriction / 8, R^T = P / 8)
        g3 = B[(3, 64, 1)]
        self._mask(a["e3"], de3, g3)
        self._layer_grads("e3", a["r3"], g3)
        dr3 = B[(3, 64, 2)]
        self._input_grad("e3", g3, dr3)
        self._prolong(2, 64, dr3, de2, 0.125, 1)
        # e2
        gg2 = B[(2, 64, 1)]
        self._mask(a["e2"], de2, gg2)
        self._layer_grads("e2", a["r2"], gg2)
        dr2 = B[(2, 32, 0)]
        self._input_grad("e2", gg2, dr2)
        self._prolong(1, 32, dr2, de1, 0.125, 1)
        # e1
        gg1 = B[(1, 32, 1)]
        self._mask(a["e1"], de1, gg1)
        self._layer_grads("e1", a["r1"], gg1)
        dr1 = B[(1, 16, 0)]
        self._input_grad("e1", gg1, dr1)
        self._prolong(0, 16, dr1, de0, 0.125, 1)
        # e0
        gg0 = B[(0, 16, 1)]
        self._mask(a["e0"], de0, gg0)
        self._layer_grads("e0", a["x"], gg0)
        return self.gw, self.gb

    def loss_and_grad(self, r, e):
        """MSE between N(r / rms(r)) and e / rms(e) (first 3 channels); fills
        the gradients.  Both sides are normalised: flexible PCG is invariant to
        the scale of z, and the unnormalised error is dominated by the
        near-void regions where it is orders of magnitude larger."""
        B = self._bufs()
        self.forward(r)
        self._norm2(e, 2)
        dout = B[(0, 16, 0)]
        _lib.call("fantom_mse_bf16", self.nv[0], _lib.ptr(self.y32), _lib.ptr(e), _lib.ptr(self.scal), 2,
                  _lib.ptr(dout), _lib.ptr(self.partials))
        dcopy = B[(0, 16, 2)]
        dcopy.copy_(dout)
        self.backward(dcopy)
        _lib.call("fantom_finalize_f64", _lib.ptr(self.partials), 256, _lib.ptr(self.scal), _lib.FIN_STORE, 1)
        return self.scal[1]

    def energy_loss_and_grad(self, t, e, vec, free):
        """Energy loss of the hybrid correction, L = 1 - (t.c)^2 / ((c.Kc)(t.e))
        with c = N(t / rms(t)) (free DOFs) and t = K e: the fraction of the
        error energy e.Ke that the Galerkin step g c (g = t.c / c.Kc) leaves.
        Scale-free in c; dL/dc = -2a/(b s) t + 2a^2/(b^2 s) Kc with a = t.c,
        b = c.Kc, s = t.e.  Fills the gradients, returns L (host float)."""
        B = self._bufs()
        y = self.forward(t).view(self.nv[0], 16)
        c = y[:, :3].reshape(-1).double() * free
        kc = torch.empty_like(c)
        vec.ku(c, kc)
        L, g = energy_loss_grad(t, e, c, kc)
        if not math.isfinite(L):
            return L
        g = g * free
        g = g / max(float(g.abs().max()), 1e-300)        # Adam is invariant to the gradient's scale
        dout = B[(0, 16, 2)]   # backward() uses (0, 16, 0) and (0, 16, 1) as work buffers
        dv = dout.view(self.nv[0], 16)
        dv.zero_()
        dv[:, :3] = g.view(-1, 3).to(torch.bfloat16).view(torch.int16)
        self.backward(dout)
        return L

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        net = self.net
        net.step += 1
        for name in net.w:
            if name == "e0":
                self.gw[name].view(27, 16, 16)[:, :, 4:] = 0.0
            if name == "out":
                self.gw[name].view(27, 16, 16)[:, 3:, :] = 0.0
                self.gb[name][3:] = 0.0
            _lib.call("fantom_adam_f32", net.w[name].numel(), _lib.ptr(net.w[name]), _lib.ptr(self.gw[name]),
                      _lib.ptr(net.m_w[name]), _lib.ptr(net.v_w[name]), _lib.ptr(net.wbf[name]), lr, b1, b2, eps,
                      net.step)
            _lib.call("fantom_adam_f32", net.b[name].numel(), _lib.ptr(net.b[name]), _lib.ptr(self.gb[name]),
                      _lib.ptr(net.m_b[name]), _lib.ptr(net.v_b[name]), None, lr, b1, b2, eps, net.step)


def energy_loss_grad(t, e, c, kc):
    """L = 1 - a^2 / (b s) with a = t.c, b = c.Kc (kc = K c), s = t.e, and
    dL/dc = -2a/(b s) t + 2a^2/(b^2 s) Kc (K symmetric).  For t = K e this is
    the fraction of e's energy left after the best multiple of c is removed.
    Returns (L, dL/dc); L is NaN when b or s is not positive."""
    a, b, sv = float(torch.dot(t, c)), float(torch.dot(c, kc)), float(torch.dot(t, e))
    if not (b > 0.0 and sv > 0.0 and math.isfinite(a)):
        return float("nan"), None
    g = (-2.0 * a / (b * sv)) * t + (2.0 * a * a / (b * b * sv)) * kc
    return 1.0 - a * a / (b * sv), g


class _Vec:
    """FP64 device vector helpers on the fantom kernels (dot, axpby, K u).

    The scalars of the PCG recursion live in the device array scal (slots
    _lib.S_*, plus S_A / S_B / S_C for the two-dot reads); the methods that
    end in a finalize leave their result there, and read() fetches several
    slots with one host synchronisation."""

    S_A, S_B, S_C = 8, 9, 10

    def __init__(self, kernel):
        self.k = kernel
        n = kernel.shape[0]
        dev = kernel.device
        L = _lib.load()
        self.nvp = L.fantom_vec_partials(n)
        kid, _ = kernel.op_args()
        self.nkp = L.fantom_ku_partials(kid, *kernel.grid)
        self.partials = torch.zeros(max(2 * self.nvp, self.nkp, 4096), dtype=torch.float64, device=dev)
        self.scal = torch.zeros(16, dtype=torch.float64, device=dev)

    def read(self, *slots):
        """Host values of the given scal slots (one synchronisation)."""
        v = self.scal.cpu()
        return [float(v[s]) for s in slots]

    def ku_dot(self, u, y, op=0, slot=8):
        """y = K u and the finalize op on u.y (default: scal[slot] = u.y)."""
        k = self.k
        kid, k0 = k.op_args()
        _lib.call("fantom_ku_dot_f64", kid, *k.grid, k.k0_ptr(k0), _lib.ptr(k.rho64),
                  _lib.ptr(k.mask) if k.has_con else None, _lib.ptr(u), _lib.ptr(y), _lib.ptr(self.partials))
        _lib.call("fantom_finalize_f64", _lib.ptr(self.partials), self.nkp, _lib.ptr(self.scal), op, slot)

    def dot_to(self, a, b, op=0, slot=8):
        """finalize op on a.b (default: scal[slot] = a.b), no host read."""
        n = a.numel()
        _lib.call("fantom_dot_partials_f64", n, _lib.ptr(a), _lib.ptr(b), _lib.ptr(self.partials))
        _lib.call("fantom_finalize_f64", _lib.ptr(self.partials), self.nvp, _lib.ptr(self.scal), op, slot)

    def dot2_to(self, a, b, c, op_b, slot_b, op_c, slot_c):
        """a.b and a.c in one pass, each through its finalize op."""
        n = a.numel()
        _lib.call("fantom_dot2_partials_f64", n, _lib.ptr(a), _lib.ptr(b), _lib.ptr(c), _lib.ptr(self.partials))
        P = self.partials.data_ptr()
        _lib.call("fantom_finalize_f64", P, self.nvp, _lib.ptr(self.scal), op_b, slot_b)
        _lib.call("fantom_finalize_f64", P + 8 * self.nvp, self.nvp, _lib.ptr(self.scal), op_c, slot_c)

    def xr(self, x, r, p, q):
        """x += alpha p, r -= alpha q with alpha = scal[S_ALPHA]; scal[S_RR] = r.r."""
        n = x.numel()
        _lib.call("fantom_pcg_xr_f64", n, _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(q), _lib.ptr(self.scal),
                  _lib.ptr(self.partials))
        _lib.call("fantom_finalize_f64", _lib.ptr(self.partials), self.nvp, _lib.ptr(self.scal), _lib.FIN_STORE,
                  _lib.S_RR)

    def dot(self, a, b):
        n = a.numel()
        _lib.call("fantom_dot_partials_f64", n, _lib.ptr(a), _lib.ptr(b), _lib.ptr(self.partials))
        _lib.call("fantom_finalize_f64", _lib.ptr(self.partials), _lib.load().fantom_vec_partials(n),
                  _lib.ptr(self.scal), _lib.FIN_STORE, 7)
        return float(self.scal[7].item())

    def axpby(self, a, x, b, y):
        _lib.call("fantom_axpby_f64", y.numel(), float(a), _lib.ptr(x), float(b), _lib.ptr(y))

    def ku(self, u, y):
        k = self.k
        kid, k0 = k.op_args()
        _lib.call("fantom_ku_f64", kid, *k.grid, k.k0_ptr(k0), _lib.ptr(k.rho64),
                  _lib.ptr(k.mask) if k.has_con else None, _lib.ptr(u), _lib.ptr(y))

    def residual(self, f, x, r):
        k = self.k
        kid, k0 = k.op_args()
        _lib.call("fantom_ku_residual_f64", kid, *k.grid, k.k0_ptr(k0), _lib.ptr(k.rho64),
                  _lib.ptr(k.mask) if k.has_con else None, _lib.ptr(f), _lib.ptr(x), _lib.ptr(r))


class LearnedPreconditioner:
    """z = M(r) for flexible PCG.

    mode 'learned': z = N(r), the network alone (replaces smoother and coarse
    correction).  mode 'hybrid': z = V(r) + g c with c = N(t), t = r - K V(r),
    one GMG V-cycle followed by a learned correction of what it leaves, scaled
    by the Galerkin step g = t.c / c.Kc (the network predicts directions, not
    magnitudes).  Constrained DOFs of z are zeroed (the network sees only the
    free part of the residual)."""

    def __init__(self, net: ConvNet, kernel, mg=None, mode="learned"):
        if mode not in ("learned", "hybrid"):
            raise ValueError("mode must be 'learned' or 'hybrid'")
        if mode == "hybrid" and mg is None:
            raise ValueError("the hybrid mode needs a MultiGrid solver for its V-cycle")
        self.runner = NetRunner(net, kernel.grid)
        self.kernel = kernel
        self.mg = mg
        self.mode = mode
        self.vec = _Vec(kernel)
        n = kernel.shape[0]
        self._t = torch.zeros(n, dtype=torch.float64, device=kernel.device)
        self._t2 = torch.zeros(n, dtype=torch.float64, device=kernel.device)
        self._c = torch.zeros(n, dtype=torch.float64, device=kernel.device)
        self._free = (~kernel.constraints).to(torch.float64)

    def set_density(self, x_el):
        self.runner.set_density(x_el.contiguous())

    def __call__(self, r, z):
        if self.mode == "learned":
            self.runner.apply(r, z)
            z.mul_(self._free)
        else:
            # z = V(r); t2 = r - K z; c = N(t2) on the free DOFs; z += (t2.c / c.Kc) c.
            # V(r) is zero on the constrained DOFs (identity rows, zero residual
            # there) and so is c, hence z needs no masking.
            v = self.vec
            self.mg.precondition_native(r, z)
            v.residual(r, z, self._t2)
            self.runner.apply(self._t2, self._c)
            self._c.mul_(self._free)
            v.ku_dot(self._c, self._t, slot=v.S_A)
            v.dot_to(self._t2, self._c, slot=v.S_B)
            ckc, t2c = v.read(v.S_A, v.S_B)
            if math.isfinite(ckc) and ckc > 0.0 and math.isfinite(t2c):
                v.axpby(t2c / ckc, self._c, 1.0, z)
        return z


def fpcg_solve(kernel, f, precond, tol=1e-8, maxiter=500, x0=None, fallback=None, stall_window=20,
               stall_factor=0.9, record=None):
    """Flexible PCG (Polak-Ribiere beta) with a nonlinear preconditioner.

    fallback: a MultiGrid solver; if the residual does not drop below
    stall_factor times its value stall_window iterations earlier, or the
    preconditioned direction breaks down (r.z or p.Kp not positive and
    finite), the solve continues with plain GMG-PCG from the current iterate.
    Without a fallback a breakdown restarts the recurrence with z = r.
    record (optional): called as record(x) after every iterate update.
    Returns (x, info) with info = {iterations, fallback, residual, history}."""
    vec = precond.vec if hasattr(precond, "vec") else _Vec(kernel)
    n = f.numel()
    dev = f.device
    x = torch.zeros_like(f) if x0 is None else x0.clone()
    r = torch.empty_like(f)
    vec.residual(f, x, r)
    nb = math.sqrt(vec.dot(f, f)) or 1.0
    z = torch.zeros_like(f)
    z_old = torch.zeros_like(f)
    p = torch.zeros_like(f)
    q = torch.zeros_like(f)
    def ok(v):
        return math.isfinite(v) and v > 0.0

    hist = [math.sqrt(vec.dot(r, r)) / nb]
    info = {"fallback": False, "fallback_iterations": 0, "restarts": 0}

    def handover(x):
        info["fallback"] = True
        fallback.tol = tol
        xf, res = _fallback_from(fallback, f, x)
        info["fallback_iterations"] = fallback.last_iterations
        hist.append(res)
        return xf

    def restart():
        z.copy_(r)
        p.copy_(z)
        info["restarts"] += 1
        return vec.dot(r, z)

    def set_rz(v):
        # the device copy of r.z that FIN_ALPHA divides by
        vec.scal[_lib.S_RZ] = v

    precond(r, z)
    rz = vec.dot(r, z)
    set_rz(rz)
    p.copy_(z)
    it = 0
    if not ok(rz) and hist[-1] > 0.0:
        if fallback is not None:
            x = handover(x)
            maxiter = 0
        else:
            rz = restart()
            set_rz(rz)
    for it in range(1, maxiter + 1):
        # q = K p with p.q, then x += alpha p, r -= alpha q (alpha = 0 on a
        # breakdown) and r.r, all on the device; one host read
        vec.ku_dot(p, q, op=_lib.FIN_ALPHA)
        vec.xr(x, r, p, q)
        pq, rr = vec.read(_lib.S_PQ, _lib.S_RR)
        breakdown = not ok(pq)
        if breakdown:
            hist.append(hist[-1])   # one history entry per iteration keeps the stall test aligned
        else:
            if record is not None:
                record(x)
            R = math.sqrt(max(rr, 0.0)) / nb
            hist.append(R)
            if R < tol:
                break
        stalled = it >= stall_window and hist[-1] > stall_factor * hist[-1 - stall_window]
        if fallback is not None and (breakdown or stalled):
            x = handover(x)
            break
        if breakdown:
            rz = restart()
            set_rz(rz)
            continue
        z, z_old = z_old, z   # z_old: the previous preconditioned residual
        precond(r, z)
        vec.dot2_to(r, z, z_old, _lib.FIN_STORE, vec.S_A, _lib.FIN_STORE, vec.S_B)
        rz_new, rz_old = vec.read(vec.S_A, vec.S_B)
        if not ok(rz_new):
            if fallback is not None:
                x = handover(x)
                break
            rz = restart()
            set_rz(rz)
            continue
        beta = (rz_new - rz_old) / rz
        rz = rz_new
        set_rz(rz)
        vec.axpby(1.0, z, beta, p)
    vec.residual(f, x, r)
    info.update(iterations=it, residual=math.sqrt(vec.dot(r, r)) / nb, history=hist)
    return x, info


def _fallback_from(mg, f, x):
    """Continue with GMG-PCG from the iterate x."""
    mg.last_x0 = x.clone()
    return mg.solve(f, use_last=True)


def density_from_moduli(E, Emin=1e-9, E0=1.0, penalty=3.0):
    """Physical (filtered) element density from the moduli of the modified
    SIMP law E = Emin + rho^p (E0 - Emin): the network's density channel,
    so that the solver needs only what FiniteElement.solve hands it."""
    return ((E - Emin) / (E0 - Emin)).clamp(min=0.0) ** (1.0 / penalty)


class LearnedSolver:
    """Flexible PCG preconditioned by one GMG V-cycle (FP64, or FP32 for
    precision='mixed') plus the learned correction of what the cycle leaves
    (LearnedPreconditioner, mode 'hybrid'), with plain FP64 GMG-PCG as the
    fallback on stagnation or breakdown.  Same protocol as MultiGrid
    (solvers/commons/__init__.py:1-67): solve(rhs, rho=None, use_last=True)
    -> (U, true relative residual), reset(), last_iterations, last_x0,
    last_fallback, last_fallback_iterations.  No counterpart in the reference."""

    def __init__(self, mesh, kernel, net: ConvNet, maxiter=500, tol=1e-8, precision="mixed", n_level=4, n_smooth=2,
                 omega=0.6, smoother="jacobi", omega_schedule="fixed", Emin=1e-9, E0=1.0, penalty=3.0,
                 stall_window=20, fallback_maxiter=1000):
        from .solvers import MultiGrid
        self.kernel = kernel
        self.maxiter = int(maxiter)
        self.tol = float(tol)
        self.Emin, self.E0, self.penalty = Emin, E0, penalty
        self.stall_window = int(stall_window)
        opts = dict(n_smooth=n_smooth, omega=omega, n_level=n_level, cycle="V", omega_schedule=omega_schedule,
                    smoother=smoother)
        self.mg = MultiGrid(mesh, kernel, maxiter=1, tol=tol, precision=precision, fallback=False, **opts)
        self.fb = MultiGrid(mesh, kernel, maxiter=fallback_maxiter, tol=tol, precision="fp64", **opts)
        self.M = LearnedPreconditioner(net, kernel, mg=self.mg, mode="hybrid")
        self._dens_state = None
        self.last_x0 = None
        self.last_iterations = 0
        self.last_fallback = False
        self.last_fallback_iterations = 0
        self.last_info = None

    def reset(self):
        self.last_x0 = None
        self._dens_state = None
        self.mg.reset()
        self.fb.reset()

    def solve(self, rhs, rho=None, use_last=True):
        k = self.kernel
        if rho is not None:
            k.set_rho(rho)
        state = (k.rho_version, k.rho64.data_ptr())
        if self._dens_state != state:
            self.M.set_density(density_from_moduli(k.rho64, self.Emin, self.E0, self.penalty))
            self._dens_state = state
        x0 = self.last_x0 if (use_last and self.last_x0 is not None and self.last_x0.numel() == rhs.numel()) else None
        x, info = fpcg_solve(k, rhs, self.M, tol=self.tol, maxiter=self.maxiter, x0=x0, fallback=self.fb,
                             stall_window=self.stall_window)
        self.last_iterations = info["iterations"]
        self.last_fallback = info["fallback"]
        self.last_fallback_iterations = info["fallback_iterations"]
        self.last_info = info
        self.last_x0 = x
        return x, info["residual"]
