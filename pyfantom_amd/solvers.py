"""Device solvers: geometric-multigrid-preconditioned CG and plain CG.

``MultiGrid`` mirrors the constructor and ``solve(rhs, rho=None,
use_last=True) -> (U, residual)`` / ``reset()`` of the reference
(solvers/CUDA/_solvers.py:330-660).  Every operation of a PCG iteration is a
gfx950 kernel from libfantom_hip.so; scalars (alpha, beta, r.z) stay on the
device, and the only host read per iteration is ||r||^2 for the stopping test.

Deliberate differences from the reference (each keeps the solution K^-1 f):
  * coarse operators are rediscretised and matrix-free: coarse E is the mean
    of the 2^d children and the element matrix is scaled by the edge ratio
    (hex8: K scales with element size; quad4 plane stress: size-free).  The
    reference assembles Galerkin CSR operators (core/CUDA/_mgm.py:19-171);
  * transfers are the plain trilinear pair P_lin^T / P_lin (the reference's
    TW normalisation cancels in the V-cycle);
  * the coarsest level is solved exactly on the device (dense SPD inverse)
    instead of a host CHOLMOD factor of K + 1e-6 I;
  * omega follows the reference's adaptive schedule by default (min_omega
    first, then min_omega + (omega - min_omega)/2 exp(500 (tol - clip(R))),
    times 1.06^level); ``omega_schedule='fixed'`` keeps omega on every level;
  * cycle='W' runs a V-cycle and cycle='w' a W-cycle at ``w_level`` -- this
    is the reference's observed behaviour (it compares with 'w');
  * the returned residual is the true ||f - K u|| / ||f||;
  * ``smoother='chebyshev'`` (native driver) runs the fourth-kind Chebyshev
    smoother that the reference's CPU MultiGrid carries but leaves commented
    out (solvers/CPU/_solvers.py:884-954): n_smooth sweeps with the weights
    of fantom_chebyshev4_weights over rho(D^-1 A) from power iterations at
    each setup (``cheb_power_iters`` = (first, later), warm-started).
"""
from __future__ import annotations

import atexit
import logging
import math
import sys
import warnings
import weakref

import numpy as np
import torch

from . import _lib

logger = logging.getLogger(__name__)

# native solver handles still alive; released before interpreter shutdown so
# that no HIP call runs after the runtime has been torn down
_LIVE = weakref.WeakSet()


@atexit.register
def _release_native_handles():
    for s in list(_LIVE):
        try:
            s._free_native()
        except Exception:
            pass


class _Level:
    """One grid of the hierarchy with its operator data and work vectors."""

    def __init__(self, grid, dim, n_nodes, dof, scale, dev):
        self.grid = grid
        self.dim = dim
        self.n_nodes = n_nodes
        self.dof = dof
        self.n_dof = n_nodes * dof
        self.scale = scale
        self.n_el = grid[0] * grid[1] * (grid[2] if dim == 3 else 1)
        z = lambda: torch.zeros(self.n_dof, dtype=torch.float64, device=dev)
        self.x, self.x2, self.b, self.r = z(), z(), z(), z()
        self.dinv = z()
        self.E = torch.empty(self.n_el, dtype=torch.float64, device=dev)
        self.mask = torch.zeros(n_nodes, dtype=torch.uint8, device=dev)
        self.K0 = None
        self.kid = None


class Solver:
    """Base interface (solvers/commons/__init__.py:1-67)."""

    def solve(self, rhs, rho=None, use_last=True):
        raise NotImplementedError

    def reset(self):
        pass


class _DeviceKrylov(Solver):
    """Shared device buffers and kernel calls of the Krylov solvers."""

    def _alloc(self, n):
        dev = self.kernel.device
        if getattr(self, "_n", None) != n:
            z = lambda: torch.zeros(n, dtype=torch.float64, device=dev)
            self._r, self._z, self._p, self._q = z(), z(), z(), z()
            k = self.kernel
            if hasattr(k, "grid"):
                ku = max(_lib.load().fantom_ku_partials(_lib.KU_TILE if k.dim == 3 else _lib.KU_NODE, *k.grid),
                         _lib.load().fantom_ku_partials(_lib.KU_NODE, *k.grid))
            else:
                ku = k.n_partials()
            self._np = max(_lib.load().fantom_vec_partials(n), ku)
            self._partials = torch.zeros(self._np, dtype=torch.float64, device=dev)
            self._scal = torch.zeros(_lib.N_SCALARS, dtype=torch.float64, device=dev)
            self._n = n

    def _dot(self, a, b, slot=_lib.S_RR, op=_lib.FIN_STORE):
        n = a.numel()
        _lib.call("fantom_dot_partials_f64", n, _lib.ptr(a), _lib.ptr(b), _lib.ptr(self._partials))
        _lib.call("fantom_finalize_f64", _lib.ptr(self._partials), _lib.load().fantom_vec_partials(n),
                  _lib.ptr(self._scal), op, slot)

    def _fine_ku_dot(self, p, q):
        """q = K p and scal[PQ/ALPHA] from p.q."""
        npart = self.kernel.ku_dot(p, q, self._partials)
        _lib.call("fantom_finalize_f64", _lib.ptr(self._partials), npart, _lib.ptr(self._scal), _lib.FIN_ALPHA, 0)

    def _true_residual(self, x, rhs, out):
        self.kernel.ku_residual(x, rhs, out)
        return self._norm(out)

    def _norm(self, v):
        self._dot(v, v, 6)
        return math.sqrt(max(float(self._scal[6].item()), 0.0))


class MultiGrid(_DeviceKrylov):
    def __init__(self, mesh, kernel, maxiter=1000, tol=1e-5, n_smooth=3, omega=0.5, n_level=3, cycle="W",
                 w_level=1, coarse_solver="cholmod", matrix_free=False, low_level_tol=1e-8,
                 low_level_maxiter=5000, min_omega=0.4, omega_boost=1.06, omega_schedule="reference",
                 check_every=1, native=True, use_graph=True, precision="fp64", inner_tol=1e-4,
                 max_outer=20, coarse=None, mixed_scheme="pcg", fallback=True, fallback_maxiter=1000,
                 smoother="jacobi", cheb_rho_scale=1.0, cheb_power_iters=(12, 4), flexible=None):
        if coarse_solver not in ["cg", "gmres", "splu", "spsolve", "cholmod"]:
            raise ValueError("Coarse solver not recognized.")
        if coarse_solver in ["splu", "spsolve"] and matrix_free:
            raise ValueError("Matrix free is not supported with splu solver use cg instead.")
        if omega_schedule not in ("fixed", "reference"):
            raise ValueError("omega_schedule must be 'fixed' or 'reference'")
        if not hasattr(kernel, "grid"):
            raise ValueError("MultiGrid needs a structured mesh and a StructuredStiffnessKernel.")
        self.mesh = mesh
        self.kernel = kernel
        self.maxiter = int(maxiter)
        self.tol = float(tol)
        self.n_smooth = int(n_smooth)
        self.omega = float(omega)
        self.max_omega = float(omega)
        self.min_omega = max(omega / 2, min_omega)
        self.d_omega = self.max_omega - self.min_omega
        self.omega_boost = omega_boost
        self.omega_schedule = omega_schedule
        self.n_level = int(n_level)
        self.cycle = cycle
        self.w_level = [w_level] if isinstance(w_level, int) else list(w_level)
        self.coarse_solver = coarse_solver
        self.matrix_free = matrix_free
        self.low_level_tol = low_level_tol
        self.low_level_maxiter = low_level_maxiter
        self.check_every = max(1, int(check_every))
        self.last_x0 = None
        self.levels = None
        self._built_version = None
        self.last_iterations = 0
        self.last_recursive_residual = None
        # the C++ driver runs V cycles with either weight schedule; the
        # Python driver below also covers the reference's live 'w' cycle
        self.native = bool(native) and cycle != "w"
        self.use_graph = bool(use_graph)
        if precision not in ("fp64", "mixed", "bf16"):
            raise ValueError("precision must be 'fp64', 'mixed' or 'bf16'")
        # 'bf16': the mixed-precision solve with the V-cycle's fine and
        # level-1 sweeps on the matrix cores (BF16 operands, FP32 sums)
        self.mixed = precision in ("mixed", "bf16")
        if self.mixed and not self.native:
            raise ValueError("the mixed-precision solve runs in the native driver (V-cycle)")
        self.precision = precision
        # flexible (Polak-Ribiere) PCG around the mixed-precision V-cycle;
        # default on for the BF16 V-cycle, whose rounding makes the
        # preconditioner vary from one application to the next
        self.flexible = (precision == "bf16") if flexible is None else bool(flexible)
        # operator kind of the BF16 V-cycle (fantom_mx_set_operator): fine
        # and level-1 sweeps on the matrix cores with the node values split
        # into hi + lo BF16 pairs (single BF16 node values stall the PCG)
        self.bf16_operator = 1 | 2 | 4
        if smoother not in ("jacobi", "chebyshev"):
            raise ValueError("smoother must be 'jacobi' or 'chebyshev'")
        if smoother == "chebyshev" and not self.native:
            raise ValueError("the Chebyshev smoother runs in the native driver (V-cycle)")
        self.smoother = smoother
        self.cheb_rho_scale = float(cheb_rho_scale)
        self.cheb_power_iters = tuple(int(v) for v in cheb_power_iters)
        # mixed precision: "pcg" = FP64 PCG preconditioned by an FP32 V-cycle;
        # "ir" = FP64 iterative refinement around an FP32 GMG-PCG
        if mixed_scheme not in ("pcg", "ir"):
            raise ValueError("mixed_scheme must be 'pcg' or 'ir'")
        self.mixed_scheme = mixed_scheme
        if coarse is None:
            coarse = "galerkin" if (self.native and len(mesh.nel) == 3) else "rediscretized"
        if coarse not in ("galerkin", "rediscretized"):
            raise ValueError("coarse must be 'galerkin' or 'rediscretized'")
        if coarse == "galerkin" and not self.native:
            raise ValueError("Galerkin coarse operators run in the native driver")
        self.coarse = coarse
        self.inner_tol = float(inner_tol)
        self.max_outer = int(max_outer)
        # mixed precision: a solve that stagnates or ends above tol is
        # finished by an FP64 GMG-PCG of the same hierarchy from its iterate
        self.fallback = bool(fallback)
        self.fallback_maxiter = int(fallback_maxiter)
        self._fb = None
        self.last_fallback = False
        self.last_fallback_iterations = 0
        self.last_outer_iterations = 0
        self._nh = None
        self._nh_key = None
        self._nh_state = None
        self._ws = None

    def reset(self):
        """Forget the hierarchy and the warm start."""
        self.levels = None
        self._built_version = None
        self.last_x0 = None
        self._free_native()
        if getattr(self, "_fb", None) is not None:
            self._fb.reset()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self._free_native()
        except Exception:
            pass

    def _free_native(self):
        if getattr(self, "_nh", None) is not None:
            if self._nh_key is not None and self._nh_key[-1] in ("mixed", "bf16"):
                _lib.load().fantom_mx_destroy(self._nh)
            else:
                _lib.load().fantom_mg_destroy(self._nh)
            self._nh = None
            self._nh_state = None

    def _n_coarse(self):
        nel = list(self.kernel.grid[:self.kernel.dim])
        n_coarse = self.n_level
        for l in range(1, self.n_level + 1):
            if any(n % (2 ** l) for n in nel):
                n_coarse = l - 1
                break
        if n_coarse < self.n_level:
            warnings.warn(f"grid {nel} halves evenly only {n_coarse} times; using {n_coarse} coarse levels")
        return n_coarse

    def precondition_native(self, r, z):
        """z = V(r): one V-cycle of the native hierarchy (FP64, or FP32 with
        FP64 input and output for precision='mixed') for the kernel's current E
        (sets the hierarchy up if needed)."""
        if not self.native:
            raise NotImplementedError("precondition_native needs the native driver")
        k = self.kernel
        if self._nh is None or self._nh_state != (k.rho_version, k.con_version, k.rho64.data_ptr()):
            self._native_prepare()
        pre = "fantom_mx_" if self.mixed else "fantom_mg_"
        _lib.call(pre + "precondition", self._nh, _lib.ptr(r), _lib.ptr(z))
        return z

    def _native_schedule(self, L, pre, h):
        adaptive = self.omega_schedule == "reference" and self.smoother == "jacobi"
        _lib.check(getattr(L, pre + "set_omega_schedule")(h, int(adaptive), self.min_omega if adaptive else 0.0,
                                                              self.omega_boost if adaptive else 1.0))
        if self.smoother == "chebyshev":
            _lib.check(getattr(L, pre + "set_smoother")(h, 1, self.cheb_rho_scale, *self.cheb_power_iters))

    def smoother_rho(self):
        """Current rho(D^-1 A) estimates of the smoothed levels (Chebyshev), fine first."""
        import ctypes
        if self._nh is None or self.smoother != "chebyshev":
            return None
        n = self._n_levels - 1
        out = (ctypes.c_double * max(n, 1))()
        pre = "fantom_mx_" if self._nh_key[-1] in ("mixed", "bf16") else "fantom_mg_"
        _lib.check(getattr(_lib.load(), pre + "smoother_rho")(self._nh, out, n))
        return list(out)[:n]

    def _native_prepare(self):
        """Create the native handle for the current options (FP64 or mixed)
        and set it up for the kernel's current E; returns (library, prefix)."""
        import ctypes
        k = self.kernel
        L = _lib.load()
        mixed = self.mixed
        if mixed and k.dim != 3:
            raise NotImplementedError("the mixed-precision solve is implemented for hex8 grids")
        gal = int(self.coarse == "galerkin")
        key = (k.grid, self.n_level, self.n_smooth, self.omega, k.kernel_kind, gal, self.omega_schedule,
               self.min_omega, self.omega_boost, self.smoother, self.cheb_rho_scale, self.cheb_power_iters,
               self.flexible, self.bf16_operator, self.precision)
        pre = "fantom_mx_" if mixed else "fantom_mg_"
        if self._nh is None or self._nh_key != key:
            self._free_native()
            n_coarse = self._n_coarse()
            nbytes = getattr(L, pre + "workspace_size")(*k.grid, n_coarse, gal)
            if nbytes <= 0:
                raise ValueError("grid does not support the requested hierarchy")
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=k.device)
            h = ctypes.c_void_p()
            K0 = np.ascontiguousarray(k.K0.cpu().numpy())
            if mixed:
                _lib.check(L.fantom_mx_create(*k.grid, n_coarse, self.n_smooth, self.omega, gal, K0.ctypes.data,
                                              self._ws.data_ptr(), nbytes, ctypes.byref(h)))
            else:
                kid = _lib.KU_TILE if k.kernel_kind == "tile" else _lib.KU_NODE
                _lib.check(L.fantom_mg_create(*k.grid, n_coarse, self.n_smooth, self.omega, kid, gal,
                                              K0.ctypes.data, self._ws.data_ptr(), nbytes, ctypes.byref(h)))
            self._nh, self._nh_key, self._nh_state = h, key, None
            self._n_levels = n_coarse + 1
            _LIVE.add(self)
            self._native_schedule(L, pre, h)
            if mixed:
                _lib.check(L.fantom_mx_set_operator(h, self.bf16_operator if self.precision == "bf16" else 0))
                _lib.check(L.fantom_mx_set_flexible(h, int(self.flexible)))
        state = (k.rho_version, k.con_version, k.rho64.data_ptr())
        if self._nh_state != state:
            _lib.call(pre + "setup", self._nh, _lib.ptr(k.rho64), _lib.ptr(k.mask) if k.has_con else None)
            self._nh_state = state
        return L, pre

    def _native_solve(self, rhs, use_last):
        import ctypes
        k = self.kernel
        mixed = self.mixed
        L, pre = self._native_prepare()
        n = rhs.numel()
        use_x0 = bool(use_last and self.last_x0 is not None and self.last_x0.numel() == n)
        x = self.last_x0 if use_x0 else torch.empty_like(rhs)
        it = ctypes.c_int(0)
        res = ctypes.c_double(0.0)
        if mixed and self.mixed_scheme == "pcg":
            _lib.call("fantom_mx_solve_pcg", self._nh, _lib.ptr(rhs), _lib.ptr(x), self.tol, self.maxiter,
                      int(use_x0), int(self.use_graph), ctypes.byref(it), ctypes.byref(res))
            self.last_outer_iterations = 1
        elif mixed:
            outer = ctypes.c_int(0)
            _lib.call("fantom_mx_solve", self._nh, _lib.ptr(rhs), _lib.ptr(x), self.tol, self.inner_tol,
                      self.max_outer, self.maxiter, int(use_x0), int(self.use_graph), ctypes.byref(outer),
                      ctypes.byref(it), ctypes.byref(res))
            self.last_outer_iterations = outer.value
        else:
            _lib.call("fantom_mg_solve", self._nh, _lib.ptr(rhs), _lib.ptr(x), self.tol, self.maxiter, int(use_x0),
                      int(self.use_graph), ctypes.byref(it), ctypes.byref(res))
        self.last_iterations = it.value
        self.last_x0 = x
        self.last_fallback = False
        self.last_fallback_iterations = 0
        if mixed and self.fallback and not res.value < self.tol:
            if self._fb is None:
                self._fb = MultiGrid(self.mesh, k, maxiter=self.fallback_maxiter, tol=self.tol, n_smooth=self.n_smooth,
                                     omega=self.omega, n_level=self.n_level, cycle=self.cycle,
                                     use_graph=self.use_graph, coarse=self.coarse, smoother=self.smoother,
                                     cheb_rho_scale=self.cheb_rho_scale, cheb_power_iters=self.cheb_power_iters)
            self._fb.tol = self.tol
            self._fb.last_x0 = x
            x, r = self._fb.solve(rhs, use_last=True)
            self.last_fallback = True
            self.last_fallback_iterations = self._fb.last_iterations
            self.last_x0 = x
            return x, r
        return x, res.value

    # ---- hierarchy ------------------------------------------------------------------
    def _make_levels(self):
        k = self.kernel
        dim = k.dim
        dev = k.device
        nel = list(k.grid[:dim])
        n_coarse = self.n_level
        for l in range(1, self.n_level + 1):
            if any(n % (2 ** l) for n in nel):
                n_coarse = l - 1
                break
        if n_coarse < self.n_level:
            warnings.warn(f"grid {nel} halves evenly only {n_coarse} times; using {n_coarse} coarse levels")
        levels = []
        for l in range(n_coarse + 1):
            g = tuple(n >> l for n in nel) + ((0,) if dim == 2 else ())
            n_nodes = int(np.prod([n + 1 for n in g[:dim]]))
            scale = float(2 ** l) if dim == 3 else 1.0
            L = _Level(g, dim, n_nodes, k.dof, scale, dev)
            L.kid, L.K0 = k.op_args(scale)
            L.K0dense = (k.K0 * scale).contiguous()
            levels.append(L)
        c = levels[-1]
        self._Ainv = torch.zeros((c.n_dof, c.n_dof), dtype=torch.float64, device=dev)
        self._gj_work = torch.zeros(c.n_dof * c.n_dof + 1024, dtype=torch.float64, device=dev)
        self.levels = levels

    def _setup(self):
        """Coarse moduli, masks, Jacobi weights and the coarsest inverse for the
        current E (once per set_rho, like the reference's per-solve _setup)."""
        k = self.kernel
        if self.levels is None or self.levels[0].n_dof != k.shape[0]:
            self._make_levels()
        L0 = self.levels[0]
        L0.E = k.rho64
        L0.mask = k.mask if k.has_con else torch.zeros_like(L0.mask)
        for l in range(1, len(self.levels)):
            f, c = self.levels[l - 1], self.levels[l]
            _lib.call("fantom_coarsen_E_f64", *f.grid, _lib.ptr(f.E), _lib.ptr(c.E))
            _lib.call("fantom_coarsen_mask", *f.grid, _lib.ptr(f.mask), _lib.ptr(c.mask))
        for L in self.levels:
            _lib.call("fantom_diag_f64", *L.grid, _lib.ptr(L.K0dense), _lib.ptr(L.E), _lib.ptr(L.mask),
                      _lib.ptr(L.dinv), 1)
        c = self.levels[-1]
        _lib.call("fantom_dense_assemble_f64", *c.grid, _lib.ptr(c.K0dense), _lib.ptr(c.E), _lib.ptr(c.mask),
                  _lib.ptr(self._Ainv))
        _lib.call("fantom_spd_inverse_f64", c.n_dof, _lib.ptr(self._Ainv), _lib.ptr(self._gj_work))
        self._built_version = k.rho_version

    # ---- V-cycle ----------------------------------------------------------------------
    def _smooth(self, L, b, x, n, omega, from_zero):
        """n Jacobi sweeps on L; returns the tensor holding the result."""
        other = L.x2 if x is L.x else L.x
        for s in range(n):
            if s == 0 and from_zero:
                _lib.call("fantom_scale_f64", L.n_dof, omega, _lib.ptr(L.dinv), _lib.ptr(b), _lib.ptr(x))
                continue
            _lib.call("fantom_ku_jacobi_f64", L.kid, *L.grid, StructuredKernelPtr(L.K0), _lib.ptr(L.E),
                      _lib.ptr(L.mask), _lib.ptr(b), _lib.ptr(L.dinv), omega, _lib.ptr(x), _lib.ptr(other))
            x, other = other, x
        return x

    def _cycle(self, l, b, omega):
        """Approximately solve A_l e = b from a zero guess; returns the result tensor."""
        L = self.levels[l]
        if l == len(self.levels) - 1:
            _lib.call("fantom_gemv_f64", L.n_dof, _lib.ptr(self._Ainv), _lib.ptr(b), _lib.ptr(L.x))
            return L.x
        w = omega * (self.omega_boost ** l if self.omega_schedule == "reference" else 1.0)
        x = self._smooth(L, b, L.x, self.n_smooth, w, True)
        n_visits = 2 if (self.cycle == "w" and l in self.w_level) else 1
        C = self.levels[l + 1]
        for visit in range(n_visits):
            _lib.call("fantom_ku_residual_f64", L.kid, *L.grid, StructuredKernelPtr(L.K0), _lib.ptr(L.E),
                      _lib.ptr(L.mask), _lib.ptr(b), _lib.ptr(x), _lib.ptr(L.r))
            _lib.call("fantom_restrict_f64", *L.grid, 0, _lib.ptr(C.mask), _lib.ptr(L.r), _lib.ptr(C.b))
            e = self._cycle(l + 1, C.b, omega)
            _lib.call("fantom_prolong_f64", *L.grid, 0, _lib.ptr(L.mask), _lib.ptr(e), _lib.ptr(x), 1)
            x = self._smooth(L, b, x, self.n_smooth, w, False)
        return x

    def precondition(self, r, z):
        """z = M^-1 r (one cycle)."""
        out = self._cycle(0, r, self._omega_now)
        if out.data_ptr() != z.data_ptr():
            z.copy_(out)
        return z

    # ---- PCG ----------------------------------------------------------------------------
    def solve(self, rhs, rho=None, use_last=True):
        k = self.kernel
        if not (k.has_rho or rho is not None):
            raise ValueError("Solver requires a density vector to be passed or set on the kernel.")
        if rho is not None:
            k.set_rho(rho)
        rhs = k._as_vec(rhs, torch.float64)
        if self.native:
            return self._native_solve(rhs, use_last)
        n = rhs.numel()
        self._alloc(n)
        if self.levels is None or self._built_version != k.rho_version or self.levels[0].n_dof != n:
            self._setup()
        if use_last and self.last_x0 is not None and self.last_x0.numel() == n:
            x = self.last_x0
        else:
            x = torch.zeros_like(rhs)
        r, z, p, q = self._r, self._z, self._p, self._q
        self._omega_now = self.omega if self.omega_schedule == "fixed" else self.min_omega
        norm_b = self._norm(rhs)
        if norm_b == 0.0:
            norm_b = 1.0
        kid, k0 = k.op_args()
        _lib.call("fantom_ku_residual_f64", kid, *k.grid, k.k0_ptr(k0), _lib.ptr(k.rho64),
                  _lib.ptr(k.mask) if k.has_con else None, _lib.ptr(rhs), _lib.ptr(x), _lib.ptr(r))
        self.precondition(r, z)
        self._dot(r, z, 0, _lib.FIN_RZ0)
        _lib.call("fantom_pcg_p_f64", n, _lib.ptr(p), _lib.ptr(z), _lib.ptr(self._scal), 1)
        it = 0
        R = None
        for it in range(1, self.maxiter + 1):
            self._fine_ku_dot(p, q)
            _lib.call("fantom_pcg_xr_f64", n, _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(q),
                      _lib.ptr(self._scal), _lib.ptr(self._partials))
            _lib.call("fantom_finalize_f64", _lib.ptr(self._partials), _lib.load().fantom_vec_partials(n),
                      _lib.ptr(self._scal), _lib.FIN_STORE, _lib.S_RR)
            if it % self.check_every == 0 or it == self.maxiter:
                R = math.sqrt(max(float(self._scal[_lib.S_RR].item()), 0.0)) / norm_b
                if R < self.tol:
                    break
                if self.omega_schedule == "reference":
                    self._omega_now = self.min_omega + self.d_omega / 2 * math.exp(
                        (-min(max(R, self.tol), 1e-1) + self.tol) * 500)
            self.precondition(r, z)
            self._dot(r, z, 0, _lib.FIN_BETA)
            _lib.call("fantom_pcg_p_f64", n, _lib.ptr(p), _lib.ptr(z), _lib.ptr(self._scal), 0)
        self.last_iterations = it
        self.last_recursive_residual = R
        residual = self._true_residual(x, rhs, q) / norm_b
        self.last_x0 = x
        return x, residual


def StructuredKernelPtr(k0):
    if isinstance(k0, np.ndarray):
        return k0.ctypes.data
    return k0.data_ptr()


class CG(_DeviceKrylov):
    """Unpreconditioned conjugate gradients on the device
    (solvers/CUDA/_solvers.py:22-132; the reference calls cupyx cg)."""

    def __init__(self, kernel, maxiter=1000, tol=1e-5):
        self.kernel = kernel
        self.maxiter = int(maxiter)
        self.tol = float(tol)
        self.last_x0 = None
        self.last_iterations = 0

    def reset(self):
        self.last_x0 = None

    def solve(self, rhs, rho=None, use_last=True):
        k = self.kernel
        if not (k.has_rho or rho is not None):
            raise ValueError("Solver requires a density vector to be passed or set on the kernel.")
        if rho is not None:
            k.set_rho(rho)
        rhs = k._as_vec(rhs, torch.float64)
        n = rhs.numel()
        self._alloc(n)
        x = self.last_x0 if (use_last and self.last_x0 is not None and self.last_x0.numel() == n) \
            else torch.zeros_like(rhs)
        r, p, q = self._r, self._p, self._q
        norm_b = self._norm(rhs) or 1.0
        k.ku_residual(x, rhs, r)
        self._dot(r, r, 0, _lib.FIN_RZ0)
        _lib.call("fantom_pcg_p_f64", n, _lib.ptr(p), _lib.ptr(r), _lib.ptr(self._scal), 1)
        it = 0
        for it in range(1, self.maxiter + 1):
            self._fine_ku_dot(p, q)
            _lib.call("fantom_pcg_xr_f64", n, _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(q),
                      _lib.ptr(self._scal), _lib.ptr(self._partials))
            _lib.call("fantom_finalize_f64", _lib.ptr(self._partials), _lib.load().fantom_vec_partials(n),
                      _lib.ptr(self._scal), _lib.FIN_BETA, 0)
            R = math.sqrt(max(float(self._scal[0].item()), 0.0)) / norm_b
            if R < self.tol:
                break
            _lib.call("fantom_pcg_p_f64", n, _lib.ptr(p), _lib.ptr(r), _lib.ptr(self._scal), 0)
        self.last_iterations = it
        residual = self._true_residual(x, rhs, q) / norm_b
        self.last_x0 = x
        return x, residual
