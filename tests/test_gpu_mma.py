"""Device MMA update (fantom_mma_update_f64 through the C ABI) against the
numpy MMA oracle on seeded data: first iterations (fixed asymptotes) and
later ones (moving asymptotes), one and two constraints."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


# n > 1024 blocks x 256 threads = 262,144 puts several elements on every
# thread (the grid-stride regime of the 14M-variable bridge); m = 3 and m = 8
# cover the larger block-reduction records of the dual system
@pytest.mark.parametrize("n,m,it,seed", [(1000, 1, 1, 0), (5000, 1, 4, 1), (3000, 2, 3, 2), (200000, 1, 5, 3),
                                         (600000, 1, 3, 4), (600000, 3, 4, 5), (400000, 8, 2, 6)])
def test_mma_update_vs_oracle(n, m, it, seed):
    from oracle import mma_oracle as M
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"mma n={n} m={m} iter={it}", flush=True)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, n)
    xold1 = np.clip(x + rng.normal(0, 0.02, n), 0, 1)
    xold2 = np.clip(xold1 + rng.normal(0, 0.02, n), 0, 1)
    low = x - rng.uniform(0.1, 0.5, n)
    upp = x + rng.uniform(0.1, 0.5, n)
    df0 = -rng.uniform(0.0, 2.0, n) + 0.1 * rng.standard_normal(n)
    dfdx = rng.uniform(0.5, 1.5, (m, n)) / n
    fval = rng.uniform(-0.02, 0.02, m)
    xmin, xmax = np.zeros(n), np.ones(n)
    a, c, d = np.zeros(m), np.full(m, 1e5), np.zeros(m)
    xr, lowr, uppr, extra = M.mmasub(it, x, xmin, xmax, xold1, xold2, df0, fval, dfdx, low.copy(), upp.copy(), a0=1.0,
                                 a=a, c=c, d=d, move=0.5)
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    xd, x1d, x2d, dfd, dgd = dev(x), dev(xold1), dev(xold2), dev(df0), dev(dfdx)
    lowd, uppd = dev(low), dev(upp)
    xn = torch.empty_like(xd)
    nbytes = _lib.load().fantom_mma_workspace_size(n, m)
    work = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device="cuda")
    arr = lambda v: (ctypes.c_double * m)(*[float(t) for t in v])
    steps = ctypes.c_int(0)
    _lib.call("fantom_mma_update_f64", n, m, it, xd.data_ptr(), x1d.data_ptr(), x2d.data_ptr(), 0.0, 1.0,
              dfd.data_ptr(), arr(fval), dgd.data_ptr(), lowd.data_ptr(), uppd.data_ptr(), 0.5, 1.0, arr(a), arr(c),
              arr(d), 1e-7, 100, xn.data_ptr(), work.data_ptr(), work.numel() * 8, ctypes.byref(steps))
    torch.cuda.synchronize()
    assert steps.value > 0
    assert np.abs(lowd.cpu().numpy() - lowr).max() <= 1e-14
    assert np.abs(uppd.cpu().numpy() - uppr).max() <= 1e-14
    err = np.abs(xn.cpu().numpy() - xr).max()
    ref_steps = extra[-1][-1]
    print(f"  newton steps {steps.value} (oracle {ref_steps}), max |x - x_oracle| = {err:.2e}", flush=True)
    assert err <= 1e-8
    # the device control flow takes the oracle's path: the same number of
    # Newton steps up to a borderline stopping test
    assert abs(steps.value - ref_steps) <= 2


def test_mma_device_control_repeatable():
    """Two updates on the same data give bit-identical points and step
    counts although their first status reads come after different numbers
    of queued units (the first batch is sized from the previous update)."""
    from pyfantom_amd import _lib
    torch = _torch()
    n, m = 300000, 2
    print(f"mma repeatable n={n} m={m}", flush=True)
    rng = np.random.default_rng(11)
    x = rng.uniform(0.05, 0.95, n)
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    xd, dfd = dev(x), dev(-rng.uniform(0.0, 2.0, n))
    dgd = dev(rng.uniform(0.5, 1.5, (m, n)) / n)
    fval = (ctypes.c_double * m)(0.01, -0.01)
    zeros = (ctypes.c_double * m)(0.0, 0.0)
    cc = (ctypes.c_double * m)(1e5, 1e5)
    nbytes = _lib.load().fantom_mma_workspace_size(n, m)
    out = []
    for trial in range(3):
        lowd, uppd = dev(x - 0.3), dev(x + 0.3)
        xn = torch.empty_like(xd)
        work = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device="cuda")
        steps = ctypes.c_int(0)
        _lib.call("fantom_mma_update_f64", n, m, 1, xd.data_ptr(), None, None, 0.0, 1.0, dfd.data_ptr(), fval,
                  dgd.data_ptr(), lowd.data_ptr(), uppd.data_ptr(), 0.2, 1.0, zeros, cc, zeros, 1e-7, 100,
                  xn.data_ptr(), work.data_ptr(), work.numel() * 8, ctypes.byref(steps))
        torch.cuda.synchronize()
        out.append((xn.cpu(), steps.value))
        if trial == 0:
            # a tiny update in between resets the first-batch size
            t = torch.full((16,), 0.5, dtype=torch.float64, device="cuda")
            w2 = torch.empty(_lib.load().fantom_mma_workspace_size(16, 1) // 8 + 1, dtype=torch.float64,
                             device="cuda")
            lo, up, xx = t - 0.5, t + 0.5, torch.empty_like(t)
            one = (ctypes.c_double * 1)(1e5)
            z1 = (ctypes.c_double * 1)(0.0)
            f1 = (ctypes.c_double * 1)(0.0)
            # operands held in names: a temporary freed right after data_ptr()
            # can hand its block to the next allocation before the launch
            g0, g1 = (-t).contiguous(), (t / 16).contiguous()
            _lib.call("fantom_mma_update_f64", 16, 1, 1, t.data_ptr(), None, None, 0.0, 1.0, g0.data_ptr(), f1,
                      g1.data_ptr(), lo.data_ptr(), up.data_ptr(), 0.2, 1.0, z1, one, z1, 1e-7, 100, xx.data_ptr(),
                      w2.data_ptr(), w2.numel() * 8, None)
            torch.cuda.synchronize()
    print(f"  steps {[o[1] for o in out]}", flush=True)
    for xo, so in out[1:]:
        assert so == out[0][1]
        assert torch.equal(xo, out[0][0])


def test_mma_rejects_bad_sizes():
    from pyfantom_amd import _lib
    _torch()
    assert _lib.load().fantom_mma_workspace_size(10, 9) == -1
    assert _lib.load().fantom_mma_workspace_size(0, 1) == -1


def test_mma_mbb_pipeline_vs_oracle(oracle):
    """MBB 120 x 40 with pf.MMA: each step's update against the oracle's
    mmasub fed the same design, sensitivities and asymptote state (the
    problem's compliance and sensitivities are pinned by the OC pipeline
    test)."""
    from oracle import mma_oracle as M
    import pyfantom_amd as pf
    torch = _torch()
    print("mma mbb pipeline", flush=True)
    nelx, nely = 120, 40
    mesh = pf.StructuredMesh2D(nelx, nely, nelx, nely, physics=pf.LinearElasticity(E=1.0, nu=0.3))
    K = pf.StructuredStiffnessKernel(mesh)
    solver = pf.MultiGrid(mesh, K, maxiter=2000, tol=1e-12, n_smooth=2, omega=0.6, n_level=3, cycle="V")
    FE = pf.FiniteElement(mesh, K, solver)
    nx = nelx + 1
    FE.add_dirichlet_boundary_condition(node_ids=np.arange(nely + 1) * nx, dofs=np.array([[1, 0]]), rhs=0)
    FE.add_dirichlet_boundary_condition(node_ids=np.array([nelx]), dofs=np.array([[0, 1]]), rhs=0)
    FE.add_point_forces(forces=np.array([[0.0, -1.0]]), node_ids=np.array([nely * nx]))
    filt = pf.StructuredFilter2D(mesh, 1.5)
    prob = pf.MinimumCompliance(FE, filt, E_mul=[1.0], void=1e-9, penalty=3.0, volume_fraction=[0.5],
                                heavyside=False, law="modified")
    opt = pf.MMA(prob, move=0.2)
    n = prob.N()
    for it in range(1, 6):
        x = prob.get_desvars().cpu().numpy()
        df = prob.nabla_f().cpu().numpy()
        dg = prob.nabla_g().cpu().numpy().reshape(1, n)
        g = prob.g().cpu().numpy().reshape(1)
        low, upp = opt.low.cpu().numpy(), opt.upp.cpu().numpy()
        x1, x2 = opt.x_1.cpu().numpy(), opt.x_2.cpu().numpy()
        xr, _, _, _ = M.mmasub(it, x, np.zeros(n), np.ones(n), x1, x2, df, g, dg, low, upp, a0=1.0, move=0.2)
        opt.iter()
        xn = prob.get_desvars().cpu().numpy()
        err = np.abs(xn - xr).max()
        print(f"  step {it}: c = {prob.f():.6e}, max |x - x_oracle| = {err:.2e}, newton steps "
              f"{opt.last_newton_steps}", flush=True)
        assert err <= 1e-7
