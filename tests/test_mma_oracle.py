"""Pins of the MMA oracle (oracle/mma_oracle.py): Svanberg's toy problem run
to convergence agrees with an independent solver (scipy SLSQP), the returned
subproblem solutions satisfy the subproblem's KKT conditions, and the
asymptote update follows its rules on hand-made sequences."""
import numpy as np
import pytest

from oracle import mma_oracle as M


def _toy(x):
    f0 = x @ x
    df0 = 2 * x
    f = np.array([(x[0] - 5) ** 2 + (x[1] - 2) ** 2 + (x[2] - 1) ** 2 - 9,
                  (x[0] - 3) ** 2 + (x[1] - 4) ** 2 + (x[2] - 3) ** 2 - 9])
    df = np.array([[2 * (x[0] - 5), 2 * (x[1] - 2), 2 * (x[2] - 1)],
                   [2 * (x[0] - 3), 2 * (x[1] - 4), 2 * (x[2] - 3)]])
    return f0, df0, f, df


def test_toy_problem_matches_slsqp():
    """min |x|^2 s.t. two spheres of radius 3, 0 <= x <= 5 (Svanberg's toy)."""
    from scipy.optimize import minimize
    x = np.array([4.0, 3.0, 2.0])
    xmin, xmax = np.zeros(3), np.full(3, 5.0)
    xold1 = x.copy()
    xold2 = x.copy()
    low, upp = xmin.copy(), xmax.copy()
    for it in range(1, 40):
        f0, df0, f, df = _toy(x)
        xnew, low, upp, _ = M.mmasub(it, x, xmin, xmax, xold1, xold2, df0, f, df, low, upp, a0=1.0,
                                     a=np.zeros(2), c=np.full(2, 1000.0), d=np.ones(2), move=1.0)
        xold2, xold1, x = xold1, x, xnew
    cons = [{"type": "ineq", "fun": lambda x, k=k: -_toy(x)[2][k]} for k in range(2)]
    ref = minimize(lambda x: x @ x, np.array([4.0, 3.0, 2.0]), jac=lambda x: 2 * x, bounds=[(0, 5)] * 3,
                   constraints=cons, method="SLSQP", options={"ftol": 1e-12, "maxiter": 500})
    assert np.abs(_toy(ref.x)[2]).max() < 1e-6          # both spheres active at the optimum
    assert np.abs(x - ref.x).max() < 1e-5
    assert _toy(x)[2].max() < 1e-6


@pytest.mark.parametrize("n,m,seed", [(50, 1, 0), (400, 1, 1), (300, 2, 2)])
def test_subproblem_kkt(n, m, seed):
    """The interior-point solution satisfies the KKT conditions of the convex
    subproblem it solves: stationarity of the Lagrangian in x (by finite
    differences of the subproblem objective), complementarity, feasibility."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.1, 0.9, n)
    xmin, xmax = np.zeros(n), np.ones(n)
    df0 = rng.standard_normal(n)
    dfdx = rng.uniform(0.5, 1.5, (m, n)) / n
    fval = rng.uniform(-0.05, 0.05, m)
    low, upp = M.asymptotes(1, x, xmin, xmax, x, x, None, None)
    alfa, beta, p0, q0, P, Q, b = M.approximations(x, xmin, xmax, low, upp, df0, fval, dfdx, move=0.5)
    a, c, d = np.zeros(m), np.full(m, 1e5), np.zeros(m)
    xs, y, z, lam, xsi, eta, mu, zet, s, steps = M.subsolv(low, upp, alfa, beta, p0, q0, P, Q, 1.0, a, b, c, d)
    assert steps > 0
    # stationarity: d/dx [sum p0/(u-x) + q0/(x-l) + lam.(P/(u-x) + Q/(x-l))] - xsi + eta = 0
    plam = p0 + P.T @ lam
    qlam = q0 + Q.T @ lam
    grad = plam / (upp - xs) ** 2 - qlam / (xs - low) ** 2
    assert np.abs(grad - xsi + eta).max() < 1e-6 * max(1.0, np.abs(grad).max())
    assert np.all(xs >= alfa - 1e-12) and np.all(xs <= beta + 1e-12)
    g = P @ (1 / (upp - xs)) + Q @ (1 / (xs - low)) - b
    assert np.all(g - y <= 1e-6)                          # feasibility (z = a.z = 0)
    assert np.abs(lam * s).max() < 1e-6 and np.abs(xsi * (xs - alfa)).max() < 1e-6


def test_asymptote_rules():
    x = np.array([0.5, 0.5, 0.5, 0.5])
    xmin, xmax = np.zeros(4), np.ones(4)
    low, upp = M.asymptotes(1, x, xmin, xmax, x, x, None, None)
    assert np.allclose(low, 0.0) and np.allclose(upp, 1.0)
    # oscillation shrinks, monotone motion widens, stagnation keeps the distance
    xold1 = np.array([0.6, 0.4, 0.5, 0.45])
    xold2 = np.array([0.5, 0.3, 0.5, 0.4])
    x = np.array([0.5, 0.5, 0.5, 0.5])
    lo0, up0 = np.full(4, 0.2), np.full(4, 0.9)
    low, upp = M.asymptotes(3, x, xmin, xmax, xold1, xold2, lo0, up0)
    sgn = (x - xold1) * (xold1 - xold2)
    fac = np.where(sgn > 0, 1.2, np.where(sgn < 0, 0.7, 1.0))
    exp_low = np.clip(x - fac * (xold1 - lo0), x - 10.0, x - 0.01)
    exp_upp = np.clip(x + fac * (up0 - xold1), x + 0.01, x + 10.0)
    assert np.allclose(low, exp_low) and np.allclose(upp, exp_upp)
    assert sgn[0] < 0 and sgn[3] > 0 and sgn[2] == 0
