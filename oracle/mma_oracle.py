"""Float64 numpy restatement of the Method of Moving Asymptotes (Svanberg
1987; the subproblem solved by a primal-dual interior-point Newton method,
Svanberg's "subsolv"), as pyFANTOM uses it: mma/CPU/_mma.py:32-165 (mmasub),
:288-528 (subsolv), :530-590 (kktcheck), driven by Optimizers/CPU/MMA.py:103-180
with m constraints, a0 = 1, a = 0, c = 1e5, d = 0, f0val = 0.

TEST INFRASTRUCTURE ONLY.  Column vectors of the reference become 1D arrays;
the m x n constraint gradient is a 2D array.  The m < n branch of the Newton
system (the one a topology problem takes) is restated; the m >= n branch is
used only by tiny toy problems and is restated too.

Pins (tests/test_mma_oracle.py): the KKT residual of the returned subproblem
solution, Svanberg's toy problem run to its known optimum, and the asymptote
update rules on hand-made sequences.  The exact iterate sequence of the
reference is "parity unpinned" (it would need the reference itself to run).
"""
import numpy as np


def asymptotes(it, x, xmin, xmax, xold1, xold2, low, upp, asyinit=0.5, asydecr=0.7, asyincr=1.2, asymin=0.01,
               asymax=10.0):
    """Moving asymptotes (mma/CPU/_mma.py:100-118)."""
    span = xmax - xmin
    if it <= 2:
        return x - asyinit * span, x + asyinit * span
    sgn = (x - xold1) * (xold1 - xold2)
    fac = np.where(sgn > 0, asyincr, np.where(sgn < 0, asydecr, 1.0))
    low = x - fac * (xold1 - low)
    upp = x + fac * (upp - xold1)
    low = np.minimum(np.maximum(low, x - asymax * span), x - asymin * span)
    upp = np.maximum(np.minimum(upp, x + asymax * span), x + asymin * span)
    return low, upp


def approximations(x, xmin, xmax, low, upp, df0dx, fval, dfdx, move=0.5, albefa=0.1, raa0=1e-5):
    """Bounds alfa/beta and the convex approximation p0, q0, P, Q, b
    (mma/CPU/_mma.py:119-158)."""
    alfa = np.maximum(np.maximum(low + albefa * (x - low), x - move * (xmax - xmin)), xmin)
    beta = np.minimum(np.minimum(upp - albefa * (upp - x), x + move * (xmax - xmin)), xmax)
    xmami_inv = 1.0 / np.maximum(xmax - xmin, 1e-5)
    ux = upp - x
    xl = x - low
    p0 = np.maximum(df0dx, 0.0)
    q0 = np.maximum(-df0dx, 0.0)
    pq0 = 0.001 * (p0 + q0) + raa0 * xmami_inv
    p0 = (p0 + pq0) * ux * ux
    q0 = (q0 + pq0) * xl * xl
    P = np.maximum(dfdx, 0.0)
    Q = np.maximum(-dfdx, 0.0)
    PQ = 0.001 * (P + Q) + raa0 * xmami_inv[None, :]
    P = (P + PQ) * (ux * ux)[None, :]
    Q = (Q + PQ) * (xl * xl)[None, :]
    b = P @ (1.0 / ux) + Q @ (1.0 / xl) - fval
    return alfa, beta, p0, q0, P, Q, b


def _residual(x, y, z, lam, xsi, eta, mu, zet, s, epsi, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d):
    ux = upp - x
    xl = x - low
    plam = p0 + P.T @ lam
    qlam = q0 + Q.T @ lam
    gvec = P @ (1.0 / ux) + Q @ (1.0 / xl)
    rex = plam / ux ** 2 - qlam / xl ** 2 - xsi + eta
    rey = c + d * y - mu - lam
    rez = a0 - zet - a @ lam
    relam = gvec - a * z - y + s - b
    rexsi = xsi * (x - alfa) - epsi
    reeta = eta * (beta - x) - epsi
    remu = mu * y - epsi
    rezet = zet * z - epsi
    res = lam * s - epsi
    r = np.concatenate([rex, rey, [rez], relam, rexsi, reeta, remu, [rezet], res])
    return np.sqrt(r @ r), np.abs(r).max()


def subsolv(low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d, epsimin=1e-7, maxiter=100):
    """Primal-dual Newton solve of the MMA subproblem (mma/CPU/_mma.py:288-528).
    Returns (x, y, z, lam, xsi, eta, mu, zet, s, newton_steps)."""
    m, n = P.shape
    epsi = 1.0
    x = 0.5 * (alfa + beta)
    y = np.ones(m)
    z = 1.0
    lam = np.ones(m)
    xsi = np.maximum(1.0 / (x - alfa), 1.0)
    eta = np.maximum(1.0 / (beta - x), 1.0)
    mu = np.maximum(np.ones(m), 0.5 * c)
    zet = 1.0
    s = np.ones(m)
    steps = 0
    outer = 0
    while epsi > epsimin and outer < maxiter:
        outer += 1
        args = (epsi, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
        rnorm, rmax = _residual(x, y, z, lam, xsi, eta, mu, zet, s, *args)
        inner = 0
        while rmax > 0.9 * epsi and inner < 200:
            inner += 1
            steps += 1
            ux = upp - x
            xl = x - low
            plam = p0 + P.T @ lam
            qlam = q0 + Q.T @ lam
            gvec = P @ (1.0 / ux) + Q @ (1.0 / xl)
            GG = P / (ux * ux)[None, :] - Q / (xl * xl)[None, :]
            delx = plam / ux ** 2 - qlam / xl ** 2 - epsi / (x - alfa) + epsi / (beta - x)
            dely = c + d * y - lam - epsi / y
            delz = a0 - a @ lam - epsi / z
            dellam = gvec - a * z - y - b + epsi / lam
            diagx = 2.0 * (plam / ux ** 3 + qlam / xl ** 3) + xsi / (x - alfa) + eta / (beta - x)
            diagy = d + mu / y
            diaglamyi = s / lam + 1.0 / diagy
            if m < n:
                blam = dellam + dely / diagy - GG @ (delx / diagx)
                Alam = np.diag(diaglamyi) + (GG / diagx[None, :]) @ GG.T
                AA = np.zeros((m + 1, m + 1))
                AA[:m, :m] = Alam
                AA[:m, m] = a
                AA[m, :m] = a
                AA[m, m] = -zet / z
                sol = np.linalg.solve(AA, np.concatenate([blam, [delz]]))
                dlam = sol[:m]
                dz = sol[m]
                dx = -delx / diagx - (GG.T @ dlam) / diagx
            else:
                dellamyi = dellam + dely / diagy
                Axx = np.diag(diagx) + (GG.T / diaglamyi[None, :]) @ GG
                azz = zet / z + a @ (a / diaglamyi)
                axz = -GG.T @ (a / diaglamyi)
                bx = delx + GG.T @ (dellamyi / diaglamyi)
                bz = delz - a @ (dellamyi / diaglamyi)
                AA = np.zeros((n + 1, n + 1))
                AA[:n, :n] = Axx
                AA[:n, n] = axz
                AA[n, :n] = axz
                AA[n, n] = azz
                sol = np.linalg.solve(AA, -np.concatenate([bx, [bz]]))
                dx = sol[:n]
                dz = sol[n]
                dlam = GG @ dx / diaglamyi - dz * (a / diaglamyi) + dellamyi / diaglamyi
            dy = -dely / diagy + dlam / diagy
            dxsi = -xsi + epsi / (x - alfa) - xsi * dx / (x - alfa)
            deta = -eta + epsi / (beta - x) + eta * dx / (beta - x)
            dmu = -mu + epsi / y - mu * dy / y
            dzet = -zet + epsi / z - zet * dz / z
            ds = -s + epsi / lam - s * dlam / lam
            xx = np.concatenate([y, [z], lam, xsi, eta, mu, [zet], s])
            dxx = np.concatenate([dy, [dz], dlam, dxsi, deta, dmu, [dzet], ds])
            stm = max(np.max(-1.01 * dxx / xx), np.max(-1.01 * dx / (x - alfa)), np.max(1.01 * dx / (beta - x)),
                      1.0)
            steg = 1.0 / stm
            old = (x, y, z, lam, xsi, eta, mu, zet, s)
            resinew = 2.0 * rnorm
            itto = 0
            while resinew > rnorm and itto < 50:
                itto += 1
                x = old[0] + steg * dx
                y = old[1] + steg * dy
                z = old[2] + steg * dz
                lam = old[3] + steg * dlam
                xsi = old[4] + steg * dxsi
                eta = old[5] + steg * deta
                mu = old[6] + steg * dmu
                zet = old[7] + steg * dzet
                s = old[8] + steg * ds
                resinew, rmax = _residual(x, y, z, lam, xsi, eta, mu, zet, s, *args)
                steg = steg / 2.0
            rnorm = resinew
        epsi = 0.1 * epsi
    return x, y, z, lam, xsi, eta, mu, zet, s, steps


def mmasub(it, x, xmin, xmax, xold1, xold2, df0dx, fval, dfdx, low, upp, a0=1.0, a=None, c=None, d=None, move=0.5,
           sub_tol=1e-7, sub_maxiter=100, **asy):
    """One MMA update.  Returns (x_new, low, upp, subproblem tuple)."""
    m = np.atleast_1d(fval).shape[0]
    a = np.zeros(m) if a is None else a
    c = np.full(m, 1e5) if c is None else c
    d = np.zeros(m) if d is None else d
    dfdx = np.atleast_2d(dfdx)
    low, upp = asymptotes(it, x, xmin, xmax, xold1, xold2, low, upp, **asy)
    alfa, beta, p0, q0, P, Q, b = approximations(x, xmin, xmax, low, upp, df0dx, np.atleast_1d(fval), dfdx, move)
    sol = subsolv(low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d, sub_tol, sub_maxiter)
    return sol[0], low, upp, (alfa, beta, p0, q0, P, Q, b, a, c, d) + (sol,)


def kkt_residual(x, y, z, lam, xsi, eta, mu, zet, s, xmin, xmax, df0dx, fval, dfdx, a0, a, c, d):
    """Norm of the KKT residual of the original problem at a point
    (mma/CPU/_mma.py:530-590)."""
    dfdx = np.atleast_2d(dfdx)
    rex = df0dx + dfdx.T @ lam - xsi + eta
    rey = c + d * y - mu - lam
    rez = a0 - zet - a @ lam
    relam = fval - a * z - y + s
    r = np.concatenate([rex, rey, [rez], relam, xsi * (x - xmin), eta * (xmax - x), mu * y, [zet * z], lam * s])
    return np.sqrt(r @ r)
