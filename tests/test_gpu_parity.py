"""GPU parity: every gfx950 kernel against the FP64 oracle on the same seeded
inputs, through the C ABI (pyfantom_amd._lib).

Tolerances: index/integer work is bit-exact; FP64 kernels that sum in the
oracle's order are compared at 1e-15 relative, kernels that change the
summation order (the Walsh-Hadamard tile kernel, tree reductions) at
1e-13 relative to the largest entry, solves by 1e-10 relative compliance
against the FP64 oracle.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _dev(a, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _np(t):
    return t.detach().cpu().numpy()


def _case(oracle, nel, seed=0, bc="random"):
    dim = len(nel)
    K0 = oracle.k0_exact(dim, 0.3)
    el = oracle.elements(nel)
    n_nodes = int(np.prod([n + 1 for n in nel]))
    rng = np.random.default_rng(seed)
    E = rng.uniform(0.01, 1.0, len(el)) ** 3
    u = rng.standard_normal(n_nodes * dim)
    con = np.zeros(n_nodes * dim, bool)
    if bc == "random":
        con[rng.choice(n_nodes * dim, size=max(1, n_nodes * dim // 7), replace=False)] = True
    elif bc == "clamp":
        nodes = oracle.nodes(nel, nel)
        con[(np.repeat(nodes[:, 0] == 0, dim))] = True
    return K0, el, E, u, con


def _mask_from_con(con, dim):
    c = con.reshape(-1, dim).astype(np.uint8)
    m = np.zeros(c.shape[0], np.uint8)
    for d in range(dim):
        m |= c[:, d] << d
    return m


def _grid(nel):
    return (nel[0], nel[1], nel[2] if len(nel) == 3 else 0)


def _ku(kernel, nel, K0, E, u, con, mode="ku", b=None, dinv=None, omega=0.6):
    from pyfantom_amd import _lib
    from pyfantom_amd.stiffness import hex8_blocks
    torch = _torch()
    dim = len(nel)
    kid = _lib.KU_TILE if kernel == "tile" else _lib.KU_NODE
    if kid == _lib.KU_TILE:
        kb = hex8_blocks(K0)
        kptr = kb.ctypes.data
    else:
        k0d = _dev(K0)
        kptr = k0d.data_ptr()
    Ed, ud = _dev(E), _dev(u)
    md = _dev(_mask_from_con(con, dim)) if con is not None else None
    y = torch.empty_like(ud)
    g = _grid(nel)
    mp = md.data_ptr() if md is not None else None
    if mode == "ku":
        _lib.call("fantom_ku_f64", kid, *g, kptr, Ed.data_ptr(), mp, ud.data_ptr(), y.data_ptr())
        return _np(y)
    if mode == "jac":
        bd = _dev(b)
        dd = _dev(dinv) if dinv is not None else None      # None: the tile sweep forms D^-1 itself
        _lib.call("fantom_ku_jacobi_f64", kid, *g, kptr, Ed.data_ptr(), mp, bd.data_ptr(),
                  dd.data_ptr() if dd is not None else None, omega, ud.data_ptr(), y.data_ptr())
        return _np(y)
    if mode == "res":
        bd = _dev(b)
        _lib.call("fantom_ku_residual_f64", kid, *g, kptr, Ed.data_ptr(), mp, bd.data_ptr(), ud.data_ptr(),
                  y.data_ptr())
        return _np(y)
    if mode == "dot":
        npart = _lib.load().fantom_ku_partials(kid, *g)
        part = torch.zeros(npart, dtype=torch.float64, device="cuda")
        scal = torch.zeros(16, dtype=torch.float64, device="cuda")
        _lib.call("fantom_ku_dot_f64", kid, *g, kptr, Ed.data_ptr(), mp, ud.data_ptr(), y.data_ptr(),
                  part.data_ptr())
        _lib.call("fantom_finalize_f64", part.data_ptr(), npart, scal.data_ptr(), _lib.FIN_STORE, 5)
        return _np(y), float(scal[5].item())
    raise ValueError(mode)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


SIZES_3D = [(1, 1, 1), (3, 2, 5), (16, 4, 15), (17, 9, 31), (33, 17, 20), (30, 14, 45)]


@pytest.mark.parametrize("nel", SIZES_3D)
@pytest.mark.parametrize("bc", ["none", "random", "clamp"])
@pytest.mark.parametrize("kernel", ["node", "tile"])
def test_ku_3d(oracle, nel, bc, kernel):
    print(f"ku_3d {kernel} {nel} {bc}", flush=True)
    K0, el, E, u, con = _case(oracle, nel, bc=bc)
    ref = oracle.matvec(K0, el, E, u, con if bc != "none" else None)
    y = _ku(kernel, nel, K0, E, u, con if bc != "none" else None)
    tol = 1e-15 if kernel == "node" else 1e-13
    assert _rel(y, ref) <= tol
    if bc != "none":
        assert np.array_equal(y[con], u[con])


@pytest.mark.parametrize("nel", [(1, 1), (5, 3), (120, 40)])
@pytest.mark.parametrize("bc", ["none", "random"])
def test_ku_2d(oracle, nel, bc):
    print(f"ku_2d {nel} {bc}", flush=True)
    K0, el, E, u, con = _case(oracle, nel, bc=bc)
    ref = oracle.matvec(K0, el, E, u, con if bc != "none" else None)
    y = _ku("node", nel, K0, E, u, con if bc != "none" else None)
    assert _rel(y, ref) <= 1e-15


@pytest.mark.parametrize("kernel", ["node", "tile"])
def test_ku_epilogues(oracle, kernel):
    nel = (17, 9, 31)
    print(f"ku_epilogues {kernel}", flush=True)
    K0, el, E, u, con = _case(oracle, nel, bc="random", seed=5)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(len(u))
    dinv = 1.0 / oracle.diag(K0, el, E, len(u), con)
    ku = oracle.matvec(K0, el, E, u, con)
    jac = _ku(kernel, nel, K0, E, u, con, "jac", b, dinv, 0.6)
    assert _rel(jac, u + 0.6 * dinv * (b - ku)) <= 1e-12
    res = _ku(kernel, nel, K0, E, u, con, "res", b)
    assert _rel(res, b - ku) <= 1e-13
    y, s = _ku(kernel, nel, K0, E, u, con, "dot")
    assert _rel(y, ku) <= 1e-13
    assert abs(s - np.dot(u, ku)) <= 1e-12 * abs(np.dot(u, ku))


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 1.0, 2.0)])
@pytest.mark.parametrize("bc", ["none", "random", "clamp"])
def test_jacobi_own_diagonal(oracle, h, bc):
    """The tiled Jacobi sweep with no D^-1 argument forms D^-1 from the moduli
    (d_c * sum of E around the node, 1 on constrained DOFs): it equals the
    sweep with the oracle's assembled diagonal, for cubic and non-cubic
    boxes, with and without Dirichlet DOFs, across tile and chunk borders."""
    nel = (33, 19, 29)
    print(f"jacobi own diagonal {h} {bc}", flush=True)
    K0, el, E, u, con = _case(oracle, nel, bc=bc, seed=11)
    K0 = np.ascontiguousarray(oracle.k0_exact(3, 0.3, *h))
    c = con if bc != "none" else None
    b = np.random.default_rng(12).standard_normal(len(u))
    dinv = 1.0 / oracle.diag(K0, el, E, len(u), c)
    ku = oracle.matvec(K0, el, E, u, c)
    ref = u + 0.6 * dinv * (b - ku)
    own = _ku("tile", nel, K0, E, u, c, "jac", b, None, 0.6)
    assert _rel(own, ref) <= 1e-12


@pytest.mark.parametrize("nel", [(3, 2, 5), (17, 9, 31), (5, 3), (120, 40)])
def test_diag_dc(oracle, nel):
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"diag_dc {nel}", flush=True)
    K0, el, E, u, con = _case(oracle, nel, bc="random", seed=2)
    dim = len(nel)
    k0d, Ed, ud, md = _dev(K0), _dev(E), _dev(u), _dev(_mask_from_con(con, dim))
    d = torch.empty_like(ud)
    _lib.call("fantom_diag_f64", *_grid(nel), k0d.data_ptr(), Ed.data_ptr(), md.data_ptr(), d.data_ptr(), 0)
    assert _rel(_np(d), oracle.diag(K0, el, E, len(u), con)) <= 1e-15
    _lib.call("fantom_diag_f64", *_grid(nel), k0d.data_ptr(), Ed.data_ptr(), md.data_ptr(), d.data_ptr(), 1)
    assert _rel(_np(d), 1 / oracle.diag(K0, el, E, len(u), con)) <= 1e-15
    dc = torch.empty_like(Ed)
    _lib.call("fantom_dc_f64", *_grid(nel), k0d.data_ptr(), ud.data_ptr(), dc.data_ptr())
    assert _rel(_np(dc), oracle.process_dk(K0, el, u)) <= 1e-14


@pytest.mark.parametrize("nel", [(2, 2, 2), (8, 4, 6), (32, 16, 16), (5, 3, 7), (9, 6, 5), (6, 4), (120, 40)])
@pytest.mark.parametrize("normalize", [0, 1])
def test_transfers(oracle, nel, normalize):
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"transfers {nel} {normalize}", flush=True)
    dim = len(nel)
    rng = np.random.default_rng(3)
    nf = int(np.prod([n + 1 for n in nel])) * dim
    nc = int(np.prod([n // 2 + 1 for n in nel])) * dim
    v, w = rng.standard_normal(nf), rng.standard_normal(nc)
    vd, wd = _dev(v), _dev(w)
    rc = torch.empty(nc, dtype=torch.float64, device="cuda")
    _lib.call("fantom_restrict_f64", *_grid(nel), normalize, None, vd.data_ptr(), rc.data_ptr())
    assert _rel(_np(rc), oracle.restrict(v, nel, dim, bool(normalize))) <= 1e-15
    xf = torch.empty(nf, dtype=torch.float64, device="cuda")
    _lib.call("fantom_prolong_f64", *_grid(nel), normalize, None, wd.data_ptr(), xf.data_ptr(), 0)
    assert _rel(_np(xf), oracle.prolong(w, nel, dim, bool(normalize))) <= 1e-15
    # accumulate into an existing vector
    _lib.call("fantom_prolong_f64", *_grid(nel), normalize, None, wd.data_ptr(), vd.data_ptr(), 1)
    assert _rel(_np(vd), v + oracle.prolong(w, nel, dim, bool(normalize))) <= 1e-15


def test_coarsen(oracle):
    from pyfantom_amd import _lib
    torch = _torch()
    for nel in ((8, 4, 6), (6, 4)):
        dim = len(nel)
        E = np.random.default_rng(0).uniform(0, 1, int(np.prod(nel)))
        Ec = torch.empty(int(np.prod([n // 2 for n in nel])), dtype=torch.float64, device="cuda")
        Ed = _dev(E)
        _lib.call("fantom_coarsen_E_f64", *_grid(nel), Ed.data_ptr(), Ec.data_ptr())
        ref = E.reshape(*sum(([n // 2, 2] for n in nel), [])).mean(axis=tuple(range(1, 2 * dim, 2))).ravel()
        assert _rel(_np(Ec), ref) <= 1e-15
        nn = int(np.prod([n + 1 for n in nel]))
        m = np.random.default_rng(1).integers(0, 8 if dim == 3 else 4, nn).astype(np.uint8)
        mc = torch.empty(int(np.prod([n // 2 + 1 for n in nel])), dtype=torch.uint8, device="cuda")
        md = _dev(m)
        _lib.call("fantom_coarsen_mask", *_grid(nel), md.data_ptr(), mc.data_ptr())
        X = oracle.nodes(nel, nel)
        Xc = oracle.nodes([n // 2 for n in nel], [n // 2 for n in nel]) * 2
        lut = {tuple(p.astype(int)): i for i, p in enumerate(X)}
        ref = np.array([m[lut[tuple(p.astype(int))]] for p in Xc], np.uint8)
        assert np.array_equal(_np(mc), ref)


# (10, 6, 6) and (16, 8, 8): 1,617 and 4,131 DOF take the 128-wide blocked path
# with its GEMM update (ragged last block), the latter being the coarsest grid
# of the 256x128x128 cantilever after four coarsenings
@pytest.mark.parametrize("nel", [(4, 2, 2), (8, 4, 4), (8, 4), (10, 6, 6), (16, 8, 8)])
def test_dense_coarse_solve(oracle, nel):
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"dense {nel}", flush=True)
    K0, el, E, u, con = _case(oracle, nel, bc="clamp", seed=4)
    dim = len(nel)
    n = len(u)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    # keep the operands alive until the asynchronous kernel has run
    k0d, Ed, md = _dev(K0), _dev(E), _dev(_mask_from_con(con, dim))
    _lib.call("fantom_dense_assemble_f64", *_grid(nel), k0d.data_ptr(), Ed.data_ptr(), md.data_ptr(), A.data_ptr())
    Aref = oracle.assemble(K0, el, E, n, con).toarray()
    assert _rel(_np(A), Aref) <= 1e-15
    work = torch.empty(n * n + 1024, dtype=torch.float64, device="cuda")
    _lib.call("fantom_spd_inverse_f64", n, A.data_ptr(), work.data_ptr())
    Ainv = np.linalg.inv(Aref)
    assert _rel(_np(A), Ainv) <= 1e-9
    b = np.random.default_rng(0).standard_normal(n)
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    bd = _dev(b)
    _lib.call("fantom_gemv_f64", n, A.data_ptr(), bd.data_ptr(), y.data_ptr())
    assert np.abs(Aref @ _np(y) - b).max() <= 1e-9 * np.abs(b).max()


@pytest.mark.parametrize("nel,r", [((5, 4, 3), 1.5), ((40, 20, 20), 1.5), ((16, 8, 8), 2.0), ((120, 40), 1.5),
                                   ((30, 20), 3.3)])
def test_filter(oracle, nel, r):
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"filter {nel} {r}", flush=True)
    offs, w = oracle.filter_taps(len(nel), r)
    n = int(np.prod(nel))
    x = np.random.default_rng(0).uniform(0, 1, n)
    od, wd, xd = _dev(offs.reshape(-1)), _dev(w), _dev(x)
    norm = torch.empty(n, dtype=torch.float64, device="cuda")
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    g = _grid(nel)
    _lib.call("fantom_filter_normalizer_f64", *g, len(w), od.data_ptr(), wd.data_ptr(), norm.data_ptr())
    assert _rel(_np(norm), oracle.filter_normalizer(nel, offs, w)) <= 1e-15
    _lib.call("fantom_filter_fwd_f64", *g, len(w), od.data_ptr(), wd.data_ptr(), xd.data_ptr(), y.data_ptr())
    assert _rel(_np(y), oracle.filter_fwd(nel, offs, w, x)) <= 1e-15
    _lib.call("fantom_filter_adj_f64", *g, len(w), od.data_ptr(), wd.data_ptr(), norm.data_ptr(), xd.data_ptr(),
              y.data_ptr())
    assert _rel(_np(y), oracle.filter_adj(nel, offs, w, x)) <= 1e-15


@pytest.mark.parametrize("n,seed", [(4, 0), (1000, 1), (524288, 2)])
def test_oc_update(oracle, n, seed):
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"oc {n}", flush=True)
    rng = np.random.default_rng(seed)
    if n == 4:
        x = np.full(4, 0.5)
        df = -np.array([4.0, 1.0, 0.25, 1.0])
    else:
        x = rng.uniform(0, 1, n)
        df = -rng.uniform(0.0, 5, n)
    dg = np.full(n, 1.0 / n)
    ref, lm = oracle.oc_update(x, df, dg, 0.4 if n > 4 else 0.5, 0.2)
    xd, dfd, dgd = _dev(x), _dev(df), _dev(dg)
    ocp = torch.empty_like(xd)
    state = torch.zeros(16, dtype=torch.float64, device="cuda")
    part = torch.zeros(16 * 1024, dtype=torch.float64, device="cuda")
    xn = torch.empty_like(xd)
    _lib.call("fantom_oc_update_f64", n, xd.data_ptr(), dfd.data_ptr(), dgd.data_ptr(), 0.4 if n > 4 else 0.5, 0.2,
              ocp.data_ptr(), state.data_ptr(), part.data_ptr(), xn.data_ptr())
    lam = float(state[2].item())
    # same sequence of midpoints unless a constraint sum lands within rounding
    # of the target; either way the multiplier agrees to the bisection width
    assert abs(lam - lm) <= 2e-8 * lm
    assert np.abs(_np(xn) - ref).max() <= 1e-7


def test_simp(oracle):
    from pyfantom_amd import _lib
    torch = _torch()
    # random densities plus the special points 0, 0.5, 1 and an overshoot
    rho = np.concatenate([np.random.default_rng(0).uniform(0, 1, 10000), [0.0, 0.5, 1.0, 1.2, 0.1]])
    rd = _dev(rho)
    E = torch.empty_like(rd)
    dE = torch.empty_like(rd)
    # (void, E_mul): the clips below and above, E_mul > 1 and E_mul < void
    for void, emul in ((1e-6, 1.0), (0.01, 2.0), (0.01, 0.001)):
        for hv in (0, 1):
            _lib.call("fantom_simp_f64", len(rho), rd.data_ptr(), 0, 3.0, void, emul, hv, 2.0, 0.5, E.data_ptr(),
                      dE.data_ptr())
            assert _rel(_np(E), oracle.penalize_reference(rho, 3.0, void, emul, bool(hv))) <= 1e-14
            assert _rel(_np(dE), oracle.penalize_grad_reference(rho, 3.0, emul, bool(hv))) <= 1e-14
    _lib.call("fantom_simp_f64", len(rho), rd.data_ptr(), 1, 3.0, 1e-9, 1.0, 0, 2.0, 0.5, E.data_ptr(), dE.data_ptr())
    assert _rel(_np(E), oracle.penalize_modified(rho, 3.0, 1.0, 1e-9)) <= 1e-14
    assert _rel(_np(dE), oracle.penalize_grad_modified(rho, 3.0, 1.0, 1e-9)) <= 1e-14


def _galerkin_ref(oracle, nel, E, con, levels=2):
    """scipy Galerkin hierarchy: A_{l+1}^U = P^T A_l^U P (unconstrained), and
    the coarse masks by the coincident-node rule."""
    import scipy.sparse as sps
    K0 = oracle.k0_exact(3, 0.3)
    AU = [oracle.assemble(K0, oracle.elements(nel), E, int(np.prod([n + 1 for n in nel])) * 3)]
    cons = [con]
    nels = [tuple(nel)]
    for l in range(levels):
        cur = nels[-1]
        P = oracle.transfer_matrix(cur, 3, False).T.tocsr()
        AU.append((P.T @ AU[-1] @ P).tocsr())
        X = oracle.nodes(cur, cur)
        nc = tuple(n // 2 for n in cur)
        Xc = oracle.nodes(nc, nc) * 2
        lut = {tuple(p.astype(int)): i for i, p in enumerate(X)}
        cc = np.zeros(len(Xc) * 3, bool)
        for i, p in enumerate(Xc):
            f = lut[tuple(p.astype(int))]
            cc[3 * i:3 * i + 3] = cons[-1][3 * f:3 * f + 3]
        cons.append(cc)
        nels.append(nc)

    def apply(l, u):
        c = cons[l]
        free = (~c).astype(float)
        return free * (AU[l] @ (free * u)) + c * u

    return apply, nels, cons


@pytest.mark.parametrize("nel", [(8, 4, 4), (16, 8, 12), (32, 16, 16)])
def test_galerkin_levels(oracle, nel):
    """Level-1 matrix-free Galerkin operator and level-2 explicit element
    matrices against scipy P^T A P on the same moduli and masks."""
    import ctypes
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"galerkin {nel}", flush=True)
    K0 = np.ascontiguousarray(oracle.k0_exact(3, 0.3))
    rng = np.random.default_rng(5)
    E = rng.uniform(0.01, 1, int(np.prod(nel))) ** 3
    _, el, _, _, con = _case(oracle, nel, bc="clamp")
    apply, nels, cons = _galerkin_ref(oracle, nel, E, con, levels=3 if all(n % 8 == 0 for n in nel) else 2)
    G = np.zeros(8 * 576)
    B = np.zeros(576)
    Gd = np.zeros(8 * 24)
    _lib.check(_lib.load().fantom_galerkin_tables(K0.ctypes.data, G.ctypes.data, B.ctypes.data, Gd.ctypes.data,
                                                  None))
    Gd_d, Bd, Ed = _dev(G), _dev(B), _dev(E)
    # level 1 (matrix-free)
    g1 = nels[1]
    n1 = int(np.prod([n + 1 for n in g1])) * 3
    u1 = rng.standard_normal(n1)
    m1 = _dev(_mask_from_con(cons[1], 3))
    u1d = _dev(u1)
    y1 = torch.empty_like(u1d)
    _lib.call("fantom_gal1_ku_f64", *g1, Bd.data_ptr(), Ed.data_ptr(), m1.data_ptr(), u1d.data_ptr(), y1.data_ptr())
    assert _rel(_np(y1), apply(1, u1)) <= 1e-13
    # the same operator with its 20 class values as kernel arguments (cubic elements)
    cls, cubic = np.zeros(20), np.zeros(1, np.int32)
    _lib.check(_lib.load().fantom_gal1_classes(B.ctypes.data, cls.ctypes.data, cubic.ctypes.data))
    assert cubic[0] == 1
    y1c = torch.empty_like(u1d)
    _lib.call("fantom_gal1_ku_cls_f64", *g1, cls.ctypes.data, Ed.data_ptr(), m1.data_ptr(), u1d.data_ptr(),
              y1c.data_ptr())
    assert _rel(_np(y1c), apply(1, u1)) <= 1e-13
    # level 2 (explicit element matrices built from the fine moduli)
    g2 = nels[2]
    n2 = int(np.prod([n + 1 for n in g2])) * 3
    Ke = torch.empty(int(np.prod(g2)) * 576, dtype=torch.float64, device="cuda")
    _lib.call("fantom_gal_build_f64", *g2, 0, Gd_d.data_ptr(), Ed.data_ptr(), None, Ke.data_ptr())
    u2 = rng.standard_normal(n2)
    m2 = _dev(_mask_from_con(cons[2], 3))
    u2d = _dev(u2)
    y2 = torch.empty_like(u2d)
    _lib.call("fantom_elem_ku_f64", *g2, Ke.data_ptr(), m2.data_ptr(), u2d.data_ptr(), y2.data_ptr())
    assert _rel(_np(y2), apply(2, u2)) <= 1e-12
    # the same matrices from the 64 level-2 tables (the drivers' build)
    G2 = np.zeros(64 * 576)
    _lib.check(_lib.load().fantom_galerkin_tables2(G.ctypes.data, G2.ctypes.data))
    G2d = _dev(G2)
    Ke2 = torch.empty_like(Ke)
    _lib.call("fantom_gal2_build_f64", *g2, G2d.data_ptr(), Ed.data_ptr(), Ke2.data_ptr())
    torch.cuda.synchronize()
    assert _rel(_np(Ke2), _np(Ke)) <= 1e-13
    _lib.call("fantom_elem_ku_f64", *g2, Ke2.data_ptr(), m2.data_ptr(), u2d.data_ptr(), y2.data_ptr())
    assert _rel(_np(y2), apply(2, u2)) <= 1e-12
    # level-1 explicit matrices agree with the matrix-free form
    Ke1 = torch.empty(int(np.prod(g1)) * 576, dtype=torch.float64, device="cuda")
    _lib.call("fantom_elem_from_fine_f64", *g1, Gd_d.data_ptr(), Ed.data_ptr(), Ke1.data_ptr())
    y1e = torch.empty_like(u1d)
    _lib.call("fantom_elem_ku_f64", *g1, Ke1.data_ptr(), m1.data_ptr(), u1d.data_ptr(), y1e.data_ptr())
    assert _rel(_np(y1e), apply(1, u1)) <= 1e-13
    # level 3 from explicit level-2 matrices (src_kind 1)
    if all(n % 8 == 0 for n in nel):
        g3 = tuple(n // 2 for n in g2)
        Ke3 = torch.empty(int(np.prod(g3)) * 576, dtype=torch.float64, device="cuda")
        _lib.call("fantom_gal_build_f64", *g3, 1, Gd_d.data_ptr(), None, Ke.data_ptr(), Ke3.data_ptr())
        # A3 = P^T A2^U P of the scipy hierarchy, with the level-3 Dirichlet rows
        n3 = int(np.prod([n + 1 for n in g3])) * 3
        u3 = rng.standard_normal(n3)
        m3 = _dev(_mask_from_con(cons[3], 3))
        u3d = _dev(u3)
        y3 = torch.empty_like(u3d)
        _lib.call("fantom_elem_ku_f64", *g3, Ke3.data_ptr(), m3.data_ptr(), u3d.data_ptr(), y3.data_ptr())
        assert _rel(_np(y3), apply(3, u3)) <= 1e-12


@pytest.mark.parametrize("nel", [(16, 8, 12), (32, 16, 16)])
def test_stencil_levels(oracle, nel):
    """Assembled 27-point stencil of the level-2 Galerkin operator against
    scipy P^T A P (K.u), and its Jacobi / residual / K.u-dot epilogues
    against the same formulas evaluated in numpy."""
    from pyfantom_amd import _lib
    torch = _torch()
    print(f"stencil {nel}", flush=True)
    K0 = np.ascontiguousarray(oracle.k0_exact(3, 0.3))
    rng = np.random.default_rng(8)
    E = rng.uniform(0.01, 1, int(np.prod(nel))) ** 3
    _, el, _, _, con = _case(oracle, nel, bc="clamp")
    apply, nels, cons = _galerkin_ref(oracle, nel, E, con)
    G = np.zeros(8 * 576)
    B = np.zeros(576)
    Gd = np.zeros(8 * 24)
    _lib.check(_lib.load().fantom_galerkin_tables(K0.ctypes.data, G.ctypes.data, B.ctypes.data, Gd.ctypes.data,
                                                  None))
    Gd_d, Ed = _dev(G), _dev(E)
    g2 = nels[2]
    nn2 = int(np.prod([n + 1 for n in g2]))
    Ke = torch.empty(int(np.prod(g2)) * 576, dtype=torch.float64, device="cuda")
    _lib.call("fantom_gal_build_f64", *g2, 0, Gd_d.data_ptr(), Ed.data_ptr(), None, Ke.data_ptr())
    m2 = _dev(_mask_from_con(cons[2], 3))
    S = torch.empty(_lib.load().fantom_stencil_size(*g2), dtype=torch.float64, device="cuda")
    _lib.call("fantom_stencil_build_f64", *g2, Ke.data_ptr(), m2.data_ptr(), S.data_ptr())
    u = rng.standard_normal(nn2 * 3)
    b = rng.standard_normal(nn2 * 3)
    dinv = rng.uniform(0.5, 2.0, nn2 * 3)
    ud, bd, dd = _dev(u), _dev(b), _dev(dinv)
    y = torch.empty_like(ud)
    ref = apply(2, u)
    _lib.call("fantom_stencil_ku_f64", 0, *g2, S.data_ptr(), ud.data_ptr(), y.data_ptr(), None, None, 0.0, None)
    assert _rel(_np(y), ref) <= 1e-12
    _lib.call("fantom_stencil_ku_f64", 1, *g2, S.data_ptr(), ud.data_ptr(), y.data_ptr(), bd.data_ptr(),
              dd.data_ptr(), 0.6, None)
    assert _rel(_np(y), u + 0.6 * dinv * (b - ref)) <= 1e-12
    _lib.call("fantom_stencil_ku_f64", 2, *g2, S.data_ptr(), ud.data_ptr(), y.data_ptr(), bd.data_ptr(), None, 0.0,
              None)
    assert _rel(_np(y), b - ref) <= 1e-12
    _lib.call("fantom_stencil_ku_f64", 4, *g2, S.data_ptr(), None, y.data_ptr(), bd.data_ptr(), dd.data_ptr(), 0.6,
              None)
    x0 = 0.6 * dinv * b
    assert _rel(_np(y), x0 + 0.6 * dinv * (b - apply(2, x0))) <= 1e-12
    npart = _lib.load().fantom_stencil_partials(*g2)
    part = torch.zeros(npart, dtype=torch.float64, device="cuda")
    scal = torch.zeros(16, dtype=torch.float64, device="cuda")
    _lib.call("fantom_stencil_ku_f64", 3, *g2, S.data_ptr(), ud.data_ptr(), y.data_ptr(), None, None, 0.0,
              part.data_ptr())
    _lib.call("fantom_finalize_f64", part.data_ptr(), npart, scal.data_ptr(), _lib.FIN_STORE, 5)
    assert _rel(_np(y), ref) <= 1e-12
    assert abs(float(scal[5].item()) - u @ ref) <= 1e-12 * abs(u @ ref)


@pytest.mark.parametrize("nel", SIZES_3D + [(64, 32, 40)])
def test_jacobi_prolong_sweep(oracle, nel):
    """The first post-smoothing sweep with the prolongation formed as x
    loads: v = x + P ec (the oracle's prolongation, normalize 0, nothing on
    constrained DOFs), then v + omega D^-1 (b - A v) with D^-1 = 1 / diag(A),
    against the oracle's assembled A.  Odd element counts give fine nodes
    with a missing parent; (64, 32, 40) runs several tiles and y chunks."""
    from pyfantom_amd import _lib
    from pyfantom_amd.stiffness import hex8_blocks
    torch = _torch()
    print(f"jacobi prolong {nel}", flush=True)
    K0, el, E, x, con = _case(oracle, nel, bc="random", seed=11)
    n = len(x)
    rng = np.random.default_rng(12)
    b = rng.standard_normal(n)
    nc = int(np.prod([m // 2 + 1 for m in nel])) * 3
    ec = rng.standard_normal(nc)
    A = oracle.assemble(K0, el, E, n, con)
    pe = oracle.prolong(ec, nel, 3, normalize=False)
    pe[con] = 0.0
    v = x + pe
    ref = v + 0.6 * (b - A @ v) / A.diagonal()
    kb = hex8_blocks(K0)
    # device operands held in names: a temporary freed right after
    # data_ptr() can hand its block to the next allocation before the launch
    md = _dev(_mask_from_con(con, 3))
    Ed, bd, xd, ecd = _dev(E), _dev(b), _dev(x), _dev(ec)
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.call("fantom_ku_jacobi_prolong_f64", *nel, kb.ctypes.data, Ed.data_ptr(), md.data_ptr(), bd.data_ptr(),
              0.6, xd.data_ptr(), ecd.data_ptr(), y.data_ptr())
    assert _rel(_np(y), ref) <= 1e-13
    f32 = torch.float32
    E32, b32, x32, ec32 = _dev(E, f32), _dev(b, f32), _dev(x, f32), _dev(ec, f32)
    y32 = torch.empty(n, dtype=f32, device="cuda")
    _lib.call("fantom_ku_jacobi_prolong_f32", *nel, kb.ctypes.data, E32.data_ptr(), md.data_ptr(), b32.data_ptr(),
              0.6, x32.data_ptr(), ec32.data_ptr(), y32.data_ptr())
    assert _rel(_np(y32).astype(np.float64), ref) <= 2e-5


@pytest.mark.parametrize("nel", [(8, 4, 4), (30, 14, 45)])
def test_jacobi0_matches_scale_then_sweep(oracle, nel):
    """The fused first sweep (x0 = omega D^-1 b formed on load) equals the
    scaling followed by one Jacobi sweep, against the oracle's assembled K."""
    from pyfantom_amd import _lib
    from pyfantom_amd.stiffness import hex8_blocks
    torch = _torch()
    print(f"jacobi0 {nel}", flush=True)
    K0, el, E, u, con = _case(oracle, nel, bc="clamp", seed=6)
    n = len(u)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(n)
    dinv = rng.uniform(0.5, 2.0, n)
    A = oracle.assemble(K0, el, E, n, con)
    x0 = 0.6 * dinv * b
    ref = x0 + 0.6 * dinv * (b - A @ x0)
    kb = hex8_blocks(K0)
    Ed, bd, dd, md = _dev(E), _dev(b), _dev(dinv), _dev(_mask_from_con(con, 3))
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.call("fantom_ku_jacobi0_f64", _lib.KU_TILE, *nel, kb.ctypes.data, Ed.data_ptr(), md.data_ptr(),
              bd.data_ptr(), dd.data_ptr(), 0.6, y.data_ptr())
    assert _rel(_np(y), ref) <= 1e-13
