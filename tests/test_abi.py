"""CPU checks of the C ABI: the library loads, exports every symbol that
include/fantom.h declares, the Python binding table matches the header, and
the host-only helper (the Walsh-Hadamard block form) is exact."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "fantom.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(fantom_[A-Za-z0-9_]+)\s*\(", src))


def test_library_exports_header():
    from pyfantom_amd import _lib
    L = _lib.load()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n


def test_binding_table_matches_header():
    from pyfantom_amd import _lib
    assert set(_lib.SIGNATURES) == _declared()


def _prototypes():
    """name -> list of parameter type strings, from the header declarations."""
    src = open(os.path.join(ROOT, "include", "fantom.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"#.*", "", src)
    out = {}
    for m in re.finditer(r"\b(fantom_[A-Za-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = params
    return out


def test_binding_arity_and_kinds_match_header():
    """Every ctypes binding has the header's parameter count, and each slot
    has the right kind (pointer / floating / integer), so that no argument
    (the trailing stream above all) is passed with ctypes' default int
    conversion."""
    import ctypes
    from pyfantom_amd import _lib
    protos = _prototypes()
    assert set(protos) == set(_lib.SIGNATURES)
    for name, (_, argtypes) in _lib.SIGNATURES.items():
        params = protos[name]
        assert len(argtypes) == len(params), (name, len(argtypes), len(params))
        for a, p in zip(argtypes, params):
            if "*" in p:
                assert a in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(a, ctypes._Pointer), (name, p, a)
            elif p.split()[0] in ("double", "float"):
                assert a in (ctypes.c_double, ctypes.c_float), (name, p, a)
            else:
                assert a in (ctypes.c_int, ctypes.c_long, ctypes.c_longlong, ctypes.c_size_t), (name, p, a)


def test_abi_version():
    from pyfantom_amd import _lib
    assert _lib.load().fantom_abi_version() == 3


def test_gal1_classes(oracle):
    """The level-1 table of a cubic element has 20 classes of equal entries
    for any Poisson ratio (host code, no GPU); a non-cubic box does not."""
    from pyfantom_amd import _lib
    lib = _lib.load()
    for nu, h, want in ((0.3, (1, 1, 1), 1), (0.4471, (1, 1, 1), 1), (0.3, (1, 1.5, 0.7), 0)):
        K0 = np.ascontiguousarray(oracle.k0_exact(3, nu, *h))
        G, B, Gd = np.zeros(8 * 576), np.zeros(576), np.zeros(8 * 24)
        _lib.check(lib.fantom_galerkin_tables(K0.ctypes.data, G.ctypes.data, B.ctypes.data, Gd.ctypes.data, None))
        cls, cubic = np.zeros(20), np.zeros(1, np.int32)
        _lib.check(lib.fantom_gal1_classes(B.ctypes.data, cls.ctypes.data, cubic.ctypes.data))
        assert cubic[0] == want
        if want:
            nz = B[np.abs(B) > 1e-12 * np.abs(B).max()]
            assert len(nz) == 279
            assert np.abs(np.abs(nz)[:, None] - np.abs(cls)[None, :]).min(axis=1).max() <= 1e-12 * np.abs(B).max()


def test_stencil_size():
    """The stencil holds 126 coefficients per node in tiles of 64 nodes (host
    arithmetic, no GPU); bad grids return -1."""
    from pyfantom_amd import _lib
    lib = _lib.load()
    for nel in ((16, 8, 8), (64, 32, 32), (1, 1, 1)):
        nn = int(np.prod([n + 1 for n in nel]))
        assert lib.fantom_stencil_size(*nel) == -(-nn // 64) * 64 * 126
    assert lib.fantom_stencil_size(16, 8, 0) == -1


def test_hex8_blocks_reconstruct(oracle):
    """K0 = (1/8)... rebuilt from the 8 blocks equals K0 (host code, no GPU)."""
    from pyfantom_amd.stiffness import hex8_blocks
    for h in ((1, 1, 1), (0.5, 1, 2)):
        K0 = oracle.k0_exact(3, 0.3, *h)
        B = hex8_blocks(K0).reshape(8, 3, 3)
        loc = [0, 1, 3, 2, 4, 5, 7, 6]
        Kn = np.zeros((24, 24))
        for b in range(8):
            for b2 in range(8):
                Kn[3 * b:3 * b + 3, 3 * b2:3 * b2 + 3] = K0[3 * loc[b]:3 * loc[b] + 3, 3 * loc[b2]:3 * loc[b2] + 3]
        H = lambda b, t: -1.0 if bin(b & t).count("1") % 2 else 1.0
        R = np.zeros((24, 24))
        for b in range(8):
            for c in range(3):
                for b2 in range(8):
                    for d in range(3):
                        R[3 * b + c, 3 * b2 + d] = sum(H(b, s ^ (1 << c)) * B[s, c, d] * H(b2, s ^ (1 << d))
                                                       for s in range(8))
        assert np.abs(R - Kn).max() <= 1e-14 * np.abs(Kn).max()


def test_hex8_blocks_rejects_non_box():
    """A K0 without the reflection symmetry is refused with ValueError."""
    from pyfantom_amd.stiffness import hex8_blocks
    K = np.random.default_rng(0).standard_normal((24, 24))
    with pytest.raises(ValueError):
        hex8_blocks(K + K.T)


def test_product_filter_taps_match_oracle(oracle):
    from pyfantom_amd.geom import filter_taps
    for dim, r, sp in ((3, 1.5, (1, 1, 1)), (2, 1.5, (1, 1)), (3, 2.0, (1, 1, 1)), (3, 2.0, (1, 0.5, 2))):
        o1, w1 = filter_taps(dim, r, sp)
        o2, w2 = oracle.filter_taps(dim, r, sp)
        assert np.array_equal(o1, o2) and np.array_equal(w1, w2)


def _galerkin_tables(K0):
    import ctypes
    from pyfantom_amd import _lib
    G = np.zeros(8 * 576)
    B = np.zeros(576)
    Gd = np.zeros(8 * 24)
    chk = np.zeros(2)
    _lib.check(_lib.load().fantom_galerkin_tables(K0.ctypes.data, G.ctypes.data, B.ctypes.data, Gd.ctypes.data,
                                                  chk.ctypes.data))
    return G.reshape(8, 24, 24), B.reshape(8, 8, 3, 3), Gd.reshape(8, 24), chk


def _natural(K0):
    loc = [0, 1, 3, 2, 4, 5, 7, 6]
    idx = np.array([3 * loc[b] + c for b in range(8) for c in range(3)])
    return K0[np.ix_(idx, idx)]


def test_galerkin_tables_vs_sparse_product(oracle):
    """Child matrices G_c = P_c^T K0 P_c equal the scipy Galerkin product on a
    single coarse element (2x2x2 fine elements, explicit trilinear P); the
    Walsh-Hadamard block form reproduces K_E u for random child moduli."""
    K0 = np.ascontiguousarray(oracle.k0_exact(3, 0.3))
    G, B, Gd, chk = _galerkin_tables(K0)
    assert chk[0] < 1e-13 and chk[1] < 1e-13
    # Galerkin on one coarse element with random child moduli, assembled
    # with scipy from the fine mesh and the explicit (unnormalised) P
    rng = np.random.default_rng(0)
    Ec = rng.uniform(0.01, 1, 8)
    nel = (2, 2, 2)
    el = oracle.elements(nel)
    # fine element index (x slowest) of child c = cx + 2cy + 4cz
    Ef = np.zeros(8)
    for c in range(8):
        cx, cy, cz = c & 1, (c >> 1) & 1, c >> 2
        Ef[(cx * 2 + cy) * 2 + cz] = Ec[c]
    A = oracle.assemble(K0, el, Ef, 27 * 3)
    P = oracle.transfer_matrix(nel, 3, False).T.toarray()
    Ac = P.T @ A.toarray() @ P          # coarse element matrix in coarse node order
    # coarse nodes of a 1x1x1 grid (reference numbering) -> natural corners
    Ac_nat = _natural_from_nodes(Ac)
    KE = np.einsum("c,cij->ij", Ec, G)
    assert np.abs(KE - Ac_nat).max() <= 1e-13 * np.abs(Ac_nat).max()
    # diagonal table
    assert np.allclose(Gd, np.einsum("cii->ci", G))
    # Walsh-Hadamard evaluation used by the level-1 kernel
    u = rng.standard_normal(24)
    H = np.array([[(-1) ** bin(b & t).count("1") for t in range(8)] for b in range(8)], float)
    w = [H.T @ u[d::3] for d in range(3)]            # w_d[t] = sum_b H[b,t] u[b,d]
    eh = H.T @ Ec
    z = np.zeros((3, 8))
    for s in range(8):
        uh = np.array([w[0][s ^ 1], w[1][s ^ 2], w[2][s ^ 4]])
        acc = np.zeros(3)
        for s2 in range(8):
            uh2 = np.array([w[0][s2 ^ 1], w[1][s2 ^ 2], w[2][s2 ^ 4]])
            acc += eh[s ^ s2] * (B[s, s2] @ uh2)
        z[0][s ^ 1], z[1][s ^ 2], z[2][s ^ 4] = acc
    y = np.zeros(24)
    for d in range(3):
        y[d::3] = H @ z[d]
    assert np.abs(y - KE @ u).max() <= 1e-13 * np.abs(KE @ u).max()


def _natural_from_nodes(A):
    """Reorder a 1x1x1-element coarse matrix from node numbering
    n = k + 2 i + 4 j (y slowest) to natural corners b = i + 2 j + 4 k."""
    perm = []
    for b in range(8):
        i, j, k = b & 1, (b >> 1) & 1, b >> 2
        n = k + 2 * i + 4 * j
        perm += [3 * n, 3 * n + 1, 3 * n + 2]
    perm = np.array(perm)
    return A[np.ix_(perm, perm)]


def test_level2_tables_vs_sparse_product(oracle):
    """The 64 level-2 tables (fantom_galerkin_tables2) give the element matrix
    of one level-2 element over 4x4x4 fine elements equal to the scipy
    two-level Galerkin product P1^T P0^T A P0 P1 with random fine moduli."""
    from pyfantom_amd import _lib
    K0 = np.ascontiguousarray(oracle.k0_exact(3, 0.3))
    G, B, Gd, chk = _galerkin_tables(K0)
    G2 = np.zeros(64 * 576)
    _lib.check(_lib.load().fantom_galerkin_tables2(np.ascontiguousarray(G).ctypes.data, G2.ctypes.data))
    G2 = G2.reshape(64, 24, 24)
    rng = np.random.default_rng(1)
    nel = (4, 4, 4)
    Ef = rng.uniform(0.01, 1, 64)          # fine moduli, reference element order (x slowest)
    A = oracle.assemble(K0, oracle.elements(nel), Ef, 125 * 3).toarray()
    P0 = oracle.transfer_matrix(nel, 3, False).T.toarray()
    P1 = oracle.transfer_matrix((2, 2, 2), 3, False).T.toarray()
    Ac = P1.T @ P0.T @ A @ P0 @ P1
    Ac_nat = _natural_from_nodes(Ac)
    # table index f = 8 c + g: level-1 child c, fine child g (bits x + 2y + 4z)
    w = np.zeros(64)
    for c in range(8):
        for g in range(8):
            fx = 2 * (c & 1) + (g & 1)
            fy = 2 * ((c >> 1) & 1) + ((g >> 1) & 1)
            fz = 2 * (c >> 2) + (g >> 2)
            w[8 * c + g] = Ef[(fx * 4 + fy) * 4 + fz]
    K2 = np.einsum("f,fij->ij", w, G2)
    assert np.abs(K2 - Ac_nat).max() <= 1e-13 * np.abs(Ac_nat).max()
