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


def test_abi_version():
    from pyfantom_amd import _lib
    assert _lib.load().fantom_abi_version() == 1


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


def test_product_connectivity_matches_oracle(oracle):
    from pyfantom_amd.geom import generate_elements
    for nel in ((3, 2, 4), (5, 3), (128, 64, 64)):
        assert np.array_equal(generate_elements(nel), oracle.elements(nel))


def test_product_filter_taps_match_oracle(oracle):
    from pyfantom_amd.geom import filter_taps
    for dim, r, sp in ((3, 1.5, (1, 1, 1)), (2, 1.5, (1, 1)), (3, 2.0, (1, 1, 1)), (3, 2.0, (1, 0.5, 2))):
        o1, w1 = filter_taps(dim, r, sp)
        o2, w2 = oracle.filter_taps(dim, r, sp)
        assert np.array_equal(o1, o2) and np.array_equal(w1, w2)
