"""Writes the committed fixtures of tests/golden from the FP64 oracle.

The reference's own Python cannot run here (it needs Numba and CuPy), so
these fixtures are produced by the oracle, whose functions are pinned
independently in tests/test_oracle.py (brute-force connectivity, exact
rational K0 integration).  Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

SIZES = ["120x40", "64x32x32", "96x48x48", "128x64x64", "192x96x96", "256x128x128"]


def main():
    hashes = {}
    for key in SIZES:
        nel = tuple(int(v) for v in key.split("x"))
        hashes[key] = hashlib.sha256(O.elements(nel).tobytes()).hexdigest()
    with open(os.path.join(HERE, "connectivity_sha256.json"), "w") as f:
        json.dump(hashes, f, indent=1)
    np.savez(os.path.join(HERE, "k0.npz"), hex8=O.k0_exact(3, 0.3), quad4=O.k0_exact(2, 0.3))


if __name__ == "__main__":
    main()
