"""Writes tests/golden/dstev_cases.json: known-answer vectors for the
symmetric tridiagonal eigenproblem (the LAPACK dstev contract the reference
calls through src/ietl/ietl2lapack.h:44-99), computed with scipy's LAPACK
binding scipy.linalg.lapack.dstev.

Run in this container (scipy is the checker, not part of the product):
    python tests/golden/make_fixtures.py
"""
import json
import os

import numpy as np
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def case(d, e, label):
    w, z, info = lapack.dstev(np.array(d, dtype=float), np.array(e, dtype=float), compute_v=1)
    assert info == 0
    return dict(label=label, d=list(map(float, d)), e=list(map(float, e)),
                eigenvalues=list(map(float, w)), zlast_abs=list(map(float, np.abs(z[-1, :]))))


def main():
    rng = np.random.default_rng(20240601)
    cases = []
    for n in (2, 3, 5, 8, 13, 21, 40, 64):
        cases.append(case(rng.normal(size=n), rng.normal(size=max(n - 1, 0)), f"normal n={n}"))
    # Lanczos-like: positive off-diagonals shrinking by orders of magnitude
    for n in (10, 30, 60):
        d = rng.uniform(-2, 2, n)
        e = 10.0 ** rng.uniform(-12, 0, n - 1)
        cases.append(case(d, e, f"graded n={n}"))
    # exact multiplicities: a direct sum of equal blocks
    blk_d = [1.0, -0.5, 2.0]
    blk_e = [0.3, 0.7]
    cases.append(case(blk_d * 3, blk_e + [0.0] + blk_e + [0.0] + blk_e, "three equal blocks"))
    # Toeplitz: eigenvalues a + 2b cos(k pi/(n+1)) in closed form
    n = 25
    cases.append(case([0.5] * n, [-1.0] * (n - 1), "toeplitz n=25"))
    with open(os.path.join(HERE, "dstev_cases.json"), "w") as f:
        json.dump(cases, f, indent=0)


if __name__ == "__main__":
    main()
