"""Writes tests/golden/xxz_large_oracle.json: the oracle's FP64 sparsediag
(lowest two levels, sparsediag.h:93-149 semantics) of random-bond XXZ chains
at the sizes of the batched and served workloads that the small-size batch
tests do not reach:

* batch-shaped: L = 18..20, bonds J_i ~ U[0.5, 1.5], Jz_i = Delta J_i with
  Delta ~ U[0, 2], open or periodic, Sz = 0 (Sz = 1/2 for odd L);
* serve-shaped: L = 22..26 with a uniform field h ~ U[0, 0.5] and a nonzero
  Sz sector (up to D = 5,311,735).

Each problem's couplings are stored with its eigenvalues, so the GPU test
rebuilds exactly these models.  The oracle is the checker; this script runs
it once in the build container, one process per problem (a few minutes and
up to about 10 GB of memory for the largest):
    python tests/golden/make_xxz_large_fixture.py
"""
import json
import math
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402

# (L, periodic, seed, field, n_up)
PROBLEMS = [
    (18, True, 101, 0.0, 9),
    (19, False, 102, 0.0, 9),
    (20, True, 103, 0.0, 10),
    (20, False, 104, 0.0, 10),
    (22, True, 105, None, 12),
    (24, False, 106, None, 14),
    (26, True, 107, None, 16),
]


def problem(L, periodic, seed, field, n_up):
    rng = np.random.default_rng(seed)
    nb = L if periodic else L - 1
    J = rng.uniform(0.5, 1.5, nb)
    delta = float(rng.uniform(0, 2))
    h = float(rng.uniform(0, 0.5)) if field is None else field
    return dict(L=L, periodic=periodic, n_up=n_up, dim=math.comb(L, n_up), jxy=J.tolist(),
                jz=(delta * J).tolist(), h=h)


def model(p):
    L = p["L"]
    s = np.arange(len(p["jxy"]))
    return O.Model(L, s, (s + 1) % L, np.array(p["jz"]), np.array(p["jxy"]), h=p["h"])


def solve(p):
    t0 = time.time()
    r = O.sparsediag(model(p), p["n_up"], nev=2, seed=1)
    assert r["error_code"] == 0 and r["dim"] == p["dim"]
    return dict(p, iterations=r["iterations"], eigenvalues=[float(v) for v in r["ev"]],
                seconds=round(time.time() - t0, 1))


def main():
    probs = [problem(*q) for q in PROBLEMS]
    with ProcessPoolExecutor(max_workers=4) as ex:
        out = list(ex.map(solve, probs))
    for r in out:
        print(f"L={r['L']} n_up={r['n_up']} dim {r['dim']} iterations {r['iterations']} "
              f"ev {r['eigenvalues']} ({r['seconds']} s)", flush=True)
    with open(os.path.join(HERE, "xxz_large_oracle.json"), "w") as f:
        json.dump(dict(nev=2, seed=1, problems=out), f, indent=1)


if __name__ == "__main__":
    main()
