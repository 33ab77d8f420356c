"""Writes tests/golden/ring24_oracle.json: the oracle's FP64 sparsediag of the
spin-1/2 Heisenberg ring L=24, J=1 (lattice "chain lattice", periodic):

* Sz=0 (n_up=12, D=2,704,156), nlowest Lanczos for the lowest 1, 2 and 4
  distinct levels (sparsediag.h:93-149 semantics: map-of-maps matrix built by
  binary-search lookup, ietl lanczos without reorthogonalization, Tmatrix
  analysis every iteration, rtol = atol = 100 eps);
* Sz=1 (n_up=13, D=2,496,144), lowest level (by SU(2) it is one of the Sz=0
  levels: the lowest triplet).

The oracle is the checker; this script runs it once in the build container
(about one to two minutes and a few GB of memory per sector) and the GPU tests
compare the engine's FP64 and FP32-storage solves with the stored values:
    python tests/golden/make_ring24_fixture.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402


def main():
    om = O.lattice_model("chain lattice", 24)
    out = dict(lattice="chain lattice", L=24, J=1.0, seed=1, runs=[])
    for n_up, nev in ((12, 1), (12, 2), (12, 4), (13, 1)):
        t0 = time.time()
        r = O.sparsediag(om, n_up, nev=nev, seed=1)
        dt = time.time() - t0
        print(f"n_up={n_up} nev={nev}: dim {r['dim']} iterations {r['iterations']} "
              f"error_code {r['error_code']} ev {r['ev'].tolist()} ({dt:.1f} s)", flush=True)
        assert r["error_code"] == 0
        out["runs"].append(dict(n_up=n_up, nev=nev, dim=r["dim"], iterations=r["iterations"],
                                eigenvalues=[float(v) for v in r["ev"]], seconds=round(dt, 1)))
    with open(os.path.join(HERE, "ring24_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
