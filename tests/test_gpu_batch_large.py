"""GPU parity of the batched solver at the sizes of the batched and served
workloads (random-bond XXZ chains, L = 18..20 at Sz = 0 or 1/2, and L = 22..26
with a uniform field in a nonzero Sz sector, up to D = 5,311,735), against
the oracle's FP64 sparsediag stored in tests/golden/xxz_large_oracle.json
(tests/golden/make_xxz_large_fixture.py).  FP32 and FP64 storage, all
problems in one batch, and the largest one also through the one-at-a-time
solve.  Bar: the lowest two levels within 1e-10 relative."""
import json
import math
import os
import sys

import numpy as np
import pytest

import alps_amd as A

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "xxz_large_oracle.json")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problems():
    with open(FIXTURE) as f:
        return json.load(f)["problems"]


def product_model(p):
    L = p["L"]
    s = np.arange(len(p["jxy"]))
    return A.SpinModel(L, s, (s + 1) % L, np.array(p["jz"]), np.array(p["jxy"]), p["n_up"],
                       h=np.full(L, p["h"]))


def test_fixture_shapes():
    ps = problems()
    assert len(ps) == 7
    for p in ps:
        assert p["dim"] == math.comb(p["L"], p["n_up"])
        assert len(p["jxy"]) == (p["L"] if p["periodic"] else p["L"] - 1)
        assert len(p["eigenvalues"]) == 2 and p["eigenvalues"][0] <= p["eigenvalues"][1]
    assert max(p["dim"] for p in ps) == 5311735


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_batch_large_against_oracle(dtype):
    ps = problems()
    say("batch", dtype, [p["dim"] for p in ps])
    b = A.Batch(dtype, hbm_bytes=24 << 30, steps_per_round=8)
    tickets = [b.submit(product_model(p), nev=2, seed=3) for p in ps]
    b.run()
    for t, p in zip(tickets, ps):
        r = b.result(t)
        say(p["L"], p["n_up"], r["iterations"], r["ev"].tolist(), p["eigenvalues"])
        assert r["error_code"] == 0 and r["dim"] == p["dim"]
        np.testing.assert_allclose(r["ev"], p["eigenvalues"], rtol=1e-10, atol=0)
    b.close()


def test_largest_one_at_a_time_f32():
    p = max(problems(), key=lambda q: q["dim"])
    say("sparsediag", p["L"], p["n_up"], p["dim"])
    r = A.sparsediag(product_model(p), nev=2, seed=3, dtype="f32")
    assert r["error_code"] == 0
    np.testing.assert_allclose(r["ev"], p["eigenvalues"], rtol=1e-10, atol=0)
