"""GPU parity of the batched solver and request loop (alps_ed_batch_*)
through the C ABI: every result against the oracle's do_subspace, and each
problem's result bit-identical alone and inside a mix."""
import math
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def xxz_chain(L, periodic, seed, field=0.0):
    rng = np.random.default_rng(seed)
    nb = L if periodic else L - 1
    s = np.arange(nb)
    J = rng.uniform(0.5, 1.5, nb)
    delta = rng.uniform(0, 2)
    return O.Model(L, s, (s + 1) % L, delta * J, J, h=field)


def to_product(m: O.Model, n_up) -> A.SpinModel:
    return A.SpinModel(m.n_sites, m.bond_s, m.bond_t, m.jz, m.jxy, n_up, k=m.k, h=m.h, d=m.d)


def mix(n, seed, Ls=(8, 10, 12, 13, 14), field=True):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        L = int(rng.choice(Ls))
        m = xxz_chain(L, bool(rng.integers(2)), seed * 1000 + i, float(rng.uniform(0, 0.5)) if field else 0.0)
        k = int(rng.integers(0, L + 1))
        out.append((m, k))
    return out


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_matches_oracle(dtype):
    probs = mix(24, 7)
    say("batch", dtype, len(probs))
    b = A.Batch(dtype, hbm_bytes=2 << 30, steps_per_round=6)
    tickets = [b.submit(to_product(m, k), nev=2, seed=3) for m, k in probs]
    b.run()
    assert sorted(b.completed()) == sorted(tickets)
    for t, (m, k) in zip(tickets, probs):
        r = b.result(t)
        ref = O.sparsediag(m, k, nev=2)
        assert r["dim"] == ref["dim"] == math.comb(m.n_sites, k)
        assert r["error_code"] == 0
        np.testing.assert_allclose(r["ev"], ref["ev"], rtol=1e-10, atol=1e-12)
    b.close()


def test_same_alone_and_in_mix():
    probs = mix(40, 11, Ls=(10, 12, 14, 16))
    target = (xxz_chain(14, True, 99, 0.2), 7)
    say("alone")
    b = A.Batch("f32", hbm_bytes=4 << 30, steps_per_round=5)
    t0 = b.submit(to_product(*target), nev=1, seed=5)
    b.run()
    alone = b.result(t0)
    b.close()
    say("in mix")
    b = A.Batch("f32", hbm_bytes=4 << 30, steps_per_round=5)
    ts = [b.submit(to_product(m, k), nev=1, seed=i) for i, (m, k) in enumerate(probs[:20])]
    t1 = b.submit(to_product(*target), nev=1, seed=5)
    # more requests arrive while the first ones run
    b.progress()
    b.progress()
    ts += [b.submit(to_product(m, k), nev=1, seed=100 + i) for i, (m, k) in enumerate(probs[20:])]
    b.run()
    mixed = b.result(t1)
    assert mixed["iterations"] == alone["iterations"]
    assert np.array_equal(mixed["ev"], alone["ev"])      # bit-identical
    ref = O.sparsediag(*target, nev=1)
    np.testing.assert_allclose(mixed["ev"], ref["ev"], rtol=1e-10)
    for t in ts:
        assert b.result(t)["error_code"] == 0
    b.close()


def test_admission_by_footprint():
    # an arena that holds about three of these problems at a time
    probs = [(xxz_chain(14, i % 2 == 0, 200 + i), 7) for i in range(9)]
    fp = A.batch_footprint(to_product(*probs[0]), "f32")
    b = A.Batch("f32", hbm_bytes=int(3.5 * fp), steps_per_round=8)
    ts = [b.submit(to_product(m, k), nev=1, seed=1) for m, k in probs]
    peak = 0
    while True:
        a, p = b.progress()
        peak = max(peak, a)
        st = b.stats()
        assert st["hbm_used"] <= st["hbm_bytes"]
        if a == 0 and p == 0:
            break
    assert 1 <= peak <= 3
    for t, (m, k) in zip(ts, probs):
        np.testing.assert_allclose(b.result(t)["ev"], O.sparsediag(m, k, nev=1)["ev"], rtol=1e-10)
    b.close()


def test_batch_edge_sectors_and_oversize():
    m = xxz_chain(10, True, 5, 0.3)
    b = A.Batch("f32", hbm_bytes=64 << 20)
    t_full = b.submit(to_product(m, 10))     # D = 1
    t_zero = b.submit(to_product(m, 0))      # D = 1 (all down)
    t_two = b.submit(to_product(O.Model(2, np.array([0]), np.array([1]), np.array([1.0]), np.array([1.0])), 1))
    big = xxz_chain(20, True, 6)
    t_big = b.submit(to_product(big, 10))    # never fits in 64 MiB
    b.run()
    for t, k in ((t_full, 10), (t_zero, 0)):
        r = b.result(t)
        ref = O.sparsediag(m, k)
        assert r["dim"] == 1 and r["ev"][0] == ref["ev"][0]   # same summation order: exact
    r2 = b.result(t_two)
    np.testing.assert_allclose(r2["ev"], [-0.75], rtol=1e-14)
    rb = b.result(t_big)
    assert rb["error_code"] == -1 and len(rb["ev"]) == 0
    b.close()


def test_batch_fixed_iterations_and_more_eigenvalues():
    m = xxz_chain(12, True, 42, 0.1)
    b = A.Batch("f64", hbm_bytes=1 << 30)
    t1 = b.submit(to_product(m, 6), max_iter=60, fixed=True, nev=3, seed=0)
    t2 = b.submit(to_product(m, 6), nev=5, seed=0)
    b.run()
    r1, r2 = b.result(t1), b.result(t2)
    ref1 = O.sparsediag(m, 6, max_iter=60, fixed=True, nev=3)
    assert r1["iterations"] == 60
    np.testing.assert_allclose(r1["ev"], ref1["ev"], rtol=1e-9)
    w = np.unique(np.round(np.linalg.eigvalsh(O.dense(m, 6)), 10))
    np.testing.assert_allclose(r2["ev"], w[:5], rtol=1e-10)
    b.close()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_graph_rounds_match_direct_launches(dtype):
    # rounds that repeat the previous round's launches run from a captured
    # HIP graph: replayed many times, with the engine's own host work and
    # synchronisations between replays, the results must equal the
    # directly launched batch bit for bit
    models = mix(12, 21, Ls=(12, 13, 14), field=True)
    out = {}
    for graphs in (False, True):
        say("graph rounds", dtype, graphs)
        b = A.Batch(dtype, steps_per_round=4, graphs=graphs)
        # fixed iteration counts keep the active set unchanged for many rounds
        t = [b.submit(to_product(m, k), max_iter=120, fixed=True, nev=2, seed=i) for i, (m, k) in enumerate(models)]
        t += [b.submit(to_product(m, k), nev=1, seed=7 + i) for i, (m, k) in enumerate(models[:4])]
        b.run()
        out[graphs] = [b.result(x) for x in t]
        st = b.stats()
        b.close()
        if graphs:
            assert st["graph_rounds"] > 10 and st["graph_captures"] >= 1
        else:
            assert st["graph_rounds"] == 0
    for r0, r1 in zip(out[False], out[True]):
        assert r0["iterations"] == r1["iterations"]
        assert np.array_equal(r0["ev"], r1["ev"])
