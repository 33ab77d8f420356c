"""GPU parity at the largest supported sector: N = 32 sites, Sz = 0
(D = C(32,16) = 601,080,390 rows), where no oracle basis fits in a test.

The checks are fixed by the mathematics, not by a second implementation
(spin_formulas: colex ranks, the spin-model row formula and the closed form
of 1^T H 1, all with Python integers).  Couplings and inputs are dyadic
rationals, so every product and sum is exact in FP64 and the comparisons
are bitwise.  The row and sum formulas are themselves checked against the
oracle's dense matrix at N = 10 (CPU)."""
import math
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O
import spin_formulas as F

N, K = 32, 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def model(n=N, k=K, extra=((0, 16), (5, 27), (11, 30), (3, 20)), seed=32):
    """Ring of n sites plus a few long bonds; dyadic couplings and fields."""
    rng = np.random.default_rng(seed)
    s = list(range(n)) + [a for a, _ in extra]
    t = [(i + 1) % n for i in range(n)] + [b for _, b in extra]
    nb = len(s)
    jz = rng.integers(-8, 9, nb) / 8.0
    jxy = rng.integers(1, 9, nb) / 8.0
    h = rng.integers(-4, 5, n) / 8.0
    d = rng.integers(-4, 5, n) / 8.0
    return A.SpinModel(n, np.array(s), np.array(t), jz, jxy, k, h=h, d=d)


def test_formulas_against_oracle_n10():
    n, k = 10, 5
    m = model(n, k, extra=((0, 5), (3, 8)), seed=3)
    om = O.Model(n, m.bond_site1, m.bond_site2, m.jz, m.jxy, h=m.h, d=m.d)
    H = O.dense(om, k)
    D = math.comb(n, k)
    assert H.shape == (D, D)
    x = np.array([F.x_of(r) for r in range(D)])
    y = H @ x
    assert all(y[r] == F.row_value(m, F.colex_unrank(r, k), r) for r in range(D))
    assert [F.colex_rank(F.colex_unrank(r, k), n) for r in range(D)] == list(range(D))
    assert H.sum() == F.closed_form_sum(m)


@pytest.mark.gpu
def test_n32_rows_ranks_and_closed_form_sum():
    m = model()
    op = A.Operator(m)
    D = math.comb(N, K)
    assert op.dim == D == 601080390
    say("operator built", D)

    # ranks and states against the colex formula at sampled rows
    rng = np.random.default_rng(7)
    rows = sorted(set([0, 1, D - 1, D - 2] + [int(v) for v in rng.integers(0, D, 120)]))
    st = np.array([F.colex_unrank(r, K) for r in rows], dtype=np.uint32)
    assert [F.colex_rank(int(v), N) for v in st] == rows
    got = np.concatenate([op.states(r, 1) for r in rows])
    assert np.array_equal(got, st)
    assert np.array_equal(op.index(st), np.array(rows, dtype=np.uint64))
    say("ranks ok")

    # x_r = dyadic function of r, built in chunks
    x = np.empty(D, dtype=np.float64)
    step = 1 << 26
    for a in range(0, D, step):
        b = min(D, a + step)
        x[a:b] = F.x_chunk(a, b)
    X = A.Vector.from_numpy(x, "f64")
    del x
    Y = op.matvec(X, A.Vector(D, "f64"))
    say("matvec f64 done")

    # 1^T H 1 from the closed form (exact: every term is a multiple of 2^-6)
    ones = A.Vector(D, "f64")
    ones.fill(1.0)
    H1 = op.matvec(ones, A.Vector(D, "f64"))
    want = F.closed_form_sum(m)
    assert ones.dot(H1) == want
    del H1, ones
    say("closed-form sum ok", want)

    y = Y.numpy()
    for r, s in zip(rows, st):
        assert y[r] == F.row_value(m, int(s), r), r
    say("f64 rows ok")
    del y, Y

    # FP32 storage: the same rows, each rounded once from the exact FP64 sum
    X32 = A.Vector(D, "f32")
    X32.copy_from(X)
    del X
    Y32 = op.matvec(X32, A.Vector(D, "f32"))
    y32 = Y32.numpy()
    for r, s in zip(rows, st):
        assert y32[r] == np.float32(F.row_value(m, int(s), r)), r
    say("f32 rows ok")
