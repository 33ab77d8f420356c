"""GPU parity at the largest supported sector: N = 32 sites, Sz = 0
(D = C(32,16) = 601,080,390 rows), where no oracle basis fits in a test.

The checks are fixed by the mathematics, not by a second implementation:
  - the rank of a state is its colex rank sum_i C(p_i, i) over the
    positions p_1 < ... < p_k of its set bits (the reference's
    lexicographic basis order, basisstates.h:176-278, is ascending integer
    order), computed here with Python integers;
  - a row of H x for the spin model (models.xml.in:342-373) is the
    diagonal sum_i (-h_i Sz_i + d_i Sz_i^2) + sum_b Jz_b Sz Sz times x_r plus
    Jxy_b / 2 times x at the rank of the flipped state for each
    antiparallel bond;
  - 1^T H 1 has a closed form: per bond, C(N-2,k-2) + C(N-2,k) parallel and
    2 C(N-2,k-1) antiparallel states; per site, C(N-1,k-1) up and C(N-1,k)
    down states.
Couplings and inputs are dyadic rationals, so every product and sum is
exact in FP64 and the comparisons are bitwise.  The row and sum formulas
are themselves checked against the oracle's dense matrix at N = 10 (CPU)."""
import math
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

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


def colex_rank(s, n=N):
    r, i = 0, 0
    for p in range(n):
        if s >> p & 1:
            i += 1
            r += math.comb(p, i)
    return r


def colex_unrank(r, k=K):
    s = 0
    for i in range(k, 0, -1):
        p = i - 1
        while math.comb(p + 1, i) <= r:
            p += 1
        r -= math.comb(p, i)
        s |= 1 << p
    return s


def x_of(r):
    """Input vector: dyadic values in [-1/2, 1/2) from the row index."""
    return float(int(r) % 1021 - 510) / 1024.0


def row_value(m, s, r):
    N = m.n_sites
    bit = lambda i: (s >> (N - 1 - i)) & 1
    sz = lambda i: 0.5 if bit(i) else -0.5
    diag = sum(-m.h[i] * sz(i) + m.d[i] * 0.25 for i in range(N))
    off = 0.0
    for b in range(len(m.bond_site1)):
        i, j = int(m.bond_site1[b]), int(m.bond_site2[b])
        diag += m.jz[b] * sz(i) * sz(j)
        if bit(i) != bit(j):
            t = s ^ (1 << (N - 1 - i)) ^ (1 << (N - 1 - j))
            off += 0.5 * m.jxy[b] * x_of(colex_rank(t, N))
    return diag * x_of(r) + off


def closed_form_sum(m):
    """1^T H 1 of the sector."""
    c, n, k = math.comb, m.n_sites, m.n_up
    par, anti = c(n - 2, k - 2) + c(n - 2, k), 2 * c(n - 2, k - 1)
    want = sum(m.jz[b] / 4 * (par - anti) + m.jxy[b] / 2 * anti for b in range(len(m.jz)))
    want += sum(-m.h[i] * 0.5 * (c(n - 1, k - 1) - c(n - 1, k)) + m.d[i] / 4 * c(n, k) for i in range(n))
    return want


def test_formulas_against_oracle_n10():
    n, k = 10, 5
    m = model(n, k, extra=((0, 5), (3, 8)), seed=3)
    om = O.Model(n, m.bond_site1, m.bond_site2, m.jz, m.jxy, h=m.h, d=m.d)
    H = O.dense(om, k)
    D = math.comb(n, k)
    assert H.shape == (D, D)
    x = np.array([x_of(r) for r in range(D)])
    y = H @ x
    assert all(y[r] == row_value(m, colex_unrank(r, k), r) for r in range(D))
    assert [colex_rank(colex_unrank(r, k), n) for r in range(D)] == list(range(D))
    assert H.sum() == closed_form_sum(m)


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
    st = np.array([colex_unrank(r) for r in rows], dtype=np.uint32)
    assert [colex_rank(int(v)) for v in st] == rows
    got = np.concatenate([op.states(r, 1) for r in rows])
    assert np.array_equal(got, st)
    assert np.array_equal(op.index(st), np.array(rows, dtype=np.uint64))
    say("ranks ok")

    # x_r = dyadic function of r, built in chunks
    x = np.empty(D, dtype=np.float64)
    step = 1 << 26
    for a in range(0, D, step):
        r = np.arange(a, min(D, a + step), dtype=np.int64)
        x[a:a + r.size] = (r % 1021 - 510) / 1024.0
    X = A.Vector.from_numpy(x, "f64")
    del x
    Y = op.matvec(X, A.Vector(D, "f64"))
    say("matvec f64 done")

    # 1^T H 1 from the closed form (exact: every term is a multiple of 2^-6)
    ones = A.Vector(D, "f64")
    ones.fill(1.0)
    H1 = op.matvec(ones, A.Vector(D, "f64"))
    want = closed_form_sum(m)
    assert ones.dot(H1) == want
    del H1, ones
    say("closed-form sum ok", want)

    y = Y.numpy()
    for r, s in zip(rows, st):
        assert y[r] == row_value(m, int(s), r), r
    say("f64 rows ok")
    del y, Y

    # FP32 storage: the same rows, each rounded once from the exact FP64 sum
    X32 = A.Vector(D, "f32")
    X32.copy_from(X)
    del X
    Y32 = op.matvec(X32, A.Vector(D, "f32"))
    y32 = Y32.numpy()
    for r, s in zip(rows, st):
        assert y32[r] == np.float32(row_value(m, int(s), r)), r
    say("f32 rows ok")
