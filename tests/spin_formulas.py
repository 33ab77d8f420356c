"""Closed forms of the spin-1/2 fixed-Sz sector used by the full-size GPU
tests, where no oracle basis fits in a test.  Everything here is fixed by
the mathematics, computed with Python integers:

  - the rank of a state is its colex rank sum_i C(p_i, i) over the
    positions p_1 < ... < p_k of its set bits (the reference's
    lexicographic basis order, basisstates.h:176-278, is ascending integer
    order);
  - a row of H x for the spin model (models.xml.in:342-373) is the
    diagonal sum_i (-h_i Sz_i + d_i Sz_i^2) + sum_b Jz_b Sz Sz times x_r plus
    Jxy_b / 2 times x at the rank of the flipped state for each
    antiparallel bond;
  - 1^T H 1 has a closed form: per bond, C(N-2,k-2) + C(N-2,k) parallel and
    2 C(N-2,k-1) antiparallel states; per site, C(N-1,k-1) up and C(N-1,k)
    down states.

Site s is bit N-1-s of the state word (up = 1).  test_gpu_maxsize checks
these formulas against the oracle's dense matrix at N = 10."""
import math


def colex_rank(s, n):
    r, i = 0, 0
    for p in range(n):
        if s >> p & 1:
            i += 1
            r += math.comb(p, i)
    return r


def colex_unrank(r, k):
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


def x_chunk(a, b):
    """x_of over rows [a, b) as a numpy array."""
    import numpy as np

    r = np.arange(a, b, dtype=np.int64)
    return (r % 1021 - 510) / 1024.0


def row_value(m, s, r):
    """(H x)_r for the state s of row r and x = x_of."""
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
