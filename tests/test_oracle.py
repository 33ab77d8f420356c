"""Pins of the CPU oracle (oracle/ed_oracle.hpp) against the reference's own
golden files, LAPACK dstev fixtures, closed forms and independent
constructions.  Each test names the oracle function it pins and the
mistake it would catch."""
import json
import math
import os
import re

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------------------
# basis order (build_basis, state_bits) -- basisstates.h:176-278
# ---------------------------------------------------------------------------
def _parse_states(path):
    rows = []
    for line in open(path):
        line = line.strip()
        if line.startswith("["):
            sz = re.findall(r"\|1/2 (-?1/2) >", line)
            rows.append([1 if s == "1/2" else 0 for s in sz])
    return rows


def _bits(row):
    n = len(row)
    return sum(1 << (n - 1 - s) for s, up in enumerate(row) if up)


def test_basis_order_fixed_sz_example12():
    # test/model/example12: square lattice L=2 (4 sites), Sz_total=0
    rows = _parse_states(os.path.join(GOLDEN, "model_example12.output"))
    assert len(rows) == 6
    assert [_bits(r) for r in rows] == list(O.basis(4, 2))


def test_basis_order_full_example10():
    # test/model/example10: no constraint, all 16 states; restricted to each
    # popcount the order must be the oracle's fixed-Sz order
    rows = _parse_states(os.path.join(GOLDEN, "model_example10.output"))
    bits = [_bits(r) for r in rows]
    assert bits == sorted(bits) == list(range(16))
    for k in range(5):
        assert [b for b in bits if bin(b).count("1") == k] == list(O.basis(4, k))


def test_basis_rank_is_colex():
    # the index of a state is sum_m C(p_m, m) over its set bits (ascending)
    for N, k in [(6, 3), (8, 2), (9, 5), (10, 5)]:
        states = O.basis(N, k)
        assert len(states) == math.comb(N, k)
        for i, s in enumerate(states):
            p = [b for b in range(N) if s >> b & 1]
            assert sum(math.comb(pm, m + 1) for m, pm in enumerate(p)) == i


# ---------------------------------------------------------------------------
# local matrix elements (site_element, bond_element) -- example3 / example4
# ---------------------------------------------------------------------------
def _eval_block(text, name, env):
    m = re.search(name + r" =\s*\n(.*?);", text, re.S)
    expr = m.group(1).replace("{", "[").replace("}", "]")
    return np.array(eval(expr, {"__builtins__": {}}, env), dtype=float)


def _two_site_model(env):
    return O.Model(2, [0], [1], env["Jz0"], env["Jxy0"], k=env["K0"], h=env["h0"], d=env["D0"])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_two_site_elements_example3_example4(seed):
    rng = np.random.default_rng(seed)
    env = {n: float(rng.uniform(-2, 2)) for n in ["Jz0", "Jxy0", "K0", "h0", "D0"]}
    env["Gamma0"] = 0.0  # Sz-conserving model
    t3 = open(os.path.join(GOLDEN, "model_example3.output")).read()
    t4 = open(os.path.join(GOLDEN, "model_example4.output")).read()
    site = _eval_block(t3, "HSpinSite", env)          # [is][js]
    bond = _eval_block(t3, "HSpinBond", env)          # [i1][i2][j1][j2]
    full4 = _eval_block(t4, "HSpinHalf", env)         # two sites, one bond
    # example3's local matrices assembled for a dimer must equal example4
    asm = np.zeros((4, 4))
    for i1 in range(2):
        for i2 in range(2):
            for j1 in range(2):
                for j2 in range(2):
                    v = bond[i1, i2, j1, j2]
                    if i2 == j2:
                        v += site[i1, j1]
                    if i1 == j1:
                        v += site[i2, j2]
                    asm[2 * i1 + i2, 2 * j1 + j2] = v
    np.testing.assert_allclose(asm, full4, rtol=1e-12, atol=1e-12)
    # the oracle's sector matrices are the diagonal blocks of HSpinHalf
    m = _two_site_model(env)
    for k, idx in [(0, [0]), (1, [1, 2]), (2, [3])]:
        A = O.dense(m, k)
        np.testing.assert_allclose(A, full4[np.ix_(idx, idx)], rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------------------
# lattice graphs (named_lattice) -- test/lattice/example6
# ---------------------------------------------------------------------------
def test_lattice_bonds_example6():
    want = [tuple(map(int, m)) for m in re.findall(r"between site (\d+) and site (\d+) has type (\d+)",
                                                    open(os.path.join(GOLDEN, "lattice_example6.output")).read())]
    N, s, t, ty = O.lattice("anisotropic square lattice", 10, 4)
    assert N == 40
    assert list(zip(s.tolist(), t.tolist(), ty.tolist())) == want


def test_lattice_shapes():
    # bond counts and boundary handling of the one-vertex unit cells
    assert O.lattice("chain lattice", 8)[1].size == 8
    assert O.lattice("open chain lattice", 8)[1].size == 7
    assert O.lattice("chain lattice", 1)[1].size == 0          # source == target dropped
    N, s, t, ty = O.lattice("frustrated square lattice", 6, 5)
    assert N == 30 and s.size == 120 and (ty == 0).sum() == 60 and (ty == 1).sum() == 60
    # site index x*W + y; the first cell's four edges in unit-cell order
    assert list(zip(s[:4], t[:4], ty[:4])) == [(0, 1, 0), (0, 5, 0), (0, 6, 1), (0, 9, 1)]
    N, s, t, ty = O.lattice("ladder", 6)
    assert N == 12 and s.size == 12 + 6                          # W defaults to 2, rungs open
    N, s, t, ty = O.lattice("open square lattice", 3, 4)
    assert N == 12 and s.size == 2 * 12 - 3 - 4


# ---------------------------------------------------------------------------
# tridiagonal eigensolver (tridiag_eig) -- LAPACK dstev fixtures
# ---------------------------------------------------------------------------
def test_tridiag_against_dstev_fixtures():
    cases = json.load(open(os.path.join(GOLDEN, "dstev_cases.json")))
    for c in cases:
        d, e = np.array(c["d"]), np.array(c["e"])
        w_ref = np.array(c["eigenvalues"])
        w, z = O.tridiag(d, e, vectors=True)
        scale = max(1.0, np.abs(w_ref).max())
        np.testing.assert_allclose(w, w_ref, atol=1e-13 * scale, rtol=0, err_msg=c["label"])
        # eigenvectors are orthonormal and satisfy T z = w z
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        np.testing.assert_allclose(z.T @ z, np.eye(len(d)), atol=1e-12)
        np.testing.assert_allclose(T @ z, z * w, atol=1e-12 * scale)
        gaps = np.diff(w_ref)
        if len(gaps) == 0 or gaps.min() > 1e-8:  # last row unique up to sign
            np.testing.assert_allclose(np.abs(z[-1]), c["zlast_abs"], atol=1e-10, err_msg=c["label"])


def test_tridiag_toeplitz_closed_form():
    n = 30
    w = O.tridiag(np.full(n, 0.25), np.full(n - 1, -0.75))
    k = np.arange(1, n + 1)
    np.testing.assert_allclose(w, np.sort(0.25 - 1.5 * np.cos(k * np.pi / (n + 1))), atol=1e-14)


# ---------------------------------------------------------------------------
# matrix build (build_matrix) -- independent constructions
# ---------------------------------------------------------------------------
_SZ = np.diag([-0.5, 0.5])      # local index 0 = Sz -1/2
_SP = np.array([[0, 0], [1, 0.]])  # S+ : |0> -> |1>
_SM = _SP.T


def _site_op(op, s, N):
    mats = [np.eye(2)] * N
    mats[s] = op
    out = mats[0]
    for m in mats[1:]:
        out = np.kron(out, m)
    return out


def kron_hamiltonian(m: O.Model):
    """Dense H as a sum of Kronecker products of local operators (site 0 the
    most significant factor), then restricted per sector."""
    N = m.n_sites
    H = np.zeros((2 ** N, 2 ** N))
    for s in range(N):
        H += -m.h[s] * _site_op(_SZ, s, N) + m.d[s] * _site_op(_SZ @ _SZ, s, N)
    for b in range(m.n_bonds):
        i, j = m.bond_s[b], m.bond_t[b]
        SzSz = _site_op(_SZ, i, N) @ _site_op(_SZ, j, N)
        ex = 0.5 * (_site_op(_SP, i, N) @ _site_op(_SM, j, N) + _site_op(_SM, i, N) @ _site_op(_SP, j, N))
        SS = SzSz + ex
        H += m.jz[b] * SzSz + m.jxy[b] * ex + m.k[b] * (SS @ SS)
    return H


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_matrix_matches_kronecker_construction(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(4, 9))
    nb = int(rng.integers(N, 2 * N + 1))
    s = rng.integers(0, N, nb)
    t = (s + rng.integers(1, N, nb)) % N      # distinct ends, repeats allowed
    m = O.Model(N, s, t, rng.uniform(-1.5, 1.5, nb), rng.uniform(-1.5, 1.5, nb), k=rng.uniform(-1, 1, nb),
                h=rng.uniform(-1, 1, N), d=rng.uniform(-1, 1, N))
    H = kron_hamiltonian(m)
    for k in range(N + 1):
        idx = O.basis(N, k).astype(np.int64)
        np.testing.assert_allclose(O.dense(m, k), H[np.ix_(idx, idx)], atol=1e-13, rtol=0)


def test_matrix_ising_limit_bruteforce():
    # Jxy = K = 0: H is diagonal with E(s) = sum_b Jz/4 (+-1) - sum_s h Sz
    rng = np.random.default_rng(7)
    N, s, t, _ = O.lattice("open square lattice", 3, 3)
    jz = rng.uniform(-1, 1, s.size)
    h = rng.uniform(-0.5, 0.5, N)
    m = O.Model(N, s, t, jz, 0.0, h=h)
    for k in (3, 4, 5):
        st = O.basis(N, k)
        up = ((st[:, None] >> (N - 1 - np.arange(N))[None, :]) & 1).astype(float)
        sz = up - 0.5
        E = (jz[None, :] * sz[:, s] * sz[:, t]).sum(1) - (h[None, :] * sz).sum(1)
        A = O.dense(m, k)
        np.testing.assert_allclose(np.diag(A), E, atol=1e-14)
        assert np.count_nonzero(A - np.diag(np.diag(A))) == 0


def _free_fermion_e0(L, J, M, periodic, h=0.0):
    T = np.zeros((L, L))
    for i in range(L - 1):
        T[i, i + 1] = T[i + 1, i] = J[i] / 2
    if periodic and L > 2:
        sgn = (-1) ** (M - 1)  # Jordan-Wigner boundary sign
        T[L - 1, 0] = T[0, L - 1] = sgn * J[L - 1] / 2
    eps = np.linalg.eigvalsh(T)
    return eps[:M].sum() - h * (M - L / 2)


@pytest.mark.parametrize("periodic", [False, True])
def test_xx_chain_free_fermions(periodic):
    rng = np.random.default_rng(11 + periodic)
    for L in (6, 8, 10, 11):
        J = rng.uniform(0.5, 1.5, L)
        h = float(rng.uniform(0, 0.5))
        nb = L if periodic else L - 1
        s = np.arange(nb)
        t = (s + 1) % L
        m = O.Model(L, s, t, 0.0, J[:nb], h=h)
        for M in range(0, L + 1):
            e0 = np.linalg.eigvalsh(O.dense(m, M))[0]
            assert abs(e0 - _free_fermion_e0(L, J, M, periodic, h)) < 1e-11, (L, M)


# ---------------------------------------------------------------------------
# Lanczos + Tmatrix (Lanczos, Tmatrix, Iteration, do_subspace)
# ---------------------------------------------------------------------------
def _distinct(w, tol=1e-9):
    out = [w[0]]
    for x in w[1:]:
        if x - out[-1] > tol:
            out.append(x)
    return np.array(out)


@pytest.mark.parametrize("name,L,W,k", [("chain lattice", 12, 0, 6), ("chain lattice", 11, 0, 5),
                                        ("square lattice", 3, 4, 6), ("frustrated square lattice", 4, 3, 6),
                                        ("ladder", 6, 0, 6)])
def test_lanczos_nlowest_matches_dense(name, L, W, k):
    m = O.lattice_model(name, L, W, by_type={1: (0.5, 0.5)} if "frustrated" in name else None)
    w = _distinct(np.linalg.eigvalsh(O.dense(m, k)))
    r = O.sparsediag(m, k, nev=4, seed=3)
    assert r["error_code"] == 0
    np.testing.assert_allclose(r["ev"], w[:4], rtol=1e-10, atol=1e-10)


def test_lanczos_ring_of_four():
    # Heisenberg ring L=4, J=1: ground state -2J (singlet), then -1 (triplet)
    r = O.sparsediag(O.lattice_model("chain lattice", 4), 2, nev=2)
    np.testing.assert_allclose(r["ev"], [-2.0, -1.0], atol=1e-14)


def _gap_by_sorting(sectors):
    # the gap read from the multiset of every sector's lowest value plus the
    # second value of the sector holding the global minimum (its first
    # occurrence): sorted, second minus first
    firsts = [v[0] for v in sectors if len(v)]
    g = min(firsts)
    ground_sector = next(v for v in sectors if len(v) and v[0] == g)
    pool = sorted(firsts + list(ground_sector[1:2]))
    return g, (pool[1] - pool[0]) if len(pool) > 1 else np.finfo(float).max


def test_ground_and_gap_ring_of_four():
    # Heisenberg ring L=4: singlet -2 in Sz=0, triplet -1 in Sz=0,+-1,
    # quintet +1 in Sz=+-2: ground -2, gap 1 (closed form)
    m = O.lattice_model("chain lattice", 4)
    sectors = [O.sparsediag(m, k, nev=2)["ev"] for k in range(5)]
    g, gap = O.ground_and_gap(sectors)
    assert abs(g + 2.0) < 1e-13 and abs(gap - 1.0) < 1e-13


def test_ground_and_gap_cases():
    rng = np.random.default_rng(11)
    cases = [[[0.0, 1.0]], [[-1.0]], [[], [2.0, 3.0], []], [[1.0, 5.0], [1.0, 2.0]],   # degenerate: gap 0
             [[-3.0, -2.5], [-2.9], [-2.0, -1.0]], [[-2.9], [-3.0, -2.5]], [[-3.0, -1.0], [-2.0]]]
    for _ in range(200):
        cases.append([sorted(rng.integers(-4, 4, rng.integers(0, 3)).astype(float).tolist())
                      for _ in range(rng.integers(1, 5))])
    for sec in cases:
        if not any(len(v) for v in sec):
            continue
        g, gap = O.ground_and_gap(sec)
        g2, gap2 = _gap_by_sorting(sec)
        assert g == g2 and gap == gap2, sec


def bethe_ground_energy(L, M):
    """Lowest energy of the spin-1/2 Heisenberg ring (J=1) with M down spins
    from the Bethe ansatz: 2 L arctan(2 l_j) = 2 pi I_j + 2 sum_k arctan(l_j -
    l_k) with I_j = -(M-1)/2, ..., (M-1)/2, and E = L/4 - sum_j 2 / (4 l_j^2 + 1)
    (magnon energy cos k - 1 with e^{ik} = (l + i/2) / (l - i/2))."""
    from scipy.optimize import fsolve

    I = np.arange(M) - (M - 1) / 2.0

    def f(lam):
        return 2 * L * np.arctan(2 * lam) - 2 * np.pi * I - 2 * np.arctan(lam[:, None] - lam[None, :]).sum(axis=1)

    lam = fsolve(f, 0.5 * np.tan(np.pi * I / L), xtol=1e-15)
    assert np.abs(f(lam)).max() < 1e-12
    return L / 4 - np.sum(2 / (4 * lam ** 2 + 1))


def test_ring24_fixture_matches_bethe_ansatz():
    """The stored oracle solves of the L=24 ring (the reference values of the
    GPU tests at that size) against the Bethe ansatz: the Sz=0 ground state
    (12 down spins), the lowest Sz=1 level (11 down spins), which SU(2) makes
    the second Sz=0 level; the Bethe form itself gives the ring of four's
    E0 = -2 exactly."""
    assert abs(bethe_ground_energy(4, 2) + 2.0) < 1e-14
    e0, e1 = bethe_ground_energy(24, 12), bethe_ground_energy(24, 11)
    for nev in (1, 2, 4):
        ev = O.ring24_fixture(12, nev)["eigenvalues"]
        assert len(ev) == nev
        assert abs(ev[0] - e0) < 1e-12 * abs(e0)
        if nev > 1:
            assert abs(ev[1] - e1) < 1e-12 * abs(e1)
            assert np.all(np.diff(ev) > 1e-3)
    assert abs(O.ring24_fixture(13, 1)["eigenvalues"][0] - e1) < 1e-12 * abs(e1)


def test_lanczos_three_term_relation():
    # alpha/beta from the recurrence reproduce H's spectrum extremes and the
    # first step's alpha is the Rayleigh quotient of the start vector
    m = O.lattice_model("chain lattice", 10)
    A = O.dense(m, 5)
    r = O.sparsediag(m, 5, max_iter=40, fixed=True, seed=5)
    D = A.shape[0]
    v0 = np.array([O.start_value(5, i) for i in range(D)])
    v0 /= np.linalg.norm(v0)
    assert abs(r["alpha"][0] - v0 @ A @ v0) < 1e-12
    w1 = A @ v0 - r["alpha"][0] * v0
    assert abs(r["beta"][0] - np.linalg.norm(w1)) < 1e-12
    assert r["iterations"] == 40 and len(r["alpha"]) == 40 and r["error_code"] == 1


def test_start_vector_splitmix64_known_answer():
    # first output of splitmix64 from state 0 is 0xe220a8397b1dcdaf
    z = 0xE220A8397B1DCDAF
    u = (z >> 11) * 2.0 ** -53
    assert O.start_value(0, 0, 0.0, 1.0) == u
    assert O.start_value(0, 0, -1.0, 1.0) == -1.0 + 2.0 * u


def _sym2(a, b, c):
    # eigenpairs of [[a, b], [b, c]]
    w, v = np.linalg.eigh(np.array([[a, b], [b, c]]))
    return w, v


def test_tmatrix_two_by_two_closed_form():
    a = [0.3, -1.1]
    b = [0.7, 0.25]
    w, v = _sym2(a[0], b[0], a[1])
    t = O.tmatrix(a, b)
    np.testing.assert_allclose(t["eig"], w, atol=1e-15)
    np.testing.assert_allclose(t["err"], np.abs(b[1] * v[1, :]), atol=1e-15)
    assert list(t["mult"]) == [1, 1]


def test_tmatrix_ghost_scan_bound():
    # T = [a0] (+) 2x2 block, coupled by a negligible beta0: the lowest
    # eigenvalue of T minus its first row/column coincides with the block's
    # lower eigenvalue, which is flagged a ghost; the upper one is outside
    # the scan (j < n-2) and survives.
    a = [-3.0, 0.5, 1.5]
    b = [1e-30, 0.4, 0.3]
    wb, vb = _sym2(a[1], b[1], a[2])
    t = O.tmatrix(a, b)
    np.testing.assert_allclose(t["eig"], [a[0], wb[1]], atol=1e-15)
    np.testing.assert_allclose(t["err"], [0.0, abs(b[2] * vb[1, 1])], atol=1e-15)


def test_tmatrix_signed_threshold():
    # all-negative spectrum: thold = 5*multol (signed max), so two values
    # 1e-10 apart stay distinct; |lambda| * eps^(2/3) would merge them
    a = [-5.0, -5.0 + 1e-10]
    b = [1e-30, 0.5]
    t = O.tmatrix(a, b)
    assert len(t["eig"]) == 2 and list(t["mult"]) == [1, 1]
    multol = 0.5 * 2 * np.finfo(float).eps * (10000 + 2)
    assert abs(t["multol"] - multol) < 1e-20
    assert abs(t["thold"] - 5 * multol) < 1e-20


def test_sparsediag_edge_dimensions():
    m = O.lattice_model("chain lattice", 6, h=0.3)
    r = O.sparsediag(m, 0)           # D = 1: H(0,0), no Lanczos
    assert r["dim"] == 1 and r["iterations"] == 0
    assert r["ev"][0] == 1.5 + 6 * 0.15   # 6 parallel bonds Jz/4, 6 sites -h*(-1/2)
    r = O.sparsediag(m, 7)           # D = 0: nothing
    assert r["dim"] == 0 and len(r["ev"]) == 0


# ---------------------------------------------------------------------------
# eigenvector measurements (expect_site_term): Kronecker products, the ring
# of four and the open XX chain as free fermions
# ---------------------------------------------------------------------------
_LOCAL = {"Sz": _SZ, "Splus": _SP, "Sminus": _SM, "Sx": 0.5 * (_SP + _SM)}


def _embed(N, k, v):
    full = np.zeros(2 ** N)
    full[O.basis(N, k).astype(np.int64)] = v
    return full


@pytest.mark.parametrize("seed", [0, 1])
def test_measurement_terms_match_kronecker_products(seed):
    rng = np.random.default_rng(100 + seed)
    N, k = 7, 3 + seed
    D = len(O.basis(N, k))
    v = rng.uniform(-1, 1, D)
    f = _embed(N, k, v)
    for a in _LOCAL:
        for i in range(N):
            want = f @ _site_op(_LOCAL[a], i, N) @ f
            assert abs(O.expect_site_term(N, k, v, a, i) - want) < 1e-13
        for b in _LOCAL:
            for i in range(N):
                for j in range(N):
                    op = (_site_op(_LOCAL[a] @ _LOCAL[b], i, N) if i == j
                          else _site_op(_LOCAL[a], i, N) @ _site_op(_LOCAL[b], j, N))
                    got = O.expect_site_term(N, k, v, a, i, b, j)
                    assert abs(got - f @ op @ f) < 1e-13, (a, b, i, j)


def test_measurement_ring_of_four_singlet():
    # ground state of the J=1 ring of four: E0 = -2, a singlet, so
    # <S_i.S_j> = -1/2 (neighbours), +1/4 (across); <SzSz> = <S.S>/3,
    # <S+_i S-_j> = 2<S.S>/3
    N, s, t, _ = O.lattice("chain lattice", 4)
    m = O.Model(N, s, t, 1.0, 1.0)
    w, V = np.linalg.eigh(O.dense(m, 2))
    assert abs(w[0] + 2) < 1e-12 and w[1] - w[0] > 0.5
    v = V[:, 0]
    for i in range(4):
        for d, ss in ((1, -0.5), (2, 0.25), (3, -0.5)):
            j = (i + d) % 4
            assert abs(O.expect_site_term(4, 2, v, "Sz", i, "Sz", j) - ss / 3) < 1e-13
            assert abs(O.expect_site_term(4, 2, v, "Splus", i, "Sminus", j) - 2 * ss / 3) < 1e-13
            assert abs(O.expect_site_term(4, 2, v, "Sx", i, "Sx", j) - ss / 3) < 1e-13
        assert abs(O.expect_site_term(4, 2, v, "Sz", i)) < 1e-13
        assert abs(O.expect_site_term(4, 2, v, "Splus", i, "Sminus", i) - 0.5) < 1e-13  # n_up = 1/2
        assert abs(O.expect_site_term(4, 2, v, "Sz", i, "Sz", i) - 0.25) < 1e-13


def test_measurement_open_xx_chain_free_fermions():
    # Jordan-Wigner: the open XX chain (H = sum J/2 (S+S- + h.c.)) is free
    # fermions with hopping J/2; in the ground state with M particles
    # G_ij = sum_{occupied} phi(i) phi(j), <n_i> = G_ii,
    # <Sz_i Sz_j> = (G_ii - 1/2)(G_jj - 1/2) - G_ij^2 (i != j, Wick),
    # <S+_i S-_{i+1}> = <c+_i c_{i+1}> = G_{i,i+1} (no string between neighbours).
    rng = np.random.default_rng(5)
    L, M = 10, 4
    J = rng.uniform(0.5, 1.5, L - 1)
    s = np.arange(L - 1)
    m = O.Model(L, s, s + 1, 0.0, J)
    T = np.zeros((L, L))
    for i in range(L - 1):
        T[i, i + 1] = T[i + 1, i] = J[i] / 2
    eps, phi = np.linalg.eigh(T)
    assert eps[M] - eps[M - 1] > 1e-3  # non-degenerate Fermi level
    G = phi[:, :M] @ phi[:, :M].T
    w, V = np.linalg.eigh(O.dense(m, M))
    assert abs(w[0] - eps[:M].sum()) < 1e-11 and w[1] - w[0] > 1e-3
    v = V[:, 0]
    for i in range(L):
        assert abs(O.expect_site_term(L, M, v, "Sz", i) - (G[i, i] - 0.5)) < 1e-11
        for j in range(L):
            if i != j:
                want = (G[i, i] - 0.5) * (G[j, j] - 0.5) - G[i, j] ** 2
                assert abs(O.expect_site_term(L, M, v, "Sz", i, "Sz", j) - want) < 1e-11, (i, j)
        if i + 1 < L:
            assert abs(O.expect_site_term(L, M, v, "Splus", i, "Sminus", i + 1) - G[i, i + 1]) < 1e-11
            assert abs(O.expect_site_term(L, M, v, "Sminus", i, "Splus", i + 1) - G[i, i + 1]) < 1e-11


# ---------------------------------------------------------------------------
# momentum sectors: translations and Bloch basis against the reference's
# test/model/example13.output (4-site ring) and example15.output (2x2 square,
# translation_momenta); Hamiltonians Hermitian and their spectra, over all
# momenta, the spectrum of the Sz sector
# ---------------------------------------------------------------------------
def _parse_bloch_fixture(path):
    blocks, cur = [], None
    for line in open(path):
        line = line.strip()
        if line.startswith("Momentum:"):
            cur = {"k": [float(x) for x in line.split()[1:]], "trans": [], "states": []}
            blocks.append(cur)
        elif line.startswith("Translation"):
            if cur is None or (cur["states"] and line.startswith("Translation 0 ")):
                cur = {"k": None, "trans": [], "states": []}
                blocks.append(cur)
            m = re.match(r"Translation (\d+) with phase \(([-\d.e]+),([-\d.e]+)\) maps (.*)$", line)
            maps = [int(x.split("->")[1]) for x in m.group(4).split()]
            cur["trans"].append((complex(float(m.group(2)), float(m.group(3))), maps))
        elif "] + translations" in line:
            norm = float(line.split("*[")[0])
            spins = re.findall(r"\|1/2 (-?1/2) >", line)
            bits = "".join("1" if s == "1/2" else "0" for s in spins)
            cur["states"].append((int(bits, 2), norm))
    return blocks


def _check_bloch_block(name, L, W, k, block):
    N = L * (W if W else 1)
    ph, pm = O.translations(name, L, W, k=k)
    assert len(ph) == len(block["trans"])
    for (want_ph, want_map), p, m in zip(block["trans"], ph, pm):
        assert abs(round(p.real, 1) - want_ph.real) < 1e-12 and abs(round(p.imag, 1) - want_ph.imag) < 1e-12
        assert list(m) == want_map
    reps, norm = O.bloch_basis(N, -1, ph, pm)
    assert [int(r) for r in reps] == [s for s, _ in block["states"]]
    for got, (_, want) in zip(norm, block["states"]):
        assert abs(got - want) <= 1e-5 * want


def test_bloch_ring_example13(golden_dir):
    blocks = _parse_bloch_fixture(os.path.join(golden_dir, "model_example13.output"))
    assert len(blocks) == 4
    for ik, b in enumerate(blocks):
        _check_bloch_block("chain lattice", 4, 0, [2 * ik * math.pi / 4], b)


def test_bloch_square_example15(golden_dir):
    blocks = _parse_bloch_fixture(os.path.join(golden_dir, "model_example15.output"))
    ks = O.translation_momenta("square lattice", 2)
    assert len(blocks) == len(ks) == 4
    for k, b in zip(ks, blocks):
        np.testing.assert_allclose(k, b["k"], atol=1e-5)
        _check_bloch_block("square lattice", 2, 2, k, b)


@pytest.mark.parametrize("name,L,W,k_up", [("chain lattice", 10, 0, [5, 4]), ("square lattice", 4, 2, [4, 3]),
                                           ("ladder", 5, 0, [5]), ("triangular lattice", 3, 3, [4]),
                                           ("frustrated square lattice", 3, 3, [4])])
def test_bloch_sectors_reproduce_sz_spectrum(name, L, W, k_up):
    rng = np.random.default_rng(L + W)
    N, s, t, types = O.lattice(name, L, W)
    # couplings by bond type keep the lattice translation invariant
    jz_t, jxy_t = rng.uniform(0.3, 1.5, 3), rng.uniform(0.3, 1.5, 3)
    m = O.Model(N, s, t, jz_t[types], jxy_t[types], h=0.3)
    for n_up in k_up:
        ev = []
        for k in O.translation_momenta(name, L, W):
            ph, pm = O.translations(name, L, W, k=k)
            H = O.bloch_dense(m, n_up, ph, pm)
            if H.size:
                assert np.abs(H - H.conj().T).max() < 1e-12
                ev += list(np.linalg.eigvalsh(H))
        full = np.linalg.eigvalsh(O.dense(m, n_up))
        assert len(ev) == len(full)
        np.testing.assert_allclose(np.sort(ev), full, atol=1e-11)
