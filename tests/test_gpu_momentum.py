"""GPU parity of the momentum (Bloch) sectors: the device H.v on realified ((re, im)
per element) vectors against the oracle's dense Bloch Hamiltonian (the restatement of
blochbasisstates.h and hamiltonian_matrix.hpp pinned by the reference's
example13/15 fixtures and by the Sz-sector spectrum), Lanczos eigenvalues per
momentum, the parameter-file momentum loop of DiagMatrix::dostep (run as
one batch), batched momentum sectors against single solves, and at L=24
size-independent checks (sector dimensions add up to the Sz sector,
the lowest level over all momenta is the Sz ground state).

Bars: H.v 1e-12 relative (FP64), 2e-6 (FP32 storage); eigenvalues 1e-10."""
import math
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_product(m: O.Model, n_up) -> A.SpinModel:
    return A.SpinModel(m.n_sites, m.bond_s, m.bond_t, m.jz, m.jxy, n_up, k=m.k, h=m.h, d=m.d)


def typed_model(name, L, W, seed):
    """Couplings by bond type (translation invariant), random values."""
    rng = np.random.default_rng(seed)
    N, s, t, types = O.lattice(name, L, W)
    jz, jxy, kk = rng.uniform(0.3, 1.5, 3), rng.uniform(0.3, 1.5, 3), rng.uniform(-0.2, 0.2, 3)
    return O.Model(N, s, t, jz[types], jxy[types], k=kk[types], h=0.2, d=0.1)


CASES = {
    "chain10": ("chain lattice", 10, 0, [5, 4]),
    "square4x2": ("square lattice", 4, 2, [4]),
    "ladder5": ("ladder", 5, 0, [5, 6]),
    "triangular3x3": ("triangular lattice", 3, 3, [4]),
    "frustrated3x3": ("frustrated square lattice", 3, 3, [5]),
}


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("case", list(CASES))
def test_bloch_hv_matches_oracle(case, dtype):
    name, L, W, ups = CASES[case]
    m = typed_model(name, L, W, 3)
    rng = np.random.default_rng(11)
    for n_up in ups:
        for k in O.translation_momenta(name, L, W):
            say("bloch hv", case, dtype, n_up, k)
            tr = O.translations(name, L, W, k=k)
            H = O.bloch_dense(m, n_up, *tr)
            dk = H.shape[0]
            op = A.Operator(to_product(m, n_up), translations=tr)
            if dk == 0:  # every orbit of the sector cancels at this momentum
                assert op.dim == 0
                continue
            assert op.dim == 2 * dk
            x = rng.uniform(-1, 1, 2 * dk)
            xv, yv = A.Vector(2 * dk, dtype), A.Vector(2 * dk, dtype)
            xv.upload(x)
            xs = xv.numpy()
            op.matvec(xv, yv)
            y = yv.numpy()
            want = H @ (xs[0::2] + 1j * xs[1::2])
            tol = 1e-12 if dtype == "f64" else 2e-6
            scale = np.abs(want).max() + 1e-300
            assert np.abs(y[0::2] - want.real).max() <= tol * scale
            assert np.abs(y[1::2] - want.imag).max() <= tol * scale


def test_momentum_eigenvalues_match_dense():
    name, L = "chain lattice", 12
    m = typed_model(name, L, 0, 5)
    for n_up in (6, 5):
        for k in O.translation_momenta(name, L):
            tr = O.translations(name, L, k=k)
            H = O.bloch_dense(m, n_up, *tr)
            w = np.linalg.eigvalsh(H)
            distinct = [w[0]] + [x for a, x in zip(w, w[1:]) if x - a > 1e-9]
            say("momentum lanczos", n_up, k, H.shape[0])
            for dtype in ("f64", "f32"):
                r = A.sparsediag(to_product(m, n_up), nev=2, dtype=dtype, translations=tr)
                assert r["dim"] == H.shape[0]
                np.testing.assert_allclose(r["ev"], distinct[:2], rtol=1e-10, atol=1e-12)


def test_parameter_file_momentum_loop():
    # the default TRANSLATION_SYMMETRY=true on a ring loops over every Sz and
    # momentum (Sz outer); each sector against the oracle's dense Bloch
    # matrix, the ground state and gap over all sectors, and the XML's
    # quantum numbers
    text = ('LATTICE="chain lattice"; L=8; J=1; Jz=0.6; h=0.05; CONSERVED_QUANTUMNUMBERS="Sz"; '
            "NUMBER_EIGENVALUES=2; {}")
    say("parameter file momenta")
    (t,) = A.run_parameter_file(text)
    (d,) = A.describe_parameter_file(text)
    ks = O.translation_momenta("chain lattice", 8)
    np.testing.assert_allclose(np.array(d["momenta"]), ks)
    m = O.Model(8, np.array(d["bond_s"]), np.array(d["bond_t"]), np.array(d["jz"]), np.array(d["jxy"]),
                h=np.array(d["h"]))
    ref, got = [], []
    for n_up in range(9):
        for k in ks:
            H = O.bloch_dense(m, n_up, *O.translations("chain lattice", 8, k=k))
            if H.shape[0] == 0:
                continue
            w = np.linalg.eigvalsh(H)
            distinct = [w[0]] + [x for a, x in zip(w, w[1:]) if x - a > 1e-9]
            ref.append((n_up, float(k[0]), H.shape[0], distinct[:2]))
    assert len(t["sectors"]) == len(ref)
    for s, (n_up, k, dk, ev) in zip(t["sectors"], ref):
        assert s["n_up"] == n_up and s["momentum"][0] == pytest.approx(k) and s["dim"] == dk
        np.testing.assert_allclose(s["eigenvalues"], ev, rtol=1e-10, atol=1e-12)
    g, gap = O.ground_and_gap([ev for *_, ev in ref])
    assert t["ground_state_energy"] == pytest.approx(g, rel=1e-10)
    assert t["energy_gap"] == pytest.approx(gap, rel=1e-8, abs=1e-10)
    # every sector's dimensions add up to 2^8
    assert sum(s["dim"] for s in t["sectors"]) == 2 ** 8
    root = ET.fromstring(t["xml"])
    q = [[(x.get("name"), x.get("value")) for x in e.findall("QUANTUMNUMBER")] for e in root.findall("EIGENVALUES")]
    assert q[0] == [("Sz", "-4"), ("TOTAL_MOMENTUM", "0")]
    assert q[1][1] == ("TOTAL_MOMENTUM", "%.20g" % t["sectors"][1]["momentum"][0])


def test_momentum_ring24_sectors():
    # L=24 Heisenberg ring, Sz=0: the 24 momentum sectors' dimensions add up
    # to C(24,12) and their lowest level is the Sz=0 ground state, at k=0
    name, L = "chain lattice", 24
    m = O.lattice_model(name, L)
    dims, lows = [], []
    for k in O.translation_momenta(name, L):
        say("ring24 momentum", k)
        r = A.sparsediag(to_product(m, 12), nev=1, dtype="f32", translations=O.translations(name, L, k=k))
        dims.append(r["dim"])
        lows.append(r["ev"][0])
    assert sum(dims) == math.comb(24, 12)
    # the oracle's FP64 sparsediag of the Sz=0 sector
    assert min(lows) == pytest.approx(O.ring24_fixture(12, 1)["eigenvalues"][0], rel=1e-10)
    assert int(np.argmin(lows)) == 0


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_momentum_sectors_match_single_solves(dtype):
    # every (Sz, k) sector of a 4x3 square lattice through one batch, mixed
    # with fixed-Sz problems, against the single-solve path; a momentum
    # problem's result is the same alone and in the mix
    name, L, W = "square lattice", 4, 3
    m = typed_model(name, L, W, 7)
    ks = O.translation_momenta(name, L, W)
    b = A.Batch(dtype, hbm_bytes=2 << 30, steps_per_round=6)
    jobs = []
    for n_up in (3, 5, 6):
        b.submit(to_product(m, n_up), nev=2)  # fixed-Sz neighbours in the mix
        for k in ks:
            tr = O.translations(name, L, W, k=k)
            jobs.append((m, n_up, tr, b.submit(to_product(m, n_up), nev=2, translations=tr)))
    # sectors of 2,700 states, above the FP32 storage threshold
    ring = typed_model("chain lattice", 18, 0, 9)
    for k in O.translation_momenta("chain lattice", 18)[:4]:
        tr = O.translations("chain lattice", 18, k=k)
        jobs.append((ring, 9, tr, b.submit(to_product(ring, 9), nev=2, translations=tr)))
    say("batch momentum", dtype, len(jobs))
    b.run()
    dims = {}
    for mm, n_up, tr, t in jobs:
        r = b.result(t)
        ref = A.sparsediag(to_product(mm, n_up), nev=2, dtype=dtype, translations=tr)
        assert r["dim"] == ref["dim"] and r["error_code"] == 0
        if mm is m:
            dims[n_up] = dims.get(n_up, 0) + r["dim"]
        if r["dim"]:
            np.testing.assert_allclose(r["ev"], ref["ev"], rtol=1e-10, atol=1e-12)
    for n_up, d in dims.items():
        assert d == math.comb(12, n_up)
    # momentum problems alone give bit-identical eigenvalues
    for mm, n_up, tr, t in (jobs[1], jobs[-1]):
        alone = A.Batch(dtype, hbm_bytes=1 << 30, steps_per_round=6)
        ta = alone.submit(to_product(mm, n_up), nev=2, translations=tr)
        alone.run()
        assert np.array_equal(alone.result(ta)["ev"], b.result(t)["ev"])


def _bloch_expand_reference(N, n_up, phases, perms, w):
    """Configuration-basis coefficients of a momentum-sector vector from the
    oracle's Bloch basis: on each state r of a kept orbit, phase_t * norm_a *
    w_a (a its representative, the smallest state of its orbit; t the first
    translation reaching it)."""
    reps, norm = O.bloch_basis(N, n_up, phases, perms)
    idx = {int(r): i for i, r in enumerate(reps)}
    states = [s for s in range(1 << N) if bin(s).count("1") == n_up]

    def image(s, perm):
        out = 0
        for i in range(N):
            if (s >> (N - 1 - i)) & 1:
                out |= 1 << (N - 1 - perm[i])
        return out

    v = np.zeros(len(states), dtype=complex)
    for r, s in enumerate(states):
        best, bt = s, 0
        for t in range(1, len(perms)):
            y = image(s, perms[t])
            if y < best:
                best, bt = y, t
        if best in idx:
            a = idx[best]
            v[r] = phases[bt] * norm[a] * w[a]
    return v


@pytest.mark.parametrize("case", ["chain10", "square4x2", "triangular3x3"])
def test_bloch_expand(case):
    # the device expansion against the oracle's basis, and an expanded
    # momentum eigenvector is an eigenvector of the fixed-Sz Hamiltonian
    name, L, W, ups = CASES[case]
    m = typed_model(name, L, W, 3)
    N = m.n_sites
    n_up = ups[0]
    rng = np.random.default_rng(2)
    plain = A.Operator(to_product(m, n_up))
    for k in O.translation_momenta(name, L, W)[:3]:
        tr = O.translations(name, L, W, k=k)
        op = A.Operator(to_product(m, n_up), translations=tr)
        dk = op.dim // 2
        if dk == 0:
            continue
        w = rng.uniform(-1, 1, 2 * dk)
        re, im = op.bloch_expand(A.Vector.from_numpy(w, "f64"))
        want = _bloch_expand_reference(N, n_up, *tr, w[0::2] + 1j * w[1::2])
        np.testing.assert_allclose(re.numpy(), want.real, rtol=0, atol=1e-14)
        np.testing.assert_allclose(im.numpy(), want.imag, rtol=0, atol=1e-14)
        # eigenvector of the dense Bloch matrix -> eigenvector of H
        H = O.bloch_dense(m, n_up, *tr)
        lam, V = np.linalg.eigh(H)
        ev = np.empty(2 * dk)
        ev[0::2], ev[1::2] = V[:, 0].real, V[:, 0].imag
        re, im = op.bloch_expand(A.Vector.from_numpy(ev, "f64"))
        for part in (re, im):
            hp = plain.matvec(part, A.Vector(plain.dim, "f64")).numpy()
            assert np.linalg.norm(hp - lam[0] * part.numpy()) < 1e-10
        assert abs(np.linalg.norm(re.numpy()) ** 2 + np.linalg.norm(im.numpy()) ** 2 - 1) < 1e-12


def test_momentum_measurements_match_sz_sector():
    # measurements in the momentum sectors (the reference's default on a
    # ring) against the same task without momenta: the non-degenerate
    # ground state of the L=10 Heisenberg ring (Sz=0, k=pi) gives the same
    # averages, correlations and structure factors; local values are not
    # reported in momentum sectors (diag.h:212)
    meas = ("MEASURE_AVERAGE[Magnetization]=Sz\n"
            "MEASURE_LOCAL[Local]=Sz\nMEASURE_CORRELATIONS[SzSz]=Sz\n"
            'MEASURE_CORRELATIONS[SpSm]="Splus:Sminus"\nMEASURE_STRUCTURE_FACTOR[Szk]=Sz\n')
    base = 'LATTICE="chain lattice"; L=10; J=1; Jz=0.8; Sz_total=0;\n' + meas
    say("momentum measurements")
    (tk,) = A.run_parameter_file(base + "{}\n")
    (t0,) = A.run_parameter_file(base + "TRANSLATION_SYMMETRY=false;\n{}\n")
    assert "measurement_error" not in tk
    low = min(tk["sectors"], key=lambda s: s["eigenvalues"][0])
    (ref,) = t0["sectors"]
    assert low["eigenvalues"][0] == pytest.approx(ref["eigenvalues"][0], rel=1e-10)
    mk, m0 = low["measurements"], ref["measurements"]
    assert mk["local"] == {} and "Local" in m0["local"]
    assert mk["average"]["Magnetization"][0] == pytest.approx(m0["average"]["Magnetization"][0], abs=1e-10)
    for group in ("correlations", "structure_factor"):
        for name, series in m0[group].items():
            assert mk[group][name]["labels"] == series["labels"]
            np.testing.assert_allclose(mk[group][name]["values"][0], series["values"][0], rtol=1e-8, atol=1e-10)

