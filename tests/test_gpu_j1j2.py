"""GPU parity at the shape of the 5x6 J1-J2 workload: spin-1/2 Heisenberg on
the periodic "frustrated square lattice" with L=6, W=5 (lattices.xml.in:
405-414, unit cell nnn2d :213-219; 30 sites, 120 bonds), J=1 on the
nearest-neighbour bonds (type 0) and J1=0.5 on the diagonal ones (type 1),
Sz=0: D = C(30,15) = 155,117,520 rows, about 63 stored elements per row.

No oracle basis of this size fits in a test, so the checks are fixed by the
mathematics (spin_formulas): colex ranks of sampled rows, sampled rows of
H x against the spin-model row formula with Python integers, the closed form
of 1^T H 1 and the symmetry <x, H z> = <H x, z>.  Couplings and inputs are
dyadic, so these comparisons are bitwise.  The Lanczos solves compare FP32
vector storage with FP64 storage after 300 fixed iterations (1e-10), and
check SU(2) symmetry: every Sz=1 level is also an Sz=0 level."""
import math
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O
import spin_formulas as F

pytestmark = pytest.mark.gpu

PARAMS = 'LATTICE="frustrated square lattice"; L=6; W=5; J=1; J1=0.5; Sz_total=0; {}'
D0 = 155117520


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def j1j2_model(n_up=15):
    """The model as the parameter-file reader resolves it, checked against
    the oracle's lattice graph and per-type couplings."""
    (t,) = A.describe_parameter_file(PARAMS)
    m = A.SpinModel(t["n_sites"], np.array(t["bond_s"]), np.array(t["bond_t"]), np.array(t["jz"]),
                    np.array(t["jxy"]), n_up)
    om = O.lattice_model("frustrated square lattice", 6, 5, by_type={1: (0.5, 0.5)})
    assert m.n_sites == om.n_sites == 30 and len(m.jz) == 120
    assert np.array_equal(m.bond_site1, om.bond_s) and np.array_equal(m.bond_site2, om.bond_t)
    assert np.array_equal(m.jz, om.jz) and np.array_equal(m.jxy, om.jxy)
    return m


def test_j1j2_rows_ranks_sum_and_symmetry():
    m = j1j2_model()
    op = A.Operator(m)
    D = op.dim
    assert D == math.comb(30, 15) == D0
    say("operator built", D, "schedule", op.schedule()[0])

    rng = np.random.default_rng(30)
    rows = sorted(set([0, 1, D - 1, D - 2] + [int(v) for v in rng.integers(0, D, 100)]))
    st = np.array([F.colex_unrank(r, 15) for r in rows], dtype=np.uint32)
    assert [F.colex_rank(int(v), 30) for v in st] == rows
    got = np.concatenate([op.states(r, 1) for r in rows])
    assert np.array_equal(got, st)
    assert np.array_equal(op.index(st), np.array(rows, dtype=np.uint64))
    say("ranks ok")

    x = np.empty(D, dtype=np.float64)
    z = np.empty(D, dtype=np.float64)
    step = 1 << 25
    for a in range(0, D, step):
        b = min(D, a + step)
        x[a:b] = F.x_chunk(a, b)
        r = np.arange(a, b, dtype=np.int64)
        z[a:b] = ((7 * r) % 1019 - 509) / 1024.0
    X = A.Vector.from_numpy(x, "f64")
    Z = A.Vector.from_numpy(z, "f64")
    del x, z
    Y = op.matvec(X, A.Vector(D, "f64"))
    y = Y.numpy()
    for r, s in zip(rows, st):
        assert y[r] == F.row_value(m, int(s), r), r
    del y
    say("f64 rows ok")

    # <x, H z> = <H x, z>: every product and partial sum is a dyadic rational
    # well inside FP64's range, so both sides are exact
    HZ = op.matvec(Z, A.Vector(D, "f64"))
    lhs, rhs = X.dot(HZ), Y.dot(Z)
    say("symmetry", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    del HZ, Z

    ones = A.Vector(D, "f64")
    ones.fill(1.0)
    H1 = op.matvec(ones, A.Vector(D, "f64"))
    want = F.closed_form_sum(m)
    assert ones.dot(H1) == want
    del H1, ones, Y
    say("closed-form sum ok", want)

    X32 = A.Vector(D, "f32")
    X32.copy_from(X)
    del X
    Y32 = op.matvec(X32, A.Vector(D, "f32"))
    y32 = Y32.numpy()
    for r, s in zip(rows, st):
        assert y32[r] == np.float32(F.row_value(m, int(s), r)), r
    say("f32 rows ok")


def test_j1j2_fixed300_fp32_vs_fp64():
    """300 fixed Lanczos iterations, lowest two levels: FP32-stored vectors
    (FP64 alpha/beta, tridiagonal solve and Rayleigh refinement) against
    FP64-stored vectors, 1e-10 relative."""
    m = j1j2_model()
    r64 = A.sparsediag(m, max_iter=300, fixed=True, nev=2, seed=1, dtype="f64")
    say("f64", r64["ev"].tolist(), r64["iterations"], r64["seconds"])
    r32 = A.sparsediag(m, max_iter=300, fixed=True, nev=2, seed=1, dtype="f32")
    say("f32", r32["ev"].tolist(), r32["iterations"], r32["seconds"], r32.get("storage"))
    assert r64["dim"] == r32["dim"] == D0
    assert r64["iterations"] == r32["iterations"] == 300
    assert len(r32["ev"]) == len(r64["ev"]) == 2
    rel = np.abs(r32["ev"] - r64["ev"]) / np.abs(r64["ev"])
    say("rel", rel.tolist())
    assert rel.max() <= 1e-10


def test_j1j2_su2_lowest_sz1_level_is_an_sz0_level():
    """Every total-spin multiplet with S >= 1 has an Sz=0 member, so the
    lowest Sz=1 level t is a level of the Sz=0 sector: t >= E0(Sz=0), and
    when t is not above the fourth-lowest Sz=0 level it is one of the lowest
    four (converged nlowest FP64 solves)."""
    m = j1j2_model()
    r0 = A.sparsediag(m, nev=4, seed=3, dtype="f64")
    say("Sz=0", r0["ev"].tolist(), r0["iterations"], r0["error_code"], r0["seconds"])
    r1 = A.sparsediag(m.with_sector(16), nev=1, seed=4, dtype="f64")
    say("Sz=1", r1["ev"].tolist(), r1["iterations"], r1["error_code"], r1["seconds"])
    assert r0["error_code"] == 0 and r1["error_code"] == 0
    assert r1["dim"] == math.comb(30, 16)
    ev, t = r0["ev"], r1["ev"][0]
    tol = 1e-9 * abs(t)
    assert ev[0] <= t + tol
    if t <= ev[-1] + tol:
        assert np.min(np.abs(ev - t)) <= tol
