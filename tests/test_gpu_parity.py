"""GPU parity: the gfx950 engine (through its C ABI) against the CPU oracle.

Bars: bit-exact for basis states, indices and matrix elements; FP64 H.v to
1e-13 relative; FP32-storage H.v to 2e-6 relative; eigenvalues to 1e-10
relative (the accuracy target of FP32-stored Lanczos vectors against FP64;
the reference's own ED comparisons use max(1e-11, 1e-11|ref|),
lib/pyalps/apptest.py:388)."""
import math
import os
import subprocess
import sys

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


def random_xxz(L, periodic, seed, field=True):
    rng = np.random.default_rng(seed)
    nb = L if periodic else L - 1
    s = np.arange(nb)
    J = rng.uniform(0.5, 1.5, nb)
    delta = rng.uniform(0, 2)
    return O.Model(L, s, (s + 1) % L, delta * J, J, h=rng.uniform(0, 0.5) if field else 0.0)


def general_model(seed):
    rng = np.random.default_rng(seed)
    N, s, t, _ = O.lattice("square lattice", 4, 3)
    nb = s.size
    return O.Model(N, s, t, rng.uniform(-1, 1, nb), rng.uniform(-1, 1, nb), k=rng.uniform(-0.5, 0.5, nb),
                   h=rng.uniform(-0.5, 0.5, N), d=rng.uniform(-0.5, 0.5, N))


CASES = {
    "heis_chain12": (lambda: O.lattice_model("chain lattice", 12), [0, 3, 6, 9, 12]),
    "heis_chain16": (lambda: O.lattice_model("chain lattice", 16), [8, 5]),
    "xxz_open13": (lambda: random_xxz(13, False, 1), [6, 4]),
    "xxz_ring14": (lambda: random_xxz(14, True, 2), [7, 3]),
    "j1j2_4x4": (lambda: O.lattice_model("frustrated square lattice", 4, 4, by_type={1: (0.5, 0.5)}), [8, 7]),
    "ladder7": (lambda: O.lattice_model("ladder", 7), [7]),
    "general_4x3": (lambda: general_model(5), [6, 5, 0, 1]),
    "ring2": (lambda: O.lattice_model("chain lattice", 2), [1]),
}


def case_params():
    for name, (mk, sectors) in CASES.items():
        for k in sectors:
            yield pytest.param(name, k, id=f"{name}-k{k}")


@pytest.mark.parametrize("name,k", list(case_params()))
def test_states_and_elements_bit_exact(name, k):
    say("case", name, k)
    om = CASES[name][0]()
    op = A.Operator(to_product(om, k))
    want = O.basis(om.n_sites, k)
    assert op.dim == len(want)
    assert np.array_equal(op.states(), want)
    assert np.array_equal(op.index(want), np.arange(len(want), dtype=np.uint64))
    rp, col, val = op.elements()
    orp, ocol, oval = O.build_csr(om, k)
    assert np.array_equal(rp, orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(val.view(np.int64), oval.view(np.int64)), "matrix elements differ bitwise"


@pytest.mark.parametrize("name,k", list(case_params()))
def test_matvec_vs_oracle(name, k):
    say("case", name, k)
    om = CASES[name][0]()
    op = A.Operator(to_product(om, k))
    rng = np.random.default_rng(k + 17)
    x = rng.uniform(-1, 1, op.dim)
    y_ref = O.spmv(om, k, x)
    nrm = max(np.linalg.norm(y_ref), 1e-300)
    y64 = op.matvec(A.Vector.from_numpy(x, "f64"), A.Vector(op.dim, "f64")).numpy()
    assert np.linalg.norm(y64 - y_ref) <= 1e-13 * nrm
    x32 = A.Vector.from_numpy(x, "f32")
    y32 = op.matvec(x32, A.Vector(op.dim, "f32")).numpy()
    y_ref32 = O.spmv(om, k, x32.numpy())  # same FP32-rounded input
    assert np.linalg.norm(y32 - y_ref32) <= 2e-6 * max(np.linalg.norm(y_ref32), 1e-300)


def test_index_rejects_states_outside_the_sector():
    op = A.Operator(to_product(O.lattice_model("chain lattice", 10), 5))
    bad = np.array([0, 0b1111, 0b11111 << 5 | 1, 1 << 10 | 0b1111], dtype=np.uint32)
    assert np.all(op.index(bad) == op.dim)


def test_full_size_l24_rows_and_symmetry():
    """C2 shape (L=24, Sz=0, D=2,704,156): sampled rows of the FP64 product
    against oracle rows built by binary-search lookup; basis round trip;
    <x,Hy> = <Hx,y>."""
    L, k = 24, 12
    om = O.lattice_model("chain lattice", L)
    op = A.Operator(to_product(om, k))
    assert op.dim == math.comb(24, 12) == 2704156
    rng = np.random.default_rng(24)
    x = rng.uniform(-1, 1, op.dim)
    X = A.Vector.from_numpy(x, "f64")
    Y = op.matvec(X, A.Vector(op.dim, "f64"))
    y = Y.numpy()
    say("matvec done")
    for r0 in (0, 1234567, op.dim - 3000):
        rp, col, val = O.build_csr(om, k, r0, r0 + 3000)
        yr = np.array([val[rp[i]:rp[i + 1]] @ x[col[rp[i]:rp[i + 1]]] for i in range(3000)])
        np.testing.assert_allclose(y[r0:r0 + 3000], yr, rtol=1e-13, atol=1e-13)
        st = op.states(r0, 3000)
        assert np.array_equal(op.index(st), np.arange(r0, r0 + 3000, dtype=np.uint64))
        assert np.all(np.diff(st.astype(np.int64)) > 0)
        say("rows", r0, "ok")
    z = rng.uniform(-1, 1, op.dim)
    Z = A.Vector.from_numpy(z, "f64")
    HZ = op.matvec(Z, A.Vector(op.dim, "f64"))
    lhs, rhs = X.dot(HZ), Z.dot(Y)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_device_lanczos_coefficients_vs_oracle():
    om = O.lattice_model("chain lattice", 14)
    op = A.Operator(to_product(om, 7))
    ref = O.sparsediag(om, 7, max_iter=40, fixed=True, seed=4)
    lz = A.Lanczos(op, "f64", max_steps=40)
    lz.start(seed=4)
    lz.steps(39)
    a, b = lz.coefficients()
    assert len(a) == 40
    np.testing.assert_allclose(a[:25], ref["alpha"][:25], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(b[:25], ref["beta"][:25], rtol=1e-9, atol=1e-9)


def test_fused_step_on_caller_vectors_matches_generic_sequence():
    om = random_xxz(12, True, 8)
    op = A.Operator(to_product(om, 6))
    n = op.dim
    rng = np.random.default_rng(3)
    v0 = rng.uniform(-1, 1, n)
    v0 /= np.linalg.norm(v0)
    v1 = rng.uniform(-1, 1, n)
    v1 -= (v1 @ v0) * v0
    v1 /= np.linalg.norm(v1)
    b0 = 0.37
    w = O.spmv(om, 6, v1)
    a_ref = v1 @ w
    w = w - a_ref * v1 - b0 * v0
    b_ref = np.linalg.norm(w)
    P, C, W = A.Vector.from_numpy(v0), A.Vector.from_numpy(v1), A.Vector(n, "f64")
    a, b = A.lanczos_step_vectors(op, P, C, W, b0)
    assert abs(a - a_ref) < 1e-12 and abs(b - b_ref) < 1e-12
    np.testing.assert_allclose(W.numpy(), w / b_ref, atol=1e-12)
    np.testing.assert_array_equal(C.numpy(), v1)  # inputs untouched


def test_recurrence_is_deterministic():
    op = A.Operator(to_product(O.lattice_model("chain lattice", 20), 10))
    runs = []
    for _ in range(2):
        lz = A.Lanczos(op, "f32", max_steps=30)
        lz.start(seed=11)
        lz.steps(29)
        runs.append(lz.coefficients())
    assert np.array_equal(runs[0][0].view(np.int64), runs[1][0].view(np.int64))
    assert np.array_equal(runs[0][1].view(np.int64), runs[1][1].view(np.int64))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_sparsediag_l16_fixed200(dtype):
    """L=16 Heisenberg ring, Sz=0 (D=12,870), 200 Lanczos iterations,
    lowest eigenvalue and the next distinct one."""
    om = O.lattice_model("chain lattice", 16)
    ref = O.sparsediag(om, 8, max_iter=200, nev=2, fixed=True, seed=1)
    got = A.sparsediag(to_product(om, 8), max_iter=200, nev=2, fixed=True, seed=1, dtype=dtype)
    say(dtype, got["ev"], ref["ev"], got["iterations"])
    assert got["dim"] == 12870 and got["iterations"] == 200
    np.testing.assert_allclose(got["ev"], ref["ev"], rtol=1e-10, atol=0)


@pytest.mark.parametrize("name,k", [("heis_chain16", 8), ("j1j2_4x4", 8), ("xxz_ring14", 7), ("general_4x3", 6)])
def test_sparsediag_nlowest_fp64_vs_oracle(name, k):
    om = CASES[name][0]()
    ref = O.sparsediag(om, k, nev=3, seed=2)
    got = A.sparsediag(to_product(om, k), nev=3, seed=2, dtype="f64")
    say(name, got["ev"], ref["ev"], got["iterations"], ref["iterations"])
    np.testing.assert_allclose(got["ev"], ref["ev"], rtol=1e-10, atol=1e-12)


def test_sparsediag_edge_dimensions():
    m = to_product(O.lattice_model("chain lattice", 6, h=0.3), 0)
    r = A.sparsediag(m)
    assert r["dim"] == 1 and r["iterations"] == 0 and r["ev"][0] == 1.5 + 6 * 0.15
    r = A.sparsediag(m.with_sector(7))
    assert r["dim"] == 0 and len(r["ev"]) == 0


@pytest.mark.parametrize("nev", [1, 2])
def test_sparsediag_l24_fp32_and_fp64_vs_oracle(nev):
    """L=24 Heisenberg ring, Sz=0 (D=2,704,156): FP64-stored and FP32-stored
    Lanczos vectors (FP64 accumulation, FP64 Rayleigh refinement) against the
    oracle's FP64 sparsediag of the same sector (tests/golden/
    ring24_oracle.json, written by make_ring24_fixture.py), 1e-10 relative on
    every reported level."""
    ref = np.array(O.ring24_fixture(12, nev)["eigenvalues"])
    m = to_product(O.lattice_model("chain lattice", 24), 12)
    r64 = A.sparsediag(m, nev=nev, seed=1, dtype="f64")
    r32 = A.sparsediag(m, nev=nev, seed=1, dtype="f32")
    say("oracle", ref.tolist())
    say("f64", r64["ev"].tolist(), r64["iterations"], r64["seconds"])
    say("f32", r32["ev"].tolist(), r32["iterations"], r32["seconds"])
    assert r64["error_code"] == 0 and r32["error_code"] == 0
    assert len(r32["ev"]) == len(r64["ev"]) == len(ref) == nev
    for r in (r64, r32):
        rel = np.abs(r["ev"] - ref) / np.abs(ref)
        say("rel err vs oracle", rel.tolist())
        assert rel.max() <= 1e-10
    # FP32 storage must not need more iterations than FP64 here
    assert r32["iterations"] <= r64["iterations"]


def test_sparsediag_l24_four_levels():
    """The lowest four levels of the L=24 ring, FP64 and FP32 storage (the
    latter with partial reorthogonalization), against the oracle's FP64
    sparsediag (1e-10), and SU(2): the lowest Sz=1 level (also from the
    oracle) is one of them."""
    ref = np.array(O.ring24_fixture(12, 4)["eigenvalues"])
    trip = O.ring24_fixture(13, 1)["eigenvalues"][0]
    assert np.min(np.abs(ref - trip)) < 1e-9 * abs(trip)
    m = to_product(O.lattice_model("chain lattice", 24), 12)
    for dtype in ("f64", "f32"):
        r = A.sparsediag(m, nev=4, seed=1, dtype=dtype)
        say(dtype, r["ev"].tolist(), r["iterations"], r["storage"], r["reorthogonalizations"], r["seconds"])
        assert len(r["ev"]) == 4 and r["storage"] == dtype
        assert (np.abs(r["ev"] - ref) / np.abs(ref)).max() <= 1e-10
    t = A.sparsediag(to_product(O.lattice_model("chain lattice", 24), 13), nev=1, seed=2, dtype="f64")
    assert abs(t["ev"][0] - trip) <= 1e-10 * abs(trip)


def test_ritz_vectors_from_kept_basis():
    """Ritz vectors over the kept basis (lanczos::eigenvectors' basis
    transformation) are eigenvectors of the oracle matrix once converged."""
    om = O.lattice_model("chain lattice", 12)
    op = A.Operator(to_product(om, 6))
    lz = A.Lanczos(op, "f64", max_steps=60, keep_basis=True)
    lz.start(seed=3)
    lz.steps(59)
    a, b = lz.coefficients()
    T = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
    w, Z = np.linalg.eigh(T)
    ys = lz.ritz(Z[:, :1].T)
    y = ys[0].numpy()
    H = O.dense(om, 6)
    e0 = np.linalg.eigvalsh(H)[0]
    assert abs(w[0] - e0) < 1e-10
    assert abs(np.linalg.norm(y) - 1.0) < 1e-8
    assert np.linalg.norm(H @ y - e0 * y) < 1e-7


IETL_GPU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "ietl_gpu_lanczos")


@pytest.mark.parametrize("dtype,rtol", [("f64", 1e-10), ("f32", 1e-6)])
def test_ietl_lanczos_on_device_vectorspace(dtype, rtol):
    """ietl::lanczos<alps_ed::hamiltonian, alps_ed::gpu_vectorspace> (the C++
    drop-in, generic recurrence with per-step host T analysis).  FP32 storage
    reports the T-matrix eigenvalues unrefined, so its bar is 1e-6."""
    om = O.lattice_model("chain lattice", 14)
    ref = O.sparsediag(om, 7, nev=3)
    out = subprocess.run([IETL_GPU, "14", "7", dtype, "1000", "3", "5"], capture_output=True, text=True,
                         timeout=120)
    say(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    ev = [float(l.split()[1]) for l in out.stdout.split("\n") if l.startswith("ev ")]
    np.testing.assert_allclose(ev[:1], ref["ev"][:1], rtol=rtol, atol=0)
    if dtype == "f64":
        np.testing.assert_allclose(ev, ref["ev"], rtol=rtol, atol=0)


@pytest.mark.parametrize("dtype,res_tol", [("f64", 1e-8), ("f32", 1e-4)])
def test_ietl_eigenvectors_on_device_vectorspace(dtype, res_tol):
    """lanczos::eigenvectors with device vectors: the recurrence is replayed
    on the GPU (it must reproduce alpha and beta bit for bit) and the Ritz
    vectors are checked against the dense eigenvectors."""
    om = O.lattice_model("chain lattice", 12)   # the program's Heisenberg ring
    H = O.dense(om, 6)
    w, Z = np.linalg.eigh(H)
    out = subprocess.run([IETL_GPU, "12", "6", dtype, "1000", "2", "9", "vectors"], capture_output=True, text=True,
                         timeout=120)
    say(out.stdout[:2000])
    assert out.returncode == 0, out.stdout + out.stderr
    infos = [l.split()[1:] for l in out.stdout.split("\n") if l.startswith("info ")]
    vecs = [np.array(list(map(float, l.split()[1:]))) for l in out.stdout.split("\n") if l.startswith("vec")]
    assert len(infos) == 2 and len(vecs) == 2
    # With FP32 storage the T values drift at ~1e-9 between lengths, above the
    # FP64 merge threshold the m1/m2 search uses, so only the ground state is
    # required to be found (the procedure reports the others as no_eigenvalue).
    for inf, v in list(zip(infos, vecs))[: (2 if dtype == "f64" else 1)]:
        status, lam, res, nrm = int(inf[3]), float(inf[4]), float(inf[5]), float(inf[6])
        # FP32 Lanczos vectors lose orthogonality sooner, so the Ritz vector is
        # not of unit norm; its direction is what is checked
        assert status == 0 and res < res_tol * nrm
        if dtype == "f64":
            assert abs(nrm - 1) < res_tol
        j = np.argmin(abs(w - lam))
        # the generic ietl path reports T eigenvalues: FP32 storage bounds them at ~1e-9
        assert abs(lam - w[j]) < (1e-9 if dtype == "f64" else 1e-7 * abs(w[j]))
        assert np.linalg.norm(H @ v - lam * v) < 10 * res_tol * np.linalg.norm(v)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_parameter_file_sectors_ground_and_gap(prec):
    # a sparsediag parameter file with Sz conserved: every sector's spectrum,
    # the ground-state energy and the gap against the oracle
    text = ('LATTICE="open ladder"; L=5; J0=1; J1=0.7; h=0.1; NUMBER_EIGENVALUES=2; '
            f'CONSERVED_QUANTUMNUMBERS="Sz"; ED_PRECISION="{prec}"; {{}}')
    say("parameter file", prec)
    (task,) = A.run_parameter_file(text)
    om = O.lattice_model("open ladder", 5, 0, h=0.1, by_type={1: (0.7, 0.7)})
    assert [s["n_up"] for s in task["sectors"]] == list(range(11))
    ref = []
    for s in task["sectors"]:
        r = O.sparsediag(om, s["n_up"], nev=2)
        ref.append(r["ev"])
        assert s["dim"] == r["dim"] and s["error_code"] == 0
        np.testing.assert_allclose(s["eigenvalues"], r["ev"], rtol=1e-10, atol=1e-12)
    g, gap = O.ground_and_gap(ref)
    assert task["ground_state_energy"] == pytest.approx(g, rel=1e-10)
    assert task["energy_gap"] == pytest.approx(gap, rel=1e-8, abs=1e-10)


def test_parameter_file_whole_space():
    # no conserved quantum numbers: the lowest distinct eigenvalues of the
    # whole 2^N space, from the dense spectra of all sectors together
    text = 'LATTICE="chain lattice"; L=9; J=1; NUMBER_EIGENVALUES=3; TRANSLATION_SYMMETRY=false; {} Sz_total=0.5; {}'
    say("parameter file, whole space")
    whole, one = A.run_parameter_file(text)
    om = O.lattice_model("chain lattice", 9)
    spec = np.sort(np.concatenate([np.linalg.eigvalsh(O.dense(om, k)) for k in range(10)]))
    distinct = [spec[0]]
    for x in spec[1:]:
        if x - distinct[-1] > 1e-9:
            distinct.append(x)
    (s,) = whole["sectors"]
    assert s["dim"] == 2 ** 9 and s["Sz_total"] is None
    np.testing.assert_allclose(s["eigenvalues"], distinct[:3], rtol=1e-10)
    # odd ring: the ground doublet sits in Sz=+-1/2, so the gap between
    # sectors is zero while the whole space reports the next distinct level
    assert whole["energy_gap"] == pytest.approx(distinct[1] - distinct[0], rel=1e-9)
    (s1,) = one["sectors"]
    assert s1["n_up"] == 5 and s1["dim"] == math.comb(9, 5)
    assert s1["eigenvalues"][0] == pytest.approx(distinct[0], rel=1e-10)


def test_cli_runs_parameter_file(tmp_path):
    import json

    f = tmp_path / "parm"
    f.write_text('LATTICE="chain lattice"\nL=8\nJ=1\nSz_total=0\nNUMBER_EIGENVALUES=2\nTRANSLATION_SYMMETRY=false\n{}\n')
    cli = os.path.join(os.path.dirname(A.LIB_PATH), "alps_ed_sparsediag")
    say("cli", cli)
    out = subprocess.run([cli, str(f)], check=True, capture_output=True, text=True, timeout=60).stdout
    (t,) = json.loads(out)["tasks"]
    r = O.sparsediag(O.lattice_model("chain lattice", 8), 4, nev=2)
    np.testing.assert_allclose(t["sectors"][0]["eigenvalues"], r["ev"], rtol=1e-10)


@pytest.mark.parametrize("name", ["heis_chain16", "xxz_open13", "j1j2_4x4", "general_4x3", "ladder7"])
def test_hv_schedules_give_identical_results(name):
    # the FP32 H.v schedules differ in waves per SIMD and chunk sizes only:
    # every one must give the same y (flip-bond counts that are not
    # multiples of the chunks exercise the padding), and match the oracle
    mk, sectors = CASES[name]
    m = mk()
    k = sectors[0]
    say("schedules", name)
    op = A.Operator(to_product(m, k))
    x = A.Vector(op.dim, "f32")
    x.fill_random(3)
    xs = x.numpy()
    outs = []
    for s in range(A.Operator.HV_SCHEDULES):
        op.set_schedule(s)
        y = A.Vector(op.dim, "f32")
        op.matvec(x, y)
        outs.append(y.numpy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    want = O.spmv(m, k, xs)
    np.testing.assert_allclose(outs[0], want, rtol=2e-6, atol=2e-6 * np.abs(want).max())


def test_hv_autotuner_full_size():
    # L=24 ring: the operator times every schedule at creation and keeps the
    # fastest; all schedules give the same y on the full sector
    op = A.Operator(to_product(O.lattice_model("chain lattice", 24), 12))
    chosen, sec = op.schedule()
    say("autotuner", chosen, sec)
    assert sec is not None and np.all(sec > 0) and sec[chosen] == sec.min()
    x = A.Vector(op.dim, "f32")
    x.fill_random(5)
    ys = []
    for s in (0, chosen, A.Operator.HV_SCHEDULES - 1):
        op.set_schedule(s)
        y = A.Vector(op.dim, "f32")
        op.matvec(x, y)
        ys.append(y.numpy())
    assert np.array_equal(ys[0], ys[1]) and np.array_equal(ys[0], ys[2])


@pytest.mark.parametrize("name", ["heis_chain16", "j1j2_4x4", "general_4x3", "xxz_open13"])
def test_hv_schedules_give_identical_lanczos_coefficients(name):
    # the dot partials of H.v (alpha) are formed per segment with the same
    # reduction tree by every schedule, so the whole FP32 recurrence is
    # bit-identical whichever schedule the autotuner picks
    mk, sectors = CASES[name]
    m = mk()
    op = A.Operator(to_product(m, sectors[0]))
    runs = []
    for s in range(A.Operator.HV_SCHEDULES):
        op.set_schedule(s)
        lz = A.Lanczos(op, "f32", max_steps=40)
        lz.start(seed=7)
        lz.steps(39)
        runs.append(lz.coefficients())
    for a, b in runs[1:]:
        assert np.array_equal(a.view(np.int64), runs[0][0].view(np.int64))
        assert np.array_equal(b.view(np.int64), runs[0][1].view(np.int64))
