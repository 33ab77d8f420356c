"""GPU parity of the eigenvector measurements (DiagMatrix::perform_measurements,
applications/diag/diag.h:191-314): the device spin sums through the C ABI
against the oracle's operator-matrix restatement (expect_site_term), and the
parameter-file measurements against the oracle on dense eigenvectors, with
size-independent sum rules at L=24.

Bars: spin sums of a given vector to 1e-12 relative (FP64 accumulation of
the same stored values); measurements of Lanczos eigenvectors to 1e-8 (FP64
vectors) and 1e-6 (FP32-stored basis, FP64 Ritz vectors) absolute."""
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


def _open_xxz(L, seed):
    rng = np.random.default_rng(seed)
    s = np.arange(L - 1)
    J = rng.uniform(0.5, 1.5, L - 1)
    return O.Model(L, s, s + 1, rng.uniform(0, 2) * J, J, h=rng.uniform(0, 0.5))


CASES = {
    "chain10": (lambda: O.lattice_model("chain lattice", 10), 5),
    "chain16": (lambda: O.lattice_model("chain lattice", 16), 8),
    "ladder7": (lambda: O.lattice_model("ladder", 7), 6),
    "square4x3": (lambda: O.lattice_model("open square lattice", 4, 3), 5),
    "xxz_open13": (lambda: _open_xxz(13, 3), 4),
}


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("name", list(CASES))
def test_spin_sums_match_oracle(name, dtype):
    mk, k = CASES[name]
    m = mk()
    N = m.n_sites
    say("spin sums", name, dtype)
    op = A.Operator(to_product(m, k))
    v = A.Vector(op.dim, dtype)
    v.fill_random(7)
    x = v.numpy()
    n2, sz, szsz = op.spin_moments(v)
    scale = float(x @ x)
    assert abs(n2 - scale) <= 1e-12 * scale
    for i in range(N):
        assert abs(sz[i] - O.expect_site_term(N, k, x, "Sz", i)) <= 1e-12 * scale, i
        for j in range(N):
            want = O.expect_site_term(N, k, x, "Sz", i, "Sz", j)
            assert abs(szsz[i, j] - want) <= 1e-12 * scale, (i, j)
    pi, pj = np.array([(i, j) for i in range(N) for j in range(N) if i != j]).T
    f = op.flip_pairs(v, pi, pj)
    for a, b, got in zip(pi, pj, f):
        want = O.expect_site_term(N, k, x, "Splus", int(a), "Sminus", int(b))
        assert abs(got - want) <= 1e-12 * scale, (a, b)
    # the same sums of the vector's transpose pairs: <S+_i S-_j> = <S+_j S-_i> for a real vector
    np.testing.assert_allclose(f, op.flip_pairs(v, pj, pi), rtol=0, atol=1e-12 * scale)


MEASURES = """MEASURE_AVERAGE[Magnetization]=Sz
MEASURE_LOCAL[Local magnetization]=Sz
MEASURE_LOCAL[Bond energy]=exchange
MEASURE_CORRELATIONS[Diagonal spin correlations]=Sz
MEASURE_CORRELATIONS[Offdiagonal spin correlations]="Splus:Sminus"
MEASURE_CORRELATIONS[Sx correlations]="Sx:Sx"
MEASURE_STRUCTURE_FACTOR[Structure Factor S]=Sz
"""


def _oracle_values(m, k, v, graph_pairs, bonds):
    N = m.n_sites
    loc = [O.expect_site_term(N, k, v, "Sz", i) for i in range(N)]
    ex = [O.expect_site_term(N, k, v, "Sz", i, "Sz", j)
          + 0.5 * (O.expect_site_term(N, k, v, "Splus", i, "Sminus", j)
                   + O.expect_site_term(N, k, v, "Sminus", i, "Splus", j)) for i, j in bonds]
    corr = {op: [O.expect_site_term(N, k, v, a, i, b, j) for i, j in graph_pairs]
            for op, (a, b) in {"Diagonal spin correlations": ("Sz", "Sz"),
                               "Offdiagonal spin correlations": ("Splus", "Sminus"),
                               "Sx correlations": ("Sx", "Sx")}.items()}
    return loc, ex, corr


@pytest.mark.parametrize("precision,tol", [("f64", 1e-8), ("f32", 1e-6)])
@pytest.mark.parametrize("lattice,size,sz", [("chain lattice", "L=10", 0), ("open chain lattice", "L=9", 0.5),
                                             ("ladder", "L=6", 0)])
def test_parameter_file_measurements_match_oracle(lattice, size, sz, precision, tol):
    say("measurements", lattice, size, sz, precision)
    text = (f'LATTICE="{lattice}"; {size}; J=1; Jz=0.8; h=0.1; Sz_total={sz}; ED_PRECISION="{precision}";\n'
            "TRANSLATION_SYMMETRY=false\n"
            + MEASURES + "{}\n")
    (t,) = A.run_parameter_file(text)
    (d,) = A.describe_parameter_file(text)
    (sec,) = t["sectors"]
    N = t["n_sites"]
    k = int(round(sz + N / 2))
    m = O.Model(N, np.array(d["bond_s"]), np.array(d["bond_t"]), np.array(d["jz"]), np.array(d["jxy"]),
                h=np.array(d["h"]))
    w, V = np.linalg.eigh(O.dense(m, k))
    assert w[1] - w[0] > 1e-6, "the test needs a non-degenerate ground state"
    assert abs(sec["eigenvalues"][0] - w[0]) < 1e-9
    v = V[:, 0]
    meas = sec["measurements"]
    # first site pair per distance label, and bonds in edge order
    labels = meas["correlations"]["Diagonal spin correlations"]["labels"]
    coords = [lab for lab in meas["local"]["Local magnetization"]["labels"]]
    pairs = [(coords.index(lab.split(" -- ")[0]), coords.index(lab.split(" -- ")[1])) for lab in labels]
    bonds = list(zip(d["bond_s"], d["bond_t"]))
    loc, ex, corr = _oracle_values(m, k, v, pairs, bonds)
    np.testing.assert_allclose(meas["local"]["Local magnetization"]["values"][0], loc, atol=tol, rtol=0)
    np.testing.assert_allclose(meas["local"]["Bond energy"]["values"][0], ex, atol=tol, rtol=0)
    for name, want in corr.items():
        np.testing.assert_allclose(meas["correlations"][name]["values"][0], want, atol=tol, rtol=0, err_msg=name)
    assert abs(meas["average"]["Magnetization"][0] - sz / N) < tol
    sf = np.array(meas["structure_factor"]["Structure Factor S"]["values"][0])
    # sum rules of the Fourier sum over the N momenta: sum_k S(k) = sum_i <Sz_i^2> = N/4
    assert abs(sf.sum() - N / 4) < tol * N
    if lattice != "open chain lattice":
        # k = 0 on a periodic lattice: (1/N) sum_ij <Sz_i Sz_j> = (Sz_total)^2 / N
        assert abs(sf[0] - sz * sz / N) < tol * N
    assert meas["unsupported"] == []


def test_measurements_full_size_ring24_sum_rules():
    # L=24 Heisenberg ring, Sz=0, FP32 storage: the ground state is a
    # translation-invariant singlet, so every bond energy is E0/24, the
    # correlations from site 0 sum to zero, <Sz_0^2> = 1/4, and the Splus:Sminus
    # structure factor sums to sum_i <S+_i S-_i> = N/2.
    say("measurements ring24")
    text = ('LATTICE="chain lattice"; L=24; J=1; Sz_total=0; ED_PRECISION="f32"; TRANSLATION_SYMMETRY=false;\n'
            "MEASURE_LOCAL[Bond energy]=exchange\nMEASURE_CORRELATIONS[SzSz]=Sz\n"
            'MEASURE_STRUCTURE_FACTOR[Sxy]="Splus:Sminus"\n{}\n')
    (t,) = A.run_parameter_file(text)
    (sec,) = t["sectors"]
    e0 = sec["eigenvalues"][0]
    e_ref = O.ring24_fixture(12, 1)["eigenvalues"][0]  # the oracle's FP64 sparsediag
    assert abs(e0 - e_ref) < 1e-10 * abs(e_ref)
    meas = sec["measurements"]
    be = np.array(meas["local"]["Bond energy"]["values"][0])
    assert abs(be.sum() - e0) < 1e-6
    assert be.max() - be.min() < 1e-6
    c = np.array(meas["correlations"]["SzSz"]["values"][0])
    assert len(c) == 24 and abs(c[0] - 0.25) < 1e-12 and abs(c.sum()) < 1e-6
    assert abs(c[1] - e0 / 24 / 3) < 1e-6  # singlet: <Sz Sz> = <S.S>/3
    sf = np.array(meas["structure_factor"]["Sxy"]["values"][0])
    assert abs(sf.sum() - 12.0) < 1e-5
    assert sf.argmax() == 12  # k = pi (the 13th momentum label)
    assert meas["structure_factor"]["Sxy"]["labels"][12] == "( 3.14159 )"


def test_measurements_need_sz_sectors():
    text = 'LATTICE="chain lattice"; L=6; J=1; TRANSLATION_SYMMETRY=false;\nMEASURE_LOCAL[M]=Sz\n{}\n'
    (t,) = A.run_parameter_file(text)
    assert "measurement_error" in t
    assert all("measurements" not in s for s in t["sectors"])


def _xml_check(doc, task):
    """The task result document against the same run's JSON."""
    import xml.etree.ElementTree as ET

    root = ET.fromstring(doc)
    assert root.tag == "SIMULATION"
    params = {p.get("name"): p.text or "" for p in root.find("PARAMETERS")}
    assert params == task["parameters"]
    avg = {a.get("name"): float(a.find("MEAN").text) for a in root.find("AVERAGES")}
    assert avg["Ground State Energy"] == task["ground_state_energy"]
    # no gap with a single level (sparsediag.h:80-84 writes it only when set)
    assert avg.get("Energy Gap") == task["energy_gap"]
    evs = root.findall("EIGENVALUES")
    states = root.findall("EIGENSTATES")
    assert len(evs) == len(states) == len(task["sectors"])
    return evs, states


def test_task_xml_matches_results(tmp_path):
    # quantum-number loop over all Sz of an odd ring: sector values are half
    # integers written as the reference's half_integer prints them
    say("task xml")
    text = ('LATTICE="chain lattice"; L=5; J=1; Jz=0.7; CONSERVED_QUANTUMNUMBERS="Sz"; NUMBER_EIGENVALUES=2;\n'
            "TRANSLATION_SYMMETRY=false\n"
            "MEASURE_LOCAL[Local magnetization]=Sz\nMEASURE_AVERAGE[Magnetization]=Sz\n{}\n")
    (t,) = A.run_parameter_file(text)
    evs, states = _xml_check(t["xml"], t)
    want_sz = ["-5/2", "-3/2", "-1/2", "1/2", "3/2", "5/2"]
    for e, st, sec, sz in zip(evs, states, t["sectors"], want_sz):
        assert [(q.get("name"), q.get("value")) for q in e.findall("QUANTUMNUMBER")] == [("Sz", sz)]
        vals = [float(x) for x in "".join(e.itertext()).split()]
        assert int(e.get("number")) == len(vals) == min(2, len(sec["eigenvalues"]))
        assert vals == sec["eigenvalues"][: len(vals)]
        for j, es in enumerate(st.findall("EIGENSTATE")):
            sc = {a.get("name"): float(a.find("MEAN").text) for a in es.findall("SCALAR_AVERAGE")}
            assert sc["Energy"] == vals[j]
            assert sc["Magnetization"] == sec["measurements"]["average"]["Magnetization"][j]
            (va,) = es.findall("VECTOR_AVERAGE")
            loc = sec["measurements"]["local"]["Local magnetization"]
            assert [s.get("indexvalue") for s in va] == loc["labels"]
            assert [float(s.find("MEAN").text) for s in va] == loc["values"][j]
            # a fixed-Sz eigenvector: the local magnetisations sum to Sz
            assert abs(sum(loc["values"][j]) - float(eval(sz))) < 1e-9


@pytest.mark.parametrize("flag,measured,states", [("", False, True), ("MEASURE_ENERGY=false;", False, False),
                                                   ("MEASURE_ENERGY=false; MEASURE_AVERAGE[M]=Sz;", True, True)])
def test_task_xml_eigenstates_follow_measure_energy(flag, measured, states):
    # DiagTask::write_xml_body (diag.hpp:141): <EIGENSTATES> only when
    # averages are measured or MEASURE_ENERGY (default true) holds
    import xml.etree.ElementTree as ET

    text = f'LATTICE="chain lattice"; L=6; J=1; Sz_total=0; TRANSLATION_SYMMETRY=false; {flag}\n{{}}\n'
    (t,) = A.run_parameter_file(text)
    root = ET.fromstring(t["xml"])
    assert len(root.findall("EIGENVALUES")) == 1
    assert len(root.findall("EIGENSTATES")) == (1 if states else 0)
    assert ("measurements" in t["sectors"][0]) == measured


def test_cli_writes_task_xml(tmp_path):
    import os
    import subprocess

    f = tmp_path / "parm.txt"
    f.write_text('LATTICE="chain lattice"\nL=8\nJ=1\nSz_total=0\nTRANSLATION_SYMMETRY=false\n{}\n{J=0.5;}\n')
    cli = os.path.join(os.path.dirname(A.LIB_PATH), "alps_ed_sparsediag")
    say("cli xml", cli)
    subprocess.run([cli, "--write-xml", str(f)], check=True, capture_output=True, text=True, timeout=60)
    tasks = A.run_parameter_file(f.read_text())
    for i, t in enumerate(tasks):
        doc = (tmp_path / f"parm.task{i + 1}.out.xml").read_text()
        evs, _ = _xml_check(doc, t)
        assert evs[0].findall("QUANTUMNUMBER") == []  # Sz_total given: no loop


def test_task_result_file_reruns(tmp_path):
    # a task result file holds the task's parameters in the task file form,
    # so it is itself an input; a job runs it and writes to its <OUTPUT>
    import os
    import subprocess

    say("xml round trip")
    text = ('LATTICE="chain lattice"; L=8; J=1; Jz=0.7; Sz_total=0; NUMBER_EIGENVALUES=2;\n'
            'TRANSLATION_SYMMETRY=false; NOTE="a<b & c"\n{}\n')
    (t,) = A.run_parameter_file(text)
    (again,) = A.run_parameter_file(t["xml"])
    assert again["parameters"] == t["parameters"] and again["parameters"]["NOTE"] == "a<b & c"
    assert [s["eigenvalues"] for s in again["sectors"]] == [s["eigenvalues"] for s in t["sectors"]]
    assert again["xml"] == t["xml"]

    (tmp_path / "t1.in.xml").write_text(t["xml"])
    job = tmp_path / "job.in.xml"
    job.write_text('<JOB><OUTPUT file="job.out.xml"/>'
                   '<TASK status="new"><INPUT file="t1.in.xml"/><OUTPUT file="t1.out.xml"/></TASK></JOB>')
    cli = os.path.join(os.path.dirname(A.LIB_PATH), "alps_ed_sparsediag")
    r = subprocess.run([cli, "--write-xml", str(job)], check=True, capture_output=True, text=True, timeout=60)
    assert '"ground_state_energy"' in r.stdout
    assert (tmp_path / "t1.out.xml").read_text() == t["xml"]

