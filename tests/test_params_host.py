"""Host-side sparsediag task input: the parameter-file grammar, expression
evaluation, lattice graphs, spin-model couplings and sector lists, through
the engine's alps_ed_describe_parameter_file (no device needed).

Fixtures: tests/golden/parameterlist.{input,output} and
expression.{input,output} are the reference's test/parameter files;
lattice_example6.{input,output} its test/lattice example.  The lattice
graphs are also checked against the oracle's, which is pinned by the same
lattice fixture in test_oracle.py.
"""
import math
import os
import re

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _read(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def _blocks(text):
    """Parameter blocks of parameterlist.output: '{', 'name = value;'..., '}'."""
    out = []
    for blk in re.findall(r"\{\n(.*?)\}", text, re.S):
        items = []
        for line in blk.strip().splitlines():
            k, v = line.rstrip(";").split(" = ", 1)
            items.append((k, v))
        out.append(items)
    return out


def test_parameterlist_golden(monkeypatch):
    # the reference's test runs with DIR=/home/alps (the golden's value of dir)
    monkeypatch.setenv("DIR", "/home/alps")
    tasks = A.describe_parameter_file(_read("parameterlist.input"))
    want = _blocks(_read("parameterlist.output"))
    # the golden prints the list twice (read, then read back from a dump)
    assert len(want) == 2 * len(tasks) == 12
    for t, w in zip(tasks, want[: len(tasks)]):
        assert list(t["parameters"].items()) == w        # values and insertion order


def _expression_cases():
    """(expression, printed value) pairs of expression.output that depend
    on no random number generator."""
    text = _read("expression.output")
    cases = re.findall(r"The value of \[(.*?)\] is (\S+)", text)
    return [(e, float(v)) for e, v in cases if "random" not in e]


def test_expression_golden():
    header = "L=10; T=0.1; beta='1/T'; LATTICE=\"open chain lattice\";\n"
    cases = _expression_cases()
    assert len(cases) == 24
    text = header + "".join("{ J=%s }\n" % e for e, _ in cases)
    tasks = A.describe_parameter_file(text)
    for (e, v), t in zip(cases, tasks):
        got = t["jz"][0]
        # the golden prints six significant digits
        assert float("%.6g" % got) == pytest.approx(v, rel=1e-12), e
        assert t["jxy"][0] == got


def test_expression_cannot_evaluate():
    # 'error = "error"' refers to itself, which the reference reports as
    # "Cannot evaluate [error]."; the engine reports an argument error
    with pytest.raises(A.EDError) as ei:
        A.describe_parameter_file('L=4; error="error"; LATTICE="chain lattice"; { J=error }')
    assert ei.value.status == -1


LATTICES = [("chain lattice", 9, 0), ("open chain lattice", 7, 0), ("nnn chain lattice", 10, 0),
            ("square lattice", 4, 3), ("open square lattice", 3, 4), ("anisotropic square lattice", 4, 3),
            ("ladder", 6, 0), ("open ladder", 5, 3), ("frustrated square lattice", 4, 4),
            ("triangular lattice", 3, 4), ("anisotropic triangular lattice", 4, 3), ("simple cubic lattice", 3, 0)]


@pytest.mark.parametrize("name,L,W", LATTICES)
def test_lattice_matches_oracle(name, L, W):
    N, s, t, ty = O.lattice(name, L, W)
    text = f'LATTICE="{name}"; L={L}; ' + (f"W={W}; " if W else "") + "J0=1.5; J1=-0.25; J2=0.125; {}"
    (task,) = A.describe_parameter_file(text)
    assert task["n_sites"] == N
    assert task["bond_s"] == s.tolist() and task["bond_t"] == t.tolist()
    # J0 -> type 0, J1 -> type 1, J2 -> type 2 (the '#' couplings)
    cpl = np.array([1.5, -0.25, 0.125])[ty]
    assert np.array_equal(task["jz"], cpl) and np.array_equal(task["jxy"], cpl)


def test_lattice_example6_input():
    # the reference's lattice example input, read as a parameter file
    want = [tuple(map(int, m)) for m in re.findall(r"between site (\d+) and site (\d+) has type (\d+)",
                                                    _read("lattice_example6.output"))]
    (task,) = A.describe_parameter_file(_read("lattice_example6.input") + "\nJ0=1; J1=2; {}")
    assert task["n_sites"] == 40
    got = list(zip(task["bond_s"], task["bond_t"], [int(j) - 1 for j in task["jz"]]))
    assert got == want


def test_spin_defaults_and_sites():
    (t,) = A.describe_parameter_file('LATTICE="chain lattice"; L=4; J=1; Jz=0.5; h=0.3; D=0.2; K=0.1; {}')
    assert t["jz"] == [0.5] * 4 and t["jxy"] == [1.0] * 4 and t["k"] == [0.0] * 4
    assert t["h"] == [0.3] * 4 and t["d"] == [0.2] * 4
    # primed couplings set type 1; Jxy1 alone overrides the xy part
    (t,) = A.describe_parameter_file("LATTICE=\"nnn chain lattice\"; L=5; J=1; J'=0.5; Jxy1=0.25; K1=0.1; {}")
    assert t["jz"] == [1.0, 0.5] * 5 and t["jxy"] == [1.0, 0.25] * 5 and t["k"] == [0.0, 0.1] * 5
    # expression chains and Pi
    (t,) = A.describe_parameter_file('LATTICE="chain lattice"; L=3; a=2; J="a^2/4"; Jxy="sin(Pi/2)*a"; {}')
    assert t["jz"] == [1.0] * 3 and t["jxy"] == [2.0] * 3


def test_sectors():
    # no conserved quantum numbers: the whole space, one subspace
    (t,) = A.describe_parameter_file('LATTICE="chain lattice"; L=6; J=1; {}')
    assert t["n_up"] == list(range(7)) and t["whole_space"]
    # Sz conserved: one subspace per Sz_total (blank- or comma-separated names)
    for q in ('"Sz"', '"N,Sz"', '"Sz N"'):
        (t,) = A.describe_parameter_file(f'LATTICE="chain lattice"; L=6; J=1; CONSERVED_QUANTUMNUMBERS={q}; {{}}')
        assert t["n_up"] == list(range(7)) and not t["whole_space"], q
    (t,) = A.describe_parameter_file('LATTICE="chain lattice"; L=6; J=1; CONSERVED_QUANTUMNUMBERS="Szz"; {}')
    assert t["whole_space"]
    # a given Sz_total restricts the basis whatever is conserved
    (t,) = A.describe_parameter_file('LATTICE="chain lattice"; L=6; J=1; Sz_total=-1; {}')
    assert t["n_up"] == [2] and not t["whole_space"]
    (t,) = A.describe_parameter_file('LATTICE="chain lattice"; L=5; J=1; Sz_total=0.5; '
                                     'CONSERVED_QUANTUMNUMBERS="Sz"; {}')
    assert t["n_up"] == [3]


@pytest.mark.parametrize("text,msg", [
    ('LATTICE="chain lattice"; L=5; Sz_total=0; {}', "Sz_total"),
    ('LATTICE="chain lattice"; L=4; MODEL="boson Hubbard"; {}', "MODEL"),
    ('LATTICE="chain lattice"; L=4; local_S=1; {}', "local_S"),
    ('LATTICE="chain lattice"; L=4; Gamma=0.5; {}', "Gamma"),
    ('LATTICE="kagome lattice"; L=4; {}', "could not find graph/lattice"),
    ("L=4; { J=1 ", "unterminated"),
    ('LATTICE="chain lattice"; { J=1 }', "L"),
])
def test_rejected_input(text, msg):
    with pytest.raises(A.EDError) as ei:
        A.describe_parameter_file(text)
    assert ei.value.status == -1 and msg in str(ei.value)


def test_no_task_without_braces():
    assert A.describe_parameter_file("A=1; B=2") == []
    assert [t["parameters"] for t in A.describe_parameter_file("A=1; {} A=3; {B=2}")] == \
        [{"A": "1"}, {"A": "3", "B": "2"}]


def test_cli_describe(tmp_path):
    import json
    import subprocess

    f = tmp_path / "parm"
    f.write_text('LATTICE="open chain lattice"; L=3; J=1; {}\n')
    cli = os.path.join(os.path.dirname(A.LIB_PATH), "alps_ed_sparsediag")
    out = subprocess.run([cli, "--describe", str(f)], check=True, capture_output=True, text=True).stdout
    (t,) = json.loads(out)["tasks"]
    assert t["bond_s"] == [0, 1] and t["bond_t"] == [1, 2]
    assert math.isclose(sum(t["jz"]), 2.0)


# ---------------------------------------------------------------------------
# XML inputs: <PARAMETERS>, <PARAMETERLIST>, task files (<SIMULATION>) and
# job files (<JOB>, through the command-line driver)
# ---------------------------------------------------------------------------
def test_parameters_xml_golden():
    # test/parameter/parameters_xml: one <PARAMETERS> document
    (t,) = A.describe_parameter_file(_read("parameters_xml.input"))
    want = [ln.rstrip(";").split(" = ") for ln in _read("parameters_xml.output").splitlines()[:2]]
    assert list(t["parameters"].items()) == [tuple(w) for w in want]


def test_parameterlist_xml_golden():
    # test/parameter/parameterlist_xml: PARAMETER children are defaults for
    # the <PARAMETERS> blocks after them; block values keep their place
    tasks = A.describe_parameter_file(_read("parameterlist_xml.input"))
    want = _blocks(_read("parameterlist_xml.output").split("[Output to alps::oxstream]")[0])
    assert len(want) == len(tasks) == 3
    for t, w in zip(tasks, want):
        assert list(t["parameters"].items()) == w


_TASK_XML = """<?xml version="1.0" encoding="UTF-8"?>
<?xml-stylesheet type="text/xsl" href="ALPS.xsl"?>
<SIMULATION xmlns:xsi="http://www.w3.org/2001/XMLSchema-instance">
  <!-- a comment with <PARAMETER name="X">1</PARAMETER> inside -->
  <PARAMETERS>
    <PARAMETER name="LATTICE">open chain lattice</PARAMETER>
    <PARAMETER name="L">5</PARAMETER>
    <PARAMETER name="a">2</PARAMETER>
    <PARAMETER name="J">a/2</PARAMETER>
    <PARAMETER name="Jz">   0.5  </PARAMETER>
    <PARAMETER name="NOTE">x &lt; y &amp;&amp; &quot;z&quot; &amp;lt;</PARAMETER>
    <PARAMETER name="EMPTY"/>
    <PARAMETER name='L'>6</PARAMETER>
  </PARAMETERS>
  <AVERAGES><SCALAR_AVERAGE name="Ground State Energy"><MEAN>-1</MEAN></SCALAR_AVERAGE></AVERAGES>
</SIMULATION>
"""


def test_task_xml_matches_parameter_file():
    (x,) = A.describe_parameter_file(_TASK_XML)
    assert list(x["parameters"].items()) == [
        ("LATTICE", "open chain lattice"), ("L", "6"), ("a", "2"), ("J", "a/2"), ("Jz", "0.5"),
        ("NOTE", 'x < y && "z" &lt;'), ("EMPTY", "")]
    (p,) = A.describe_parameter_file('LATTICE="open chain lattice"; L=6; a=2; J="a/2"; Jz=0.5; {}')
    for key in ("n_sites", "bond_s", "bond_t", "jz", "jxy", "k", "h", "d", "n_up", "whole_space"):
        assert x[key] == p[key], key
    assert x["jz"] == [0.5] * 5 and x["jxy"] == [1.0] * 5


def test_xml_errors():
    for text, msg in [("<SIMULATION><AVERAGES/></SIMULATION>", "without <PARAMETERS>"),
                      ("<PARAMETERS><PARAMETER>1</PARAMETER></PARAMETERS>", "name attribute"),
                      ("<PARAMETERS><PARAMETER name='L'>1</PARAMETERS>", "unterminated <PARAMETER>"),
                      ("<PARAMETERS><PARAMETER name='L'>1</PARAMETER>", "unterminated <PARAMETERS>"),
                      ("<PARAMETERLIST><FOO/></PARAMETERLIST>", "unknown tag")]:
        with pytest.raises(A.EDError) as ei:
            A.describe_parameter_file(text)
        assert msg in str(ei.value), text


def _cli():
    return os.path.join(os.path.dirname(A.LIB_PATH), "alps_ed_sparsediag")


def test_cli_job_file(tmp_path):
    import json
    import subprocess

    # a job as parameter2xml writes it (tool/parameter2xml.C:95-138), task
    # files in a subdirectory named relative to the job file
    (tmp_path / "sub").mkdir()
    for i, L in enumerate((3, 4, 5), 1):
        (tmp_path / "sub" / f"j.task{i}.in.xml").write_text(
            _TASK_XML.replace(">6<", f">{L}<"))
    job = tmp_path / "j.in.xml"
    job.write_text("""<?xml version="1.0" encoding="UTF-8"?>
<JOB xmlns:xsi="http://www.w3.org/2001/XMLSchema-instance" name="j">
  <OUTPUT file="j.out.xml"/>
  <TASK status="new"><INPUT file="sub/j.task1.in.xml"/><OUTPUT file="sub/j.task1.out.xml"/></TASK>
  <TASK status="finished"><INPUT file="sub/j.task2.in.xml"/><OUTPUT file="sub/j.task2.out.xml"/></TASK>
  <TASK><INPUT file = "sub/j.task3.in.xml"></INPUT></TASK>
</JOB>
""")
    out = subprocess.run([_cli(), "--describe", str(job)], check=True, capture_output=True, text=True).stdout
    tasks = json.loads(out)["tasks"]
    # the finished task is not run
    assert [t["parameters"]["L"] for t in tasks] == ["3", "5"]
    assert [len(t["bond_s"]) for t in tasks] == [2, 4]

    for body, msg in [("<TASK><OUTPUT file='a'/></TASK>", "missing <INPUT>"),
                      ("<TASK status='done'><INPUT file='a'/></TASK>", "illegal status"),
                      ("<TASK><INPUT file='sub/j.task1.in.xml'/>", "missing </TASK>")]:
        job.write_text("<JOB>" + body + "</JOB>")
        r = subprocess.run([_cli(), "--describe", str(job)], capture_output=True, text=True)
        assert r.returncode == 1 and msg in r.stderr, (body, r.stderr)
    job.write_text("<JOB><TASK><INPUT file='sub/none.xml'/></TASK></JOB>")
    r = subprocess.run([_cli(), "--describe", str(job)], capture_output=True, text=True)
    assert r.returncode == 2 and "cannot read" in r.stderr


# ---------------------------------------------------------------------------
# measurement requests (MeasurementOperators, measurement_operators.C:25-66)
# and the lattice labels they are reported with
# ---------------------------------------------------------------------------
_MEAS = """MEASURE_AVERAGE[Magnetization]=Sz
MEASURE_AVERAGE[Exchange]=exchange
MEASURE_LOCAL[Local magnetization]=Sz
MEASURE_CORRELATIONS[Diagonal spin correlations]=Sz
MEASURE_CORRELATIONS[Offdiagonal spin correlations]="Splus:Sminus"
MEASURE_CORRELATIONS[Broken]=Sz=
MEASURE_STRUCTURE_FACTOR[Structure Factor S]=Sz
"""


def test_measure_requests_and_labels():
    text = ('MODEL="spin"; CONSERVED_QUANTUMNUMBERS="Sz"; J=1;\n' + _MEAS +
            '{LATTICE="dimer";}\n{LATTICE="chain lattice";L=4;}\n{LATTICE="open chain lattice";L=3;}\n'
            '{LATTICE="triangular lattice";L=3;W=2;}\n')
    dimer, ring, open3, tri = A.describe_parameter_file(text)
    m = ring["measure"]
    assert m["average"] == {"Exchange": "exchange", "Magnetization": "Sz"}
    assert m["correlations"]["Offdiagonal spin correlations"] == ["Splus", "Sminus"]
    assert m["correlations"]["Diagonal spin correlations"] == ["Sz", "Sz"]
    assert m["correlations"]["Broken"] == ["Sz=", "Sz="]  # evaluated as unsupported
    assert {u["name"] for u in m["unsupported"]} == {"Exchange", "Broken"}
    # periodic chain: one distance per offset, from site 0; momenta 2 pi x / L
    # shifted to (-pi, pi]
    assert m["distance_labels"] == ["( 0 ) -- ( 0 )", "( 0 ) -- ( 1 )", "( 0 ) -- ( 2 )", "( 0 ) -- ( 3 )"]
    assert m["momentum_labels"] == ["( 0 )", "( 1.5708 )", "( 3.14159 )", "( -1.5708 )"]
    assert m["flip_pairs"] == 3  # (0,1), (0,2), (0,3)
    # open boundaries: every ordered pair is its own distance (L^2 of them)
    assert dimer["n_sites"] == 2 and dimer["bond_s"] == [0] and dimer["bond_t"] == [1]
    assert dimer["measure"]["distance_labels"] == ["( 0 ) -- ( 0 )", "( 0 ) -- ( 1 )", "( 1 ) -- ( 0 )",
                                                   "( 1 ) -- ( 1 )"]
    assert len(open3["measure"]["distance_labels"]) == 9 and open3["measure"]["flip_pairs"] == 6
    assert open3["measure"]["momentum_labels"] == ["( 0 )", "( 2.0944 )", "( -2.0944 )"]
    # triangular lattice: Cartesian coordinates a1 = (1,0), a2 = (1/2, sqrt(3)/2)
    # and momenta in the reciprocal basis (lattices.xml.in:45-49)
    t = tri["measure"]
    assert t["distance_labels"][1] == "( 0,0 ) -- ( 0.5,0.866025 )"
    assert t["momentum_labels"][1] == "( 0,3.6276 )"  # offset (0,1): k = (0, pi) -> b2 / 2
    assert t["momentum_labels"][2] == "( 2.0944,-1.2092 )"  # offset (1,0): b1 / 3


def test_measure_absent_without_requests():
    (t,) = A.describe_parameter_file('LATTICE="chain lattice"; L=4; {}')
    assert "measure" not in t


def test_batch_footprint_of_momentum_sectors():
    # host-only sizing: a momentum sector of the L=12 ring holds its Bloch
    # basis (8 bytes per state of the Sz sector) and vectors of 2 * d_k
    # elements; it needs less than the whole Sz sector's solve at FP64, and
    # bad translations are rejected
    bs = np.arange(12)
    m = A.SpinModel(12, bs, (bs + 1) % 12, 1.0, 1.0, 6)
    full = A.batch_footprint(m, "f64", nev=1)
    for k in O.translation_momenta("chain lattice", 12)[:3]:
        tr = O.translations("chain lattice", 12, k=k)
        f = A.batch_footprint(m, "f64", nev=1, translations=tr)
        assert 8 * 924 < f < full
    ph, perms = O.translations("chain lattice", 12, k=[0.0])
    bad = perms.copy()
    bad[1, 0] = bad[1, 1]
    with pytest.raises(A.EDError):
        A.batch_footprint(m, "f64", translations=(ph, bad))
    with pytest.raises(A.EDError):
        A.batch_footprint(m, "f64", translations=(ph[1:], perms[1:]))  # first must be the identity

