"""Explicit graphs in parameter files (GRAPH, LATTICE_LIBRARY) and the
lattice geometry the task description reports, on the host (no device).

The reference picks a task's graph in graph_helper::make_graph
(src/alps/lattice/graph_helper.h:651-700) and reads <GRAPH> elements with
read_graph_xml (src/alps/lattice/graph.h:120-215); the spin couplings of an
edge of type t are Jz<t>, Jxy<t> (default J<t>, default 0) and the field of
a vertex of type t is h<t> (lib/xml/models.xml.in:342-373).

Fixture: tests/golden/lattice_example5.output is the reference's
test/lattice/example5 output (the sites of an anisotropic square lattice,
L=10, W=4, with their types and coordinates)."""
import os
import re

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_lattice_example5_golden():
    # test/lattice/example5.input: its LATTICE_LIBRARY is the reference's
    # lattices.xml, whose lattices are built in here
    text = ('LATTICE_LIBRARY = "../../lib/xml/lattices.xml"\nLATTICE = "anisotropic square lattice"\n'
            "L = 10\nW = 4\n{}\n")
    (t,) = A.describe_parameter_file(text)
    with open(os.path.join(GOLDEN, "lattice_example5.output")) as f:
        golden = f.read()
    n = int(re.search(r"The graph has (\d+) sites", golden).group(1))
    sites = re.findall(r"Site (\d+) has type (\d+) and has coordinates ([\d. -]+?) ?\n", golden)
    assert t["n_sites"] == n == len(sites) == 40
    for s, ty, coord in sites:
        s = int(s)
        assert t["site_type"][s] == int(ty)
        assert t["coordinates"][s] == [float(c) for c in coord.split()]


def test_builtin_graphs():
    # lib/xml/lattices.xml.in:617-641, 1-based ids there
    want = {
        "5-site dimerized": (5, [0] * 5, [(1, 2, 0), (2, 3, 1), (3, 4, 0), (4, 5, 1)]),
        "4-site mixed": (4, [0, 1, 0, 1], [(1, 2, 0), (2, 3, 0), (3, 4, 0), (4, 1, 0), (1, 3, 1), (2, 4, 1)]),
        "triangle": (3, [0] * 3, [(1, 2, 0), (2, 3, 0), (3, 1, 0)]),
    }
    for name, (n, vtypes, edges) in want.items():
        (t,) = A.describe_parameter_file(f'GRAPH="{name}"; J0=1; J1=0.25; h0=0.5; h1=-0.5; {{}}')
        assert t["n_sites"] == n
        assert t["site_type"] == vtypes
        assert [(s + 1, d + 1, int(ty)) for s, d, ty in zip(t["bond_s"], t["bond_t"], t["bond_type"])] == edges
        assert t["jz"] == [1.0 if ty == 0 else 0.25 for _, _, ty in edges] == t["jxy"]
        assert t["h"] == [0.5 if ty == 0 else -0.5 for ty in vtypes]
        assert "momenta" not in t and "coordinates" not in t
    # a built-in GRAPH name is also found through LATTICE (graph_helper.h:672-677)
    (t,) = A.describe_parameter_file('LATTICE="triangle"; J=1; {}')
    assert t["n_sites"] == 3 and len(t["bond_s"]) == 3


def random_bond_library(L, periodic, seed):
    """A lattice library with one GRAPH: a chain of L sites whose edge i has
    type i, and the couplings of each type drawn as the batched
    random-bond workload draws them (J ~ U[0.5,1.5], Delta ~ U[0,2])."""
    rng = np.random.default_rng(seed)
    n_e = L if periodic else L - 1
    J = rng.uniform(0.5, 1.5, n_e)
    delta = rng.uniform(0.0, 2.0)
    lines = ['<LATTICES>', '<!-- random-bond XXZ chain -->',
             f'<GRAPH name="random-bond chain" vertices="{L}">']
    for i in range(n_e):
        lines.append(f'  <EDGE type="{i}" source="{i + 1}" target="{(i + 1) % L + 1}"/>')
    lines += ['</GRAPH>', '</LATTICES>']
    params = "; ".join(f"Jxy{i}={float(J[i])!r}; Jz{i}={float(J[i] * delta)!r}" for i in range(n_e))
    return "\n".join(lines) + "\n", params, J, delta


@pytest.mark.parametrize("periodic", [True, False])
def test_random_bond_graph_matches_oracle_chain(tmp_path, periodic):
    L = 13
    xml, params, J, delta = random_bond_library(L, periodic, seed=11 + periodic)
    lib = tmp_path / "chains.xml"
    lib.write_text(xml)
    (t,) = A.describe_parameter_file(f'LATTICE_LIBRARY="{lib}"; GRAPH="random-bond chain"; '
                                     f"Sz_total=0.5; DISORDER_SEED=7; {params}; {{}}")
    n, s, d, ty = O.lattice("chain lattice" if periodic else "open chain lattice", L)
    assert t["n_sites"] == n == L
    np.testing.assert_array_equal(t["bond_s"], s)
    np.testing.assert_array_equal(t["bond_t"], d)
    assert t["bond_type"] == list(range(len(s)))
    np.testing.assert_array_equal(t["jxy"], J)
    np.testing.assert_array_equal(t["jz"], J * delta)
    assert t["n_up"] == [7] and t["parameters"]["DISORDER_SEED"] == "7"


def test_graph_vertices_grow_and_default_ids(tmp_path):
    lib = tmp_path / "g.xml"
    lib.write_text('<?xml version="1.0"?>\n<LATTICES><GRAPH name="g">\n'
                   '<VERTEX type="2"><COORDINATE>0 0</COORDINATE></VERTEX>\n<VERTEX/>\n'
                   '<EDGE source="1" target="2"></EDGE>\n<EDGE source="2" target="4" type="3"/>\n'
                   '</GRAPH></LATTICES>\n')
    (t,) = A.describe_parameter_file(f'LATTICE_LIBRARY="{lib}"; GRAPH="g"; J=1; J3=2; h2=1; {{}}')
    assert t["n_sites"] == 4 and t["site_type"] == [2, 0, 0, 0]
    assert t["bond_s"] == [0, 1] and t["bond_t"] == [1, 3] and t["jz"] == [1.0, 2.0]
    assert t["h"] == [1.0, 0.0, 0.0, 0.0]


@pytest.mark.parametrize("xml,params,message", [
    (None, 'GRAPH="triangle"; LATTICE="chain lattice"; L=4', "both GRAPH and LATTICE"),
    (None, 'GRAPH="hexagon"', "could not find graph/lattice"),
    (None, 'GRAPH="chain lattice"; L=4', "could not find graph/lattice"),
    (None, 'GRAPH="triangle"; TOTAL_MOMENTUM="0"', "TOTAL_MOMENTUM needs a LATTICE"),
    (None, 'GRAPH="triangle"; J=random()', "random couplings"),
    (None, 'GRAPH="triangle"; MEASURE_CORRELATIONS[C]="Sz"', "need a LATTICE"),
    ('<GRAPH name="g" vertices="2"><VERTEX id="3"/></GRAPH>', "", "too many vertices"),
    ('<GRAPH name="g" vertices="2"><SITE/></GRAPH>', "", "illegal element <SITE>"),
    ('<GRAPH name="g" vertices="2"><EDGE source="1" target="2"><X/></EDGE></GRAPH>', "", "Nonempty <EDGE>"),
    ('<GRAPH name="g" vertices="2"><EDGE source="1" target="1"/></GRAPH>', "", "to itself"),
    ('<GRAPH name="g" vertices="2"><EDGE source="1" target="2"/>', "", "</GRAPH> tag missing"),
    ('<GRAPH name="g"><EDGE source="a" target="2"/></GRAPH>', "", "bad EDGE source"),
])
def test_graph_errors(tmp_path, xml, params, message):
    if xml is not None:
        lib = tmp_path / "bad.xml"
        lib.write_text(xml)
        params = f'LATTICE_LIBRARY="{lib}"; GRAPH="g"; ' + params
    with pytest.raises(A.EDError) as e:
        A.describe_parameter_file(params + "; {}")
    assert message in str(e.value)


def test_missing_library_file(tmp_path):
    with pytest.raises(A.EDError) as e:
        A.describe_parameter_file(f'LATTICE_LIBRARY="{tmp_path}/none.xml"; GRAPH="g"; {{}}')
    assert "Could not find lattice library file" in str(e.value)
