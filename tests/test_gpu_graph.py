"""A random-bond XXZ chain given as an explicit GRAPH (one edge type per
bond, lib/xml/lattices.xml.in-style library file), solved through the
parameter-file path and checked against the oracle's dense spectrum of the
same bonds (graph_helper.h:651-700, sparsediag.h:93-149)."""
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O
from test_graph_host import random_bond_library

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("periodic", [True, False])
def test_random_bond_graph_spectrum_vs_oracle(tmp_path, periodic):
    L = 12
    xml, params, J, delta = random_bond_library(L, periodic, seed=3 + periodic)
    lib = tmp_path / "chains.xml"
    lib.write_text(xml)
    (t,) = A.run_parameter_file(f'LATTICE_LIBRARY="{lib}"; GRAPH="random-bond chain"; Sz_total=0; '
                                f"NUMBER_EIGENVALUES=3; {params}; {{}}")
    n, s, d, _ = O.lattice("chain lattice" if periodic else "open chain lattice", L)
    om = O.Model(n, s, d, J * delta, J)
    dense = np.unique(np.round(np.linalg.eigvalsh(O.dense(om, L // 2)), 10))[:3]
    (sec,) = t["sectors"]
    print(periodic, sec["eigenvalues"], dense.tolist(), flush=True)
    sys.stdout.flush()
    np.testing.assert_allclose(sec["eigenvalues"], dense, rtol=0, atol=1e-10)
    assert abs(t["ground_state_energy"] - dense[0]) < 1e-10
