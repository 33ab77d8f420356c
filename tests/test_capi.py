"""CPU checks of the engine library: it loads, exports every entry point
declared in include/*.h, and its host-only T-matrix analysis agrees with the
oracle.  No device work runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(alps_ed_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(A.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_declarations():
    A.lib()  # sets argtypes for every binding; each must exist in the header
    declared = set(declared_functions())
    bound = {n for n in dir(A.lib()) if n.startswith("alps_ed_")}
    assert bound <= declared


@pytest.mark.parametrize("n", [2, 5, 17, 60, 120])
def test_tmatrix_analysis_matches_oracle(n):
    m = O.lattice_model("chain lattice", 12)
    r = O.sparsediag(m, 6, max_iter=120, fixed=True, seed=9)
    a, b = r["alpha"][:n], r["beta"][:n]
    t1 = O.tmatrix(a, b)
    t2 = A.tmatrix_analyze(a, b)
    np.testing.assert_array_equal(t1["mult"], t2["mult"])
    np.testing.assert_allclose(t2["eig"], t1["eig"], atol=1e-13, rtol=0)
    # error bounds are last-row eigenvector entries times beta: rounding-level
    np.testing.assert_allclose(t2["err"], t1["err"], atol=5e-14, rtol=0)


def test_tmatrix_ghosts_removed_after_convergence():
    # long recurrence on a small matrix: copies of converged values appear and
    # the distinct non-ghost list still equals the exact distinct spectrum
    m = O.lattice_model("chain lattice", 8)
    w = np.linalg.eigvalsh(O.dense(m, 4))
    d = [w[0]]
    for x in w[1:]:
        if x - d[-1] > 1e-9:
            d.append(x)
    r = O.sparsediag(m, 4, max_iter=300, fixed=True, seed=2)
    t = A.tmatrix_analyze(r["alpha"], r["beta"])
    assert len(t["eig"]) == len(d)
    np.testing.assert_allclose(t["eig"], d, atol=1e-9)


def test_errors_are_reported_without_device():
    with pytest.raises(A.EDError):
        A.tmatrix_analyze(np.zeros(0), np.zeros(0))
