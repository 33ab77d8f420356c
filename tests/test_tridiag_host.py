"""Host checks of the T-matrix helpers of the solves (tests/cpp/
tridiag_host.cpp, no device), and the host programs built under
AddressSanitizer and UndefinedBehaviorSanitizer (tests/cpp/Makefile).

- the lowest eigenvalue of the O(m) convergence test against plain Sturm
  bisection on 3000 tridiagonal matrices (random, nearly decoupled, nearly
  double lowest values, Wilkinson-like, scales 1e-3..1e5), with and without a
  bracket hint;
- its error bound |beta_{m-1} z_{m-1}| never above the one of cyclic
  Jacobi's eigenvector (an overestimate would stop a solve later than the
  nlowest rule, iteration.h:116-130);
- tridiag_eigenvector against Jacobi's eigenvector;
- the test as a necessary condition: wherever the full rule holds on the T
  matrix of a real FP64 Lanczos run, the test passes;
- the bottom-up analysis of the lowest entries (analyze_low) against the
  full T-matrix analysis on Lanczos T matrices with ghosts and on random
  ones: the same entries (values to 1e-14 |T|, indices), and the same FP32
  stopping decisions."""
import os
import subprocess

import pytest

BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build")


def run(name):
    out = subprocess.run([os.path.join(BUILD, name)], capture_output=True, text=True, check=True,
                         timeout=300).stdout
    vals = {}
    for line in out.splitlines():
        f = line.split()
        vals.update(zip(f[::2], f[1::2]))
    return vals


@pytest.mark.parametrize("name", ["tridiag_host", "tridiag_host_san"])
def test_tridiag_helpers(name):
    v = run(name)
    assert int(v["cases"]) == 3000
    assert float(v["lam_rel"]) < 1e-15          # of |T|
    assert int(v["hint_bad"]) == 0
    assert float(v["err_over"]) < 1e-3
    assert float(v["vec_bad"]) < 1e-12
    assert int(v["gate_checks"]) > 5000 and int(v["rule_true"]) > 500
    assert int(v["gate_true"]) >= int(v["rule_true"])
    assert int(v["gate_bad"]) == 0
    assert int(v["low_checks"]) > 10000 and int(v["ghost_seen"]) > 1000
    assert int(v["low_count_bad"]) == 0 and int(v["low_index_bad"]) == 0
    assert float(v["low_ev"]) < 1e-14
    assert int(v["low_rule_bad32"]) == 0


def test_layout_host_under_sanitizers():
    v = run("layout_host_san")
    assert float(v["diag_err"]) < 1e-13
    assert int(v["parse_bad"]) == 0 and int(v["proc_bad"]) == 0
    assert int(v["flip_bad"]) == 0 and int(v["bloch_sectors"]) == 417
