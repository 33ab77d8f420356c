"""Host checks of the engine's sector layout and momentum-sector builders
(tests/cpp/layout_host.cpp, no device): the diagonal tables against a direct
sum over sites and bonds on random models, and the lattice translations and
Bloch bases against the oracle's restatement of blochbasisstates.h."""
import os
import subprocess

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "layout_host")


def test_layout_diagonal_and_bloch_bases():
    out = subprocess.run([BIN], capture_output=True, text=True, check=True, timeout=120).stdout.split("\n")
    vals = dict(line.split() for line in out if line)
    assert float(vals["diag_err"]) < 1e-13
    # 5 lattices x every Sz sector x every momentum, minus none
    assert int(vals["bloch_sectors"]) == 417
