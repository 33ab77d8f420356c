"""Host checks of the engine's sector layout and momentum-sector builders
(tests/cpp/layout_host.cpp, no device): the diagonal tables against a direct
sum over sites and bonds on random models, the lattice translations and
Bloch bases against the oracle's restatement of blochbasisstates.h, the
flip tables read with the H.v kernel's addressing against the flipped words'
ranks, and the per-high-word flip lists against a recomputation from the
states."""
import os
import subprocess

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "layout_host")


def test_layout_diagonal_and_bloch_bases():
    out = subprocess.run([BIN], capture_output=True, text=True, check=True, timeout=120).stdout.split("\n")
    vals = {}
    for line in out:
        f = line.split()
        vals.update(zip(f[::2], f[1::2]))
    assert float(vals["diag_err"]) < 1e-13
    # 5 lattices x every Sz sector x every momentum, minus none
    assert int(vals["bloch_sectors"]) == 417
    assert int(vals["flip_bad"]) == 0 and int(vals["flip_checked"]) > 10**6
    assert int(vals["proc_bad"]) == 0
    assert int(vals["list_bad"]) == 0 and int(vals["list_checked"]) > 10**5
    assert int(vals["parse_bad"]) == 0
