"""The product's ietl templates (include/ietl/*.h) instantiated with host
vectors, against the oracle: same start vector, same recurrence, so alpha,
beta and the reported eigenvalues agree to rounding."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "ietl_host_lanczos")


def run(m, k, max_iter, nev, seed, fixed, tmp_path):
    rp, col, val = O.build_csr(m, k)
    f = tmp_path / "csr.txt"
    with open(f, "w") as fh:
        fh.write(f"{len(rp) - 1} {len(col)}\n")
        fh.write(" ".join(map(str, rp.tolist())) + "\n")
        fh.write(" ".join(map(str, col.tolist())) + "\n")
        fh.write(" ".join(repr(float(v)) for v in val) + "\n")
    out = subprocess.run([BIN, str(f), str(max_iter), str(nev), str(seed), str(int(fixed))],
                         capture_output=True, text=True, check=True).stdout.split("\n")
    ab = np.array([list(map(float, l.split()[1:])) for l in out if l.startswith("ab ")])
    ev = np.array([float(l.split()[1]) for l in out if l.startswith("ev ")])
    its = int(out[0].split()[1])
    return ab, ev, its


@pytest.mark.parametrize("name,L,k,fixed", [("chain lattice", 12, 6, False), ("chain lattice", 10, 4, True),
                                            ("frustrated square lattice", 4, 8, False)])
def test_host_ietl_lanczos_matches_oracle(name, L, k, fixed, tmp_path):
    m = O.lattice_model(name, L, by_type={1: (0.5, 0.5)} if "frustrated" in name else None)
    ref = O.sparsediag(m, k, max_iter=60 if fixed else 0, nev=3, seed=7, fixed=fixed)
    ab, ev, its = run(m, k, 60 if fixed else min(10 * len(O.basis(m.n_sites, k)), 1000), 3, 7, fixed, tmp_path)
    assert its == ref["iterations"]
    assert len(ab) == len(ref["alpha"])
    np.testing.assert_allclose(ab[:, 0], ref["alpha"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ab[:, 1], ref["beta"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ev, ref["ev"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name,L,k", [("chain lattice", 10, 5), ("open chain lattice", 11, 5)])
def test_host_ietl_eigenvectors(name, L, k, tmp_path):
    # lanczos::eigenvectors: Ritz vectors of the reported eigenvalues, checked
    # against the dense eigenvectors (up to sign) and by their residuals
    m = O.lattice_model(name, L, jz=0.7)
    rp, col, val = O.build_csr(m, k)
    f = tmp_path / "csr.txt"
    with open(f, "w") as fh:
        fh.write(f"{len(rp) - 1} {len(col)}\n")
        fh.write(" ".join(map(str, rp.tolist())) + "\n")
        fh.write(" ".join(map(str, col.tolist())) + "\n")
        fh.write(" ".join(repr(float(v)) for v in val) + "\n")
    out = subprocess.run([BIN, str(f), "1000", "2", "3", "0", "vectors"], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    infos = [l.split()[1:] for l in out if l.startswith("info ")]
    vecs = [np.array(list(map(float, l.split()[1:]))) for l in out if l.startswith("vec")]
    H = O.dense(m, k)
    w, Z = np.linalg.eigh(H)
    assert len(vecs) == 2
    for i, (inf, v) in enumerate(zip(infos, vecs)):
        m1, m2, ma, status = map(int, inf[:4])
        lam, res = float(inf[4]), float(inf[5])
        assert status == 0 and 2 <= m1 <= ma
        assert abs(lam - w[np.argmin(abs(w - lam))]) < 1e-10
        assert res < 1e-8 and abs(np.linalg.norm(v) - 1) < 1e-8
        assert np.linalg.norm(H @ v - lam * v) < 1e-8
        j = np.argmin(abs(w - lam))
        if np.min(abs(np.delete(w, j) - w[j])) > 1e-6:   # non-degenerate: the eigenvector itself
            assert abs(abs(v @ Z[:, j]) - 1) < 1e-8


@pytest.mark.parametrize("name,L,k", [("chain lattice", 10, 5), ("frustrated square lattice", 4, 8)])
def test_host_ietl_save_load_resumes_bit_for_bit(name, L, k, tmp_path):
    # lanczos::save / load (src/ietl/lanczos.h:86-98): the first half of a
    # fixed run archived, a fresh object loaded from the archive continues
    # with more_eigenvalues, and alpha/beta equal the uninterrupted run's
    m = O.lattice_model(name, L, by_type={1: (0.5, 0.5)} if "frustrated" in name else None)
    rp, col, val = O.build_csr(m, k)
    f = tmp_path / "csr.txt"
    with open(f, "w") as fh:
        fh.write(f"{len(rp) - 1} {len(col)}\n")
        fh.write(" ".join(map(str, rp.tolist())) + "\n")
        fh.write(" ".join(map(str, col.tolist())) + "\n")
        fh.write(" ".join(repr(float(v)) for v in val) + "\n")
    out = subprocess.run([BIN, str(f), "70", "2", "5", "1", "resume"], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    line = [l for l in out if l.startswith("resume_equal")][0].split()
    assert line[1] == "1" and int(line[3]) > 16 * (len(rp) - 1)
