"""Partial reorthogonalization of FP32-stored Lanczos vectors (the omega
recurrence and Gram-Schmidt passes on the device).

The reference's recurrence reorthogonalizes nothing (src/ietl/lanczos.h:
392-406); with FP32 storage its vectors lose orthogonality as soon as a Ritz
value converges, and copies of converged levels appear in T.  These tests
check that the kept basis stays semi-orthogonal with the partial scheme and
is far from it without, that the recurrence stays deterministic, and that
four levels of the L=24 ring in FP32 storage match the oracle's FP64
sparsediag (tests/golden/ring24_oracle.json) to 1e-10."""
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ring(L):
    om = O.lattice_model("chain lattice", L)
    return A.SpinModel(om.n_sites, om.bond_s, om.bond_t, om.jz, om.jxy, L // 2)


def kept_basis(lz, m):
    """The normalized Lanczos vectors v_0..v_{m-1} (FP64 copies of the FP32
    basis) as the columns of a matrix."""
    vs = lz.ritz(np.eye(m))
    return np.stack([v.numpy() for v in vs], axis=1)


@pytest.mark.parametrize("reorth", [True, False])
def test_kept_basis_orthogonality(reorth):
    op = A.Operator(ring(16))
    m = 160
    lz = A.Lanczos(op, "f32", max_steps=m, keep_basis=True, reorth=reorth)
    lz.start(seed=5)
    lz.steps(m - 1)
    V = kept_basis(lz, m)
    G = V.T @ V
    loss = np.abs(G - np.eye(m)).max()
    n_re = lz.reorthogonalizations()
    say("reorth", reorth, "max |V^T V - I|", loss, "reorthogonalized steps", n_re)
    if reorth:
        # semi-orthogonality: sqrt(2^-24) = 2.4e-4 is the trigger level
        assert 0 < n_re < m
        assert loss < 2e-3
    else:
        assert n_re == 0
        assert loss > 1e-1


def test_reorth_recurrence_is_deterministic():
    op = A.Operator(ring(18))
    runs = []
    for _ in range(2):
        lz = A.Lanczos(op, "f32", max_steps=120, keep_basis=True, reorth=True)
        lz.start(seed=9)
        lz.steps(119)
        a, b = lz.coefficients()
        runs.append((a, b, lz.reorthogonalizations()))
    assert runs[0][2] == runs[1][2] > 0
    assert np.array_equal(runs[0][0].view(np.int64), runs[1][0].view(np.int64))
    assert np.array_equal(runs[0][1].view(np.int64), runs[1][1].view(np.int64))


def test_ring24_four_levels_fp32_storage_vs_oracle():
    """The lowest four levels of the L=24 ring (D=2,704,156) with FP32
    storage: served in FP32 with partial reorthogonalization, 1e-10 against
    the oracle's FP64 sparsediag, in no more iterations than FP64 storage."""
    ref = np.array(O.ring24_fixture(12, 4)["eigenvalues"])
    m = ring(24)
    r32 = A.sparsediag(m, nev=4, seed=1, dtype="f32")
    r64 = A.sparsediag(m, nev=4, seed=1, dtype="f64")
    say("oracle", ref.tolist())
    say("f32", r32["ev"].tolist(), r32["iterations"], r32["reorthogonalizations"], r32["seconds"])
    say("f64", r64["ev"].tolist(), r64["iterations"], r64["seconds"])
    assert r32["storage"] == "f32" and r64["storage"] == "f64"
    assert r32["error_code"] == 0 and len(r32["ev"]) == 4
    assert r32["reorthogonalizations"] > 0
    rel = np.abs(r32["ev"] - ref) / np.abs(ref)
    say("rel err", rel.tolist())
    assert rel.max() <= 1e-10
    assert r32["iterations"] <= r64["iterations"]


def test_small_sectors_report_fp64_storage():
    # fewer than 1000 states: an FP32 request is served (and reported) in FP64
    r = A.sparsediag(ring(10), nev=2, seed=1, dtype="f32")
    assert r["dim"] == 252 and r["storage"] == "f64" and r["reorthogonalizations"] == 0
