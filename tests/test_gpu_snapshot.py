"""Snapshot and restore of a device Lanczos recurrence (alps_ed_lanczos_save
/ load; the reference archives its recurrence with lanczos::save / load,
src/ietl/lanczos.h:86-98, and continues it with more_eigenvalues).

A run interrupted by a snapshot, with every device object destroyed and the
recurrence continued from the snapshot in a fresh object, must give the
uninterrupted run's alpha and beta bit for bit, and the same Ritz vectors
when the basis is kept.  Mismatched snapshots are rejected."""
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ring(L, jz=1.0):
    om = O.lattice_model("chain lattice", L, jz=jz)
    return A.SpinModel(om.n_sites, om.bond_s, om.bond_t, om.jz, om.jxy, L // 2)


def bits(x):
    return np.asarray(x, dtype=np.float64).view(np.int64)


def run_split(make_op, dtype, keep, reorth, first, total, seed=3):
    op = make_op()
    ref = A.Lanczos(op, dtype, max_steps=total + 1, keep_basis=keep, reorth=reorth)
    ref.start(seed=seed)
    ref.steps(total - 1)
    a_ref, b_ref = ref.coefficients()
    lz = A.Lanczos(op, dtype, max_steps=total + 1, keep_basis=keep, reorth=reorth)
    lz.start(seed=seed)
    lz.steps(first - 1)
    snap = lz.save()
    lz.close()
    op.close()
    op2 = make_op()
    lz2 = A.Lanczos(op2, dtype, max_steps=total + 1, keep_basis=keep, reorth=reorth)
    lz2.load(snap)
    lz2.steps(total - first)
    a, b = lz2.coefficients()
    return ref, lz2, (a_ref, b_ref), (a, b), len(snap)


@pytest.mark.parametrize("dtype,keep,reorth", [("f32", False, False), ("f64", False, False),
                                               ("f32", True, False), ("f32", True, True)])
def test_snapshot_resumes_bit_for_bit(dtype, keep, reorth):
    ref, lz2, (a_ref, b_ref), (a, b), nbytes = run_split(lambda: A.Operator(ring(18)), dtype, keep, reorth, 40, 110)
    say(dtype, keep, reorth, "snapshot bytes", nbytes, "steps", len(a))
    assert len(a) == len(a_ref) == 110
    assert np.array_equal(bits(a), bits(a_ref)) and np.array_equal(bits(b), bits(b_ref))
    if reorth:
        assert lz2.reorthogonalizations() == ref.reorthogonalizations() > 0
    if keep:
        c = np.zeros((1, 110))
        c[0, :] = 1.0 / np.sqrt(110.0)
        (v_ref,) = ref.ritz(c)
        (v,) = lz2.ritz(c)
        assert np.array_equal(bits(v.numpy()), bits(v_ref.numpy()))


def test_snapshot_in_a_momentum_sector():
    L = 16
    s = np.arange(L)
    m = A.SpinModel(L, s, (s + 1) % L, 1.0, 1.0, L // 2)
    perms = np.array([[(i + j) % L for i in range(L)] for j in range(L)])
    phases = np.exp(2j * np.pi * 3 * np.arange(L) / L)
    make = lambda: A.Operator(m, translations=(phases, perms))
    _, _, (a_ref, b_ref), (a, b), _ = run_split(make, "f32", False, False, 25, 80)
    assert np.array_equal(bits(a), bits(a_ref)) and np.array_equal(bits(b), bits(b_ref))


def test_snapshot_rejections():
    op = A.Operator(ring(14))
    lz = A.Lanczos(op, "f32", max_steps=50)
    lz.start(seed=1)
    lz.steps(10)
    snap = lz.save()
    # a started object
    with pytest.raises(A.EDError) as e:
        lz.load(snap)
    assert e.value.status == -5
    # another model (same dimension), dtype, max_steps, keep_basis
    other = A.Operator(ring(14, jz=0.5))
    cases = [A.Lanczos(other, "f32", max_steps=50), A.Lanczos(op, "f64", max_steps=50),
             A.Lanczos(op, "f32", max_steps=60), A.Lanczos(op, "f32", max_steps=50, keep_basis=True)]
    for fresh in cases:
        with pytest.raises(A.EDError) as e:
            fresh.load(snap)
        assert e.value.status == -1
    fresh = A.Lanczos(op, "f32", max_steps=50)
    for bad in (snap[: len(snap) // 2], b"x" * len(snap), snap[:10]):
        with pytest.raises(A.EDError) as e:
            fresh.load(bad)
        assert e.value.status == -1
    fresh.load(snap)  # the intact snapshot still loads into a fresh object
    fresh.steps(5)
