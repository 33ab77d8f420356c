"""Failure handling of the solves, forced from the host so that every kernel
still ends normally: an exactly zero beta (the recurrence reaches an
invariant subspace), non-finite couplings (NaN alpha), and a sector smaller
than the requested iteration count.

The reference's recurrence checks none of these (src/ietl/lanczos.h:392-406:
no beta = 0 test; NaN reaches LAPACK, which fails with "LAPACK error, stev
function failed", tmatrix.h:144-145).  Here a zero beta ends the solve with
the exact T matrix, and non-finite coefficients are reported as a solver
error (single solve) or error_code -2 (batch)."""
import math
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def chain(L, jz=1.0, jxy=1.0, n_up=None, h=0.0):
    s = np.arange(L)
    return A.SpinModel(L, s, (s + 1) % L, jz, jxy, L // 2 if n_up is None else n_up, h=h)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("fixed", [False, True])
def test_zero_operator_ends_at_the_invariant_subspace(dtype, fixed):
    # H = 0: w = H v0 = 0, alpha_0 = 0 and beta_0 = 0 exactly; T = [0] is exact
    m = chain(12, jz=0.0, jxy=0.0)
    r = A.sparsediag(m, nev=2, seed=3, dtype=dtype, fixed=fixed, max_iter=40)
    say(dtype, fixed, r)
    assert r["dim"] == math.comb(12, 6)
    assert r["error_code"] == 0 and r["iterations"] == 1
    assert list(r["ev"]) == [0.0]


def test_nan_coupling_is_a_solver_error():
    jz = np.ones(12)
    jz[3] = np.nan
    with pytest.raises(A.EDError) as e:
        A.sparsediag(chain(12, jz=jz), nev=1, seed=1, dtype="f64")
    say(e.value)
    assert e.value.status == -4 and "non-finite" in str(e.value)


def test_batch_nan_and_zero_problems_next_to_a_normal_one():
    jz = np.ones(12)
    jz[5] = np.nan
    b = A.Batch("f64", hbm_bytes=1 << 30)
    t_nan = b.submit(chain(12, jz=jz), nev=1, seed=1)
    t_zero = b.submit(chain(12, jz=0.0, jxy=0.0), nev=1, seed=1)
    t_ok = b.submit(chain(12), nev=1, seed=1)
    b.run()
    r_nan, r_zero, r_ok = b.result(t_nan), b.result(t_zero), b.result(t_ok)
    say(r_nan, r_zero, r_ok)
    assert r_nan["error_code"] == -2 and len(r_nan["ev"]) == 0
    assert r_zero["error_code"] == 0 and list(r_zero["ev"]) == [0.0] and r_zero["iterations"] == 1
    alone = A.sparsediag(chain(12), nev=1, seed=1, dtype="f64")
    assert r_ok["error_code"] == 0 and abs(r_ok["ev"][0] - alone["ev"][0]) < 1e-12
    b.close()


def test_tiny_sector_fixed_iterations_past_its_dimension():
    # ring of four, Sz=0: D = 6 < 50 fixed iterations; the recurrence runs on
    # rounding noise past the Krylov dimension, as the reference's does, and
    # the ghost-filtered T eigenvalues are the exact levels (E0 = -2, then -1).
    # Ending at max_iter is error code 1, as fixed_lanczos_iteration reports
    # it (iteration.h:74-84,209-216).
    om = O.lattice_model("chain lattice", 4)
    m = A.SpinModel(om.n_sites, om.bond_s, om.bond_t, om.jz, om.jxy, 2)
    r = A.sparsediag(m, nev=2, seed=1, dtype="f64", fixed=True, max_iter=50)
    say(r)
    assert r["dim"] == 6 and r["iterations"] == 50 and r["error_code"] == 1
    dense = np.unique(np.round(np.linalg.eigvalsh(O.dense(om, 2)), 10))
    np.testing.assert_allclose(r["ev"], dense[:2], atol=1e-10)


def test_operator_destroyed_before_its_lanczos_object():
    # the caller destroys the operator while a Lanczos object still uses it:
    # the object keeps it alive until the object is destroyed, and the
    # recurrence and Ritz vectors are those of an undisturbed run
    def run(close_early):
        op = A.Operator(chain(12))
        lz = A.Lanczos(op, "f64", max_steps=40, keep_basis=True)
        lz.start(seed=5)
        if close_early:
            op.close()
            other = A.Operator(chain(14, jz=0.3))  # reuses freed memory if the operator were gone
        lz.steps(30)
        a, b = lz.coefficients()
        (v,) = lz.ritz(np.full((1, 30), 1.0 / math.sqrt(30.0)))
        out = (a.copy(), b.copy(), v.numpy().copy())
        lz.close()
        if not close_early:
            op.close()
        return out

    ref, early = run(False), run(True)
    for x, y in zip(ref, early):
        assert np.array_equal(x, y)
