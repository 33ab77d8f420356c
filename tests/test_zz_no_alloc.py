"""The allocation-free Lanczos step (alps_ed_lanczos_step_vectors_noalloc,
static device scratch) against the allocating entry point it mirrors
(alps_ed_lanczos_step_vectors): the same kernels in the same order, so alpha,
beta and the output vector must agree bit for bit.  Two calls in a row on
different operators, the second one smaller, so the second call runs on the
scratch the first one left behind.  Also: an operator whose H.v needs more
partials than the scratch holds is rejected with an argument error."""
import sys

import numpy as np
import pytest

import alps_amd as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def product_model(om, n_up):
    return A.SpinModel(om.n_sites, om.bond_s, om.bond_t, om.jz, om.jxy, n_up, k=om.k, h=om.h, d=om.d)


def xxz_ring(L, seed):
    rng = np.random.default_rng(seed)
    s = np.arange(L)
    J = rng.uniform(0.5, 1.5, L)
    return O.Model(L, s, (s + 1) % L, rng.uniform(0, 2) * J, J, h=rng.uniform(0, 0.5))


def step_inputs(n, seed):
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(-1, 1, n)
    v0 /= np.linalg.norm(v0)
    v1 = rng.uniform(-1, 1, n)
    v1 -= (v1 @ v0) * v0
    v1 /= np.linalg.norm(v1)
    return v0, v1


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_step_matches_allocating_step(dtype):
    # larger operator first (L=20, D=184,756), then a smaller one (L=14)
    for L, seed in ((20, 1), (14, 2)):
        say("case", dtype, "L", L)
        om = xxz_ring(L, seed)
        op = A.Operator(product_model(om, L // 2))
        n = op.dim
        v0, v1 = step_inputs(n, seed + 10)
        b0 = 0.41
        P, C = A.Vector.from_numpy(v0, dtype), A.Vector.from_numpy(v1, dtype)
        W_ref, W = A.Vector(n, dtype), A.Vector(n, dtype)
        a_ref, b_ref = A.lanczos_step_vectors(op, P, C, W_ref, b0)
        a, b = A.lanczos_step_vectors(op, P, C, W, b0, noalloc=True)
        say("alpha", a, a_ref, "beta", b, b_ref)
        assert a == a_ref and b == b_ref
        assert np.array_equal(W.numpy().view(np.int64), W_ref.numpy().view(np.int64))
        # and against the oracle's H.v in FP64 (the step's own arithmetic)
        if dtype == "f64":
            w = O.spmv(om, L // 2, v1)
            ar = v1 @ w
            w = w - ar * v1 - b0 * v0
            assert abs(a - ar) < 1e-11 and abs(b - np.linalg.norm(w)) < 1e-11


def test_noalloc_rejects_operator_beyond_scratch():
    # N=31, 15 up spins: 522,937 segments of the split ranking (high words of
    # 19 bits with 3..15 set bits), more H.v partials than the static scratch
    # holds; the operator is checked before the vectors
    N = 31
    s = np.arange(N)
    op = A.Operator(A.SpinModel(N, s, (s + 1) % N, 1.0, 1.0, 15))
    P, C, W = (A.Vector(1, "f32") for _ in range(3))
    with pytest.raises(A.EDError) as e:
        A.lanczos_step_vectors(op, P, C, W, 0.0, noalloc=True)
    assert e.value.status == -1 and "static scratch" in str(e.value)
