"""Host-side parts of the batched solver that need no device: the HBM
footprint used for admission."""
import math

import numpy as np

import alps_amd as A


def chain(L, n_up):
    s = np.arange(L)
    return A.SpinModel(L, s, (s + 1) % L, 1.0, 1.0, n_up)


def test_footprint_counts_the_kept_basis():
    m = chain(20, 10)
    D = math.comb(20, 10)
    stride = (D + 63) // 64 * 64
    f32 = A.batch_footprint(m, "f32")
    f64 = A.batch_footprint(m, "f64")
    # FP32: the whole basis (max_iter + 2 = 1002 slots) plus FP64 refinement
    # buffers (one Ritz vector and H y); FP64: a ring of three vectors
    assert f32 >= 1002 * stride * 4 + 2 * D * 8
    assert f32 < 1002 * stride * 4 + 2 * D * 8 + (8 << 20)
    assert 3 * stride * 8 <= f64 < 3 * stride * 8 + (8 << 20)
    # max_iter bounds the kept basis
    assert A.batch_footprint(m, "f32", max_iter=200) < f32 / 4


def test_small_sectors_use_fp64_storage():
    m = chain(10, 5)  # D = 252 < 1000
    assert A.batch_footprint(m, "f32") == A.batch_footprint(m, "f64")


def test_footprint_rejects_bad_models():
    import pytest

    with pytest.raises(A.EDError):
        A.batch_footprint(A.SpinModel(40, np.array([0]), np.array([1]), 1.0, 1.0, 20), "f32")
