"""CSR with block values on the GPU against the oracle, bit for bit: the
reference's 73 x 73 system of 3 x 3 diagonal blocks (test/shared-blocked.cpp)
and random blocks of sizes 1-4 in f64 and f32."""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O

from test_oracle import blocked_reference_product, blocked_reference_system

pytestmark = pytest.mark.gpu


def test_bcsr_reference_system():
    rp, cols, vals = blocked_reference_system()
    A = la.BlockCsrMatrix(rp, cols, vals, np.float32)
    y = (A * la.DenseVector(np.ones(73 * 3, np.float32))).numpy()
    np.testing.assert_array_equal(y, blocked_reference_product())


@pytest.mark.parametrize("bsz", [1, 2, 3, 4])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bcsr_spmv_equals_oracle(bsz, dtype):
    rng = np.random.default_rng(10 + bsz)
    nb = 5000
    lens = rng.integers(0, 9, nb)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    cols = rng.integers(0, nb, int(rp[-1])).astype(np.int64)
    vals = rng.standard_normal((cols.size, bsz, bsz))
    x = rng.standard_normal(nb * bsz).astype(dtype)
    A = la.BlockCsrMatrix(rp, cols, vals, dtype)
    y = (A * la.DenseVector(x)).numpy()
    np.testing.assert_array_equal(y, O.bcsr_spmv(rp, cols, vals, x))


def test_bcsr_rejects_bad_input():
    with pytest.raises(la.LinealError):
        la.BlockCsrMatrix(np.array([0, 1]), np.array([3]), np.ones((1, 2, 2)))
    with pytest.raises(la.LinealError):
        la.BlockCsrMatrix(np.array([0, 1]), np.array([0]), np.ones((1, 5, 5)))
