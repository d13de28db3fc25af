"""The sticky wait-error word of the wavefront sweeps: an expired bounded
wait (forced here from the host, so that every kernel still exits normally)
is reported as SYNC_TIMEOUT by the entry point that ran the sweep, the word
is cleared, and the next sweep is exact again."""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu


def _expect_timeout(fn):
    with pytest.raises(la.LinealError) as e:
        fn()
    assert e.value.status == la.SYNC_TIMEOUT


@pytest.mark.parametrize("kind", ["stencil", "csr"])
def test_sor_reports_and_clears_the_wait_error(kind):
    dims = (12, 10, 9)
    c = problems.lognormal(dims, 1.0, 3)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    S = O.Stencil(dims, coeff=c)
    if kind == "csr":
        A = A.to_csr()
    b = S.rhs()
    x = problems.uniform(S.N, 4)
    out = la.DenseVector(S.N)
    A.ctx.set_wait_error()
    _expect_timeout(lambda: A.sor(la.DenseVector(x), out, la.DenseVector(b), 1.0, True))
    # the word was cleared: the next sweeps are exact
    for fwd in (True, False):
        got = A.sor(la.DenseVector(x), out, la.DenseVector(b), 1.0, fwd).numpy()
        np.testing.assert_array_equal(got, O.sor(S, x, b, 1.0, fwd))


def test_krylov_and_vcycle_report_and_clear_the_wait_error():
    dims = (16, 12, 10)
    c = problems.lognormal(dims, 1.0, 5)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    S = O.Stencil(dims, coeff=c)
    b = la.DenseVector(S.N)
    b.upload(S.rhs())
    M = la.MultiGridSolver(coarse_segment_rows=0).instantiate(A)
    ctx = A.ctx
    ctx.set_wait_error()
    _expect_timeout(lambda: la.ConjugateGradientsSolver(8, M).solve(A, la.DenseVector(S.N, fill=0.0), b))
    ctx.set_wait_error()
    _expect_timeout(lambda: M.apply(b, la.DenseVector(S.N)))
    # cleared: the cycle equals the oracle's and the solve converges
    H = O.Hierarchy(S)
    r = S.rhs()
    np.testing.assert_array_equal(M.apply(b, la.DenseVector(S.N)).numpy(), H.vcycle(r))
    x = la.DenseVector(S.N, fill=0.0)
    res = la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=200)
    assert res.converged
