"""GPU parity of SSOR (gauss-seidel-variant/solver.hpp:45-96) as a stationary
solver and as a Krylov preconditioner (test/aux/iterative.hpp:122-165)."""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu


def same(a, b):
    assert a.shape == b.shape
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


def pair(dims, kind):
    if kind == "values":
        c = problems.lognormal(dims, 1.0, 6)
        return la.StencilMatrix.values(la.BoxInfo(dims), c), O.Stencil(dims, coeff=c)
    if kind == "lookup":
        m = problems.diagonal_lines(dims)
        return la.StencilMatrix.lookup(la.BoxInfo(dims), m), O.Stencil(dims, materials=m)
    c = problems.lognormal(dims, 1.0, 6)
    S = O.Stencil(dims, coeff=c)
    return la.StencilMatrix.values(la.BoxInfo(dims), c).to_csr(), S.to_csr()[0]


@pytest.mark.parametrize("dims", [(4, 9, 7), (13, 17, 5), (33, 17, 65)])
@pytest.mark.parametrize("kind", ["values", "lookup", "csr"])
@pytest.mark.parametrize("ultra", [False, True])
def test_ssor_matches_oracle(dims, kind, ultra):
    A, S = pair(dims, kind)
    n = int(np.prod(dims))
    x = problems.uniform(n, 2)
    b = problems.uniform(n, 3)
    for iters in (1, 3):
        inst = la.SsorSolver(1.2, ultra, iters).instantiate(A)
        out = inst.apply(la.DenseVector(x), la.DenseVector(n), la.DenseVector(b))
        same(out.numpy(), O.ssor(S, x, b, 1.2, ultra, iters))
    # from zero, and in place
    inst = la.SsorSolver(1.2, ultra).instantiate(A)
    same(inst.apply(None, la.DenseVector(n), la.DenseVector(b)).numpy(), O.ssor(S, None, b, 1.2, ultra))
    v = la.DenseVector(x)
    inst.apply(v, v, la.DenseVector(b))
    same(v.numpy(), O.ssor(S, x, b, 1.2, ultra))


@pytest.mark.parametrize("kind", ["values", "lookup"])
def test_ssor_f32_matches_oracle_sweeps(kind):
    dims = (13, 17, 5)
    A, S = pair(dims, kind)
    n = int(np.prod(dims))
    x = problems.uniform(n, 4, np.float32)
    b = problems.uniform(n, 5, np.float32)
    out = la.SsorSolver(1.0).instantiate(A).apply(la.DenseVector(x), la.DenseVector(n, np.float32),
                                                  la.DenseVector(b))
    w = np.float32(1.0)
    same(out.numpy(), O.sor(S, O.sor(S, x, b, w, True, False), b, w, False, False))


# Krylov results: the GPU reduces dot products in another order than the
# oracle's sequential sums, and the rounding difference grows over the
# iterations.  CG's history agrees to 1e-7 relative (7e-9 seen after 40
# iterations here).  BiCGStab's erratic residual amplifies it more (a factor
# 2 in late iterations), so its first ten entries agree to 1e-6.  Solutions
# agree to 1e-7 of their maximum, and iteration counts to within one.
@pytest.mark.parametrize("method", ["cg", "bicgstab"])
@pytest.mark.parametrize("ultra", [False, True])
def test_ssor_preconditioned_krylov_matches_oracle(method, ultra):
    dims = (13, 17, 11)
    A, S = pair(dims, "values")
    n = int(np.prod(dims))
    b = S.rhs()
    P = la.SsorSolver(1.0, ultra).instantiate(A)
    x = la.DenseVector(n, fill=0.0)
    solver = la.ConjugateGradientsSolver(8, P) if method == "cg" else la.BiCgStabSolver(8, P)
    res = solver.solve(A, x, la.DenseVector(b), rel_tol=1e-10, max_it=300)
    fn = O.pcg if method == "cg" else O.pbicgstab
    xo, it, hist = fn(S, b, None, precon=3, ultra=ultra, rel_tol=1e-10, max_it=300)
    assert res.converged and abs(res.iterations - it) <= 1
    k = min(len(hist), len(res.residuals)) if method == "cg" else 10
    np.testing.assert_allclose(res.residuals[:k], hist[:k], rtol=1e-7 if method == "cg" else 1e-6, atol=0)
    assert np.abs(x.numpy() - xo).max() <= 1e-7 * np.abs(xo).max()


@pytest.mark.parametrize("ultra", [False, True])
def test_reference_thresholds_ssor_on_gpu(ultra):
    """test/shared-stencil.hpp:198-224 with test/aux/iterative.hpp:122-165: the
    4x9x7 diagonal-lines system with its Dirichlet rhs; ssor1 / ssor5 after
    129 applications within 1.1e-14, pcg_ssor1 / pbicgstab_ssor1 within
    1e-13."""
    dims = (4, 9, 7)
    A, S = pair(dims, "lookup")
    n = int(np.prod(dims))
    b = A.rhs()
    r = la.DenseVector(n)
    for iters in (1, 5):
        inst = la.SsorSolver(1.0, ultra, iters).instantiate(A)
        x = la.DenseVector(n, fill=0.0)
        for _ in range(129):
            inst.apply(x, x, b)
        A.residual(b, x, r)
        rv = r.numpy()
        assert np.linalg.norm(rv) <= 1.1e-14 and np.abs(rv).max() <= 1.1e-14
    for cls in (la.ConjugateGradientsSolver, la.BiCgStabSolver):
        P = la.SsorSolver(1.0, ultra).instantiate(A)
        x = la.DenseVector(n, fill=0.0)
        res = cls(8, P).solve(A, x, b, rel_tol=1e-300, max_it=129)
        A.residual(b, x, r)
        assert res.residuals[-1] <= 1e-13 and np.abs(r.numpy()).max() <= 1e-13
