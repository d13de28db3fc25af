"""The solve-phase workloads at their full sizes:

* 64^3 lognormal (sigma 1) PCG-AMG to 1e-8 against the oracle's PCG-AMG, and
  sigma 2 in mixed precision against the oracle's FP64 solution;
* the 256^3 operator and smoother kernels (matrix-free and CSR, f64 and f32)
  from seeded uniform[-1, 1] vectors, bit for bit against the oracle, and one
  FP32 V-cycle at 256^3 against the oracle's;
* the 256^3 two-phase field (spheres of radius 4-12, 35% solid, coefficients
  1 and 1e-6): FP64 AMG-CG, and BiCGStab with upwind advection at Peclet 10,
  checked by their true residuals;
* 512^3 lognormal sigma 2: the mixed-precision AMG-CG (FP32 hierarchy inside
  FP64 CG) and the FP64 one, checked by the true FP64 residual, their
  agreement and their iteration counts.
"""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu


def true_rel_residual(A, b, x):
    r = la.DenseVector(b.n)
    A.residual(b, x, r)
    return float(np.linalg.norm(r.numpy()) / np.linalg.norm(b.numpy()))


# ---------------------------------------------------------------------------
# 64^3


def test_pcg_amg_64_lognormal_against_oracle():
    dims = (64, 64, 64)
    c = problems.lognormal(dims, 1.0, 1)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    S = O.Stencil(dims, coeff=c)
    b = S.rhs()
    H = O.Hierarchy(S, coarse_segment=64)
    x_ref, it_ref, _ = O.pcg(S, b, H, precon=1, rel_tol=1e-8, max_it=500)
    bv = la.DenseVector(b)
    for prec in (np.float64, np.float32):
        M = la.MultiGridSolver(precision=prec).instantiate(A)
        x = la.DenseVector(S.N, fill=0.0)
        res = la.ConjugateGradientsSolver(8, M).solve(A, x, bv, rel_tol=1e-8, max_it=500)
        assert res.converged and abs(res.iterations - it_ref) <= 1, (res.iterations, it_ref)
        err = np.abs(x.numpy() - x_ref).max() / np.abs(x_ref).max()
        assert err <= (1e-9 if prec == np.float64 else 1e-7), err
        assert true_rel_residual(A, bv, x) <= 1.01e-8


def test_mixed_precision_64_sigma2_against_fp64_oracle():
    """sigma 2 (contrast e^+-8 at two standard deviations): the mixed solve's
    relative error against the oracle's FP64 solution stays below 1e-7."""
    dims = (64, 64, 64)
    c = problems.lognormal(dims, 2.0, 1)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    S = O.Stencil(dims, coeff=c)
    b = S.rhs()
    x_ref, it_ref, _ = O.pcg(S, b, O.Hierarchy(S, coarse_segment=64), precon=1, rel_tol=1e-8, max_it=500)
    M = la.MultiGridSolver(precision=np.float32).instantiate(A)
    x = la.DenseVector(S.N, fill=0.0)
    bv = la.DenseVector(b)
    res = la.ConjugateGradientsSolver(8, M).solve(A, x, bv, rel_tol=1e-8, max_it=500)
    assert res.converged and abs(res.iterations - it_ref) <= 2
    err = np.abs(x.numpy() - x_ref).max() / np.abs(x_ref).max()
    assert err <= 1e-7, err
    assert true_rel_residual(A, bv, x) <= 1.01e-8


# ---------------------------------------------------------------------------
# 256^3 kernels


@pytest.fixture(scope="module")
def field256():
    dims = (256, 256, 256)
    c = problems.lognormal(dims, 1.0, 1)
    return dims, c, la.StencilMatrix.values(la.BoxInfo(dims), c), O.Stencil(dims, coeff=c)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_256_matrix_free_and_csr(field256, dtype):
    dims, c, A, S = field256
    x = problems.uniform(S.N, 2).astype(dtype)
    xv = la.DenseVector(x)
    np.testing.assert_array_equal((A * xv).numpy(), O.spmv(S, x))
    C = A.to_csr(dtype)
    Cr, _, _ = S.to_csr()
    if dtype == np.float32:
        Cr = O.Csr(Cr.rowptr, Cr.cols, Cr.vals.astype(np.float32).astype(np.float64), is_f32=True)
    np.testing.assert_array_equal((C * xv).numpy(), O.spmv(Cr, x))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_jacobi_256_ten_sweeps(field256, dtype):
    """10 Jacobi sweeps (omega 2/3) chained from uniform[-1, 1]."""
    dims, c, A, S = field256
    x = problems.uniform(S.N, 2).astype(dtype)
    b = S.rhs(dtype)
    bv, xa, xb = la.DenseVector(b), la.DenseVector(x), la.DenseVector(S.N, dtype)
    ref = x
    for _ in range(10):
        A.jacobi(xa, xb, bv, 2.0 / 3.0)
        xa, xb = xb, xa
        ref = O.jacobi(S, ref, b, 2.0 / 3.0)
    np.testing.assert_array_equal(xa.numpy(), ref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("cache", [True, False])
def test_sor_256_ten_sweeps(field256, dtype, cache):
    """10 lexicographic SOR sweeps (omega 1), alternating forward and backward,
    with the cached face weights and diagonal and with them evaluated on the
    fly."""
    dims, c, A, S = field256
    A.set_sor_cache(cache)
    x = problems.uniform(S.N, 2).astype(dtype)
    b = S.rhs(dtype)
    bv, xa, xb = la.DenseVector(b), la.DenseVector(x), la.DenseVector(S.N, dtype)
    ref = x
    for k in range(10):
        fwd = k % 2 == 0
        A.sor(xa, xb, bv, 1.0, fwd)
        xa, xb = xb, xa
        ref = O.sor(S, ref, b, 1.0, fwd)
    np.testing.assert_array_equal(xa.numpy(), ref)
    A.set_sor_cache(True)


@pytest.mark.timeout(600)
def test_fp32_vcycle_256_equals_oracle(field256):
    """The FP32 hierarchy and one FP32 V-cycle at 256^3, bit for bit."""
    dims, c, A, S = field256
    M = la.MultiGridSolver(precision=np.float32).instantiate(A)
    H = O.Hierarchy(S, coarse_f32=True, coarse_segment=64)
    assert [M.level_size(l) for l in range(M.level_num)] == [H.size(l) for l in range(H.level_num)]
    r = problems.uniform(S.N, 4).astype(np.float32)
    z = M.apply(la.DenseVector(r), la.DenseVector(S.N, np.float32)).numpy()
    np.testing.assert_array_equal(z, H.vcycle(r))


# ---------------------------------------------------------------------------
# 256^3 two-phase field


@pytest.fixture(scope="module")
def spheres256():
    dims = (256, 256, 256)
    m = problems.spheres(dims, 3)
    assert 0.35 <= m.mean() < 0.37
    return dims, m


def test_two_phase_256_fp64_amg_cg(spheres256):
    dims, m = spheres256
    A = la.StencilMatrix.lookup(la.BoxInfo(dims), m, (1.0, 1e-6, 0.0, 0.0))
    b = A.rhs()
    M = la.MultiGridSolver().instantiate(A)
    x = la.DenseVector(A.row_num, fill=0.0)
    res = la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=1000)
    assert res.converged and res.iterations < 200
    assert true_rel_residual(A, b, x) <= 1.01e-8


def test_two_phase_256_advection_bicgstab(spheres256):
    """Upwind advection v = 10 / L_x (Pe = v L_x / D_ref = 10 with D_ref = 1);
    BiCGStab preconditioned with the AMG of the nonsymmetric operator (LU
    coarsest solve).  BiCGStab stops on its recursive residual; the true
    residual of its iterate stays within twice the tolerance."""
    dims, m = spheres256
    A = la.StencilMatrix.lookup(la.BoxInfo(dims, velocity_x=10.0 / dims[0]), m, (1.0, 1e-6, 0.0, 0.0))
    b = A.rhs()
    M = la.MultiGridSolver().instantiate(A)
    x = la.DenseVector(A.row_num, fill=0.0)
    res = la.BiCgStabSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=1000)
    assert res.converged and res.iterations < 200
    assert true_rel_residual(A, b, x) <= 2e-8


# ---------------------------------------------------------------------------
# 512^3 lognormal sigma 2


@pytest.fixture(scope="module")
def lognormal512():
    dims = (512, 512, 512)
    c = problems.lognormal(dims, 2.0, 1)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    del c
    return A, A.rhs(), {}


@pytest.mark.timeout(600)
def test_mixed_512_sigma2(lognormal512):
    A, b, keep = lognormal512
    M = la.MultiGridSolver(precision=np.float32).instantiate(A)
    x = la.DenseVector(A.row_num, fill=0.0)
    res = la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=1000)
    del M
    assert res.converged
    assert true_rel_residual(A, b, x) <= 1.01e-8
    keep["mixed"] = (x.numpy(), res.iterations)


@pytest.mark.timeout(600)
def test_fp64_512_sigma2_agrees_with_mixed(lognormal512):
    A, b, keep = lognormal512
    assert "mixed" in keep, "needs test_mixed_512_sigma2 first"
    M = la.MultiGridSolver(precision=np.float64).instantiate(A)
    x = la.DenseVector(A.row_num, fill=0.0)
    res = la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=1000)
    del M
    assert res.converged
    assert true_rel_residual(A, b, x) <= 1.01e-8
    xm, it_m = keep.pop("mixed")
    xf = x.numpy()
    err = np.abs(xm - xf).max() / np.abs(xf).max()
    print(f"\n512^3 sigma 2: mixed {it_m} / FP64 {res.iterations} iterations, rel. difference {err:.2e}",
          flush=True)
    assert abs(it_m - res.iterations) <= 2 and err <= 1e-7
