"""run_iterative_solver with iteration callbacks (linear-solver/runner.hpp:17-84):
the reference's test suite (test/aux/iterative.hpp:25-181) ported onto the
GPU classes, with its `cg` and `gs` callbacks and the residual bounds of
test/shared-stencil.hpp:154-190; and the per-iteration Krylov instances
against the fused solves."""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu


def cg_callback(iter_num):
    """test/aux/iterative.hpp:25-45: stop after iter_num iterations or at
    rho == 0, returning the system's residual norm."""

    def op(iteration, inst):
        res, rho = inst.residual_norm(), inst.rho()
        assert np.isfinite(res)
        return res if (iteration > iter_num or rho == 0) else None

    return op


def gs_callback(lhs, sol, rhs, iter_num):
    """test/aux/iterative.hpp:47-65: ||A sol - rhs|| after every apply."""
    r = la.DenseVector(sol.n)

    def op(iteration, _inst):
        res = float(np.linalg.norm(lhs.residual(rhs, sol, r).numpy()))
        return res if iteration > iter_num else None

    return op


# (name, residual bound) of test/shared-stencil.hpp:157-168 (the same for
# both SOR variants)
BOUNDS = [("cg", 1e-12), ("bicgstab", 1e-12), ("sor", 1e-8), ("ssor1", 1e-13), ("sor5", 1e-13),
          ("ssor5", 1e-13), ("pcg_ssor1", 1e-13), ("pbicgstab_ssor1", 1e-12), ("pcg_amg", 1e-13),
          ("pbicgstab_amg", 1e-13)]


def reference_system():
    dims = (4, 9, 7)
    m = problems.diagonal_lines(dims)
    A = la.StencilMatrix.lookup(la.BoxInfo(dims), m, (0.0, 1.0, 2.0, 3.0))
    two = la.DenseVector(np.full(A.row_num, 2.0))
    return A, A * two


@pytest.mark.parametrize("ultra", [False, True])
@pytest.mark.parametrize("name,bound", BOUNDS)
def test_reference_suite_through_callbacks(name, bound, ultra):
    """test::suite (test/aux/iterative.hpp:67-181) with known solution 2."""
    A, b = reference_system()
    sol = la.DenseVector(A.row_num, fill=0.0)
    ssor1 = la.SsorSolver(1.0, ultra, 1).instantiate(A)
    amg = la.MultiGridSolver(ultra=ultra).instantiate(A)
    solvers = {
        "cg": (la.ConjugateGradientsSolver(8), 128),
        "bicgstab": (la.BiCgStabSolver(8), 128),
        "sor": (la.SorSolver(1.0, True, ultra), 128),
        "ssor1": (la.SsorSolver(1.0, ultra, 1), 128),
        "sor5": (la.SorSolver(1.0, True, ultra, iterations=5), 128),
        "ssor5": (la.SsorSolver(1.0, ultra, 5), 128),
        "pcg_ssor1": (la.ConjugateGradientsSolver(8, ssor1), 128),
        "pbicgstab_ssor1": (la.BiCgStabSolver(8, ssor1), 128),
        "pcg_amg": (la.ConjugateGradientsSolver(8, amg), 32),
        "pbicgstab_amg": (la.BiCgStabSolver(8, amg), 32),
    }
    solver, iters = solvers[name]
    if isinstance(solver, la._Krylov):
        op = cg_callback(iters)
    else:
        op = gs_callback(A, sol, b, iters)
    res = la.run_iterative_solver(solver, A, sol, b, op)
    err = np.abs(sol.numpy() - 2.0).max()
    assert res <= bound and err <= bound, (res, err)


@pytest.mark.parametrize("bicg", [False, True])
@pytest.mark.parametrize("precision", [None, np.float64, np.float32])
def test_system_instance_iterates_like_the_fused_solve(bicg, precision):
    """iterate()/residual_norm() of a system instance give the fused solve's
    residual history and solution bit for bit (the same kernels in the same
    order; the fused loop only queues each next iteration's head early)."""
    dims = (20, 16, 12)
    c = problems.lognormal(dims, 1.0, 3)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    b = A.rhs()
    M = None if precision is None else la.MultiGridSolver(precision=precision).instantiate(A)
    cls = la.BiCgStabSolver if bicg else la.ConjugateGradientsSolver
    x1 = la.DenseVector(A.row_num, fill=0.0)
    fused = cls(8, M).solve(A, x1, b, rel_tol=1e-10, max_it=300)
    assert fused.converged
    x2 = la.DenseVector(A.row_num, fill=0.0)
    sys = cls(8, M).instantiate(A).instantiate(x2, b)
    hist = [sys.residual_norm()]
    for _ in range(fused.iterations):
        sys.iterate()
        hist.append(sys.residual_norm())
    assert sys.iteration == fused.iterations
    np.testing.assert_array_equal(np.array(hist), fused.residuals)
    np.testing.assert_array_equal(x2.numpy(), x1.numpy())
    assert sys.solution() is x2


def test_two_system_instances_iterate_in_turn():
    """Each instance owns its vectors and scalars: interleaved iterations of
    two systems equal their separate runs."""
    dims = (12, 10, 9)
    A = la.StencilMatrix.values(la.BoxInfo(dims), problems.lognormal(dims, 1.0, 4))
    b1 = A.rhs()
    b2 = la.DenseVector(problems.uniform(A.row_num, 5))
    M = la.MultiGridSolver().instantiate(A)
    solver = la.ConjugateGradientsSolver(8, M).instantiate(A)
    xa, xb = la.DenseVector(A.row_num, fill=0.0), la.DenseVector(A.row_num, fill=0.0)
    sa, sb = solver.instantiate(xa, b1), solver.instantiate(xb, b2)
    for _ in range(6):
        sa.iterate()
        sb.iterate()
    ya = la.DenseVector(A.row_num, fill=0.0)
    ref = solver.instantiate(ya, b1)
    for _ in range(6):
        ref.iterate()
    np.testing.assert_array_equal(xa.numpy(), ya.numpy())
    assert sa.rho() == ref.rho() and sa.residual_norm() == ref.residual_norm()


def test_multigrid_as_a_stationary_solver():
    """MultiGridSolver in run_iterative_solver's stationary branch: each apply
    is one cycle from the current iterate (multigrid/solver.hpp:41-46),
    equal to the oracle's cycle from that iterate."""
    dims = (16, 12, 10)
    c = problems.lognormal(dims, 1.0, 6)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    S = O.Stencil(dims, coeff=c)
    H = O.Hierarchy(S, coarse_segment=64)
    b = A.rhs()
    sol = la.DenseVector(A.row_num, fill=0.0)
    x = np.zeros(A.row_num)
    bh = S.rhs()
    seen = []

    def op(iteration, _inst):
        nonlocal x
        x = H.vcycle(bh, x_in=x)
        np.testing.assert_array_equal(sol.numpy(), x)
        seen.append(iteration)
        return float(np.linalg.norm(O.residual(S, bh, x))) if iteration >= 12 else None

    res = la.run_iterative_solver(la.MultiGridSolver(), A, sol, b, op)
    assert seen == list(range(1, 13)) and res <= 1e-2 * np.linalg.norm(bh)
