"""GPU parity of the AMG hierarchy, the V-cycle and the Krylov solvers."""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu


def pair(dims, kind="values", sigma=1.0, seed=1):
    box = la.BoxInfo(dims)
    if kind == "values":
        c = problems.lognormal(dims, sigma, seed)
        return la.StencilMatrix.values(box, c), O.Stencil(dims, coeff=c)
    if kind == "spheres":
        m = problems.spheres(dims, seed, rmin=2, rmax=4)
        mc = (1.0, 1e-6, 0.0, 0.0)
    else:
        m = problems.diagonal_lines(dims)
        mc = (0.0, 1.0, 2.0, 3.0)
    return la.StencilMatrix.lookup(box, m, mc), O.Stencil(dims, materials=m, mat_coeffs=mc)


CASES = [((4, 9, 7), "lookup"), ((24, 20, 16), "values"), ((32, 32, 32), "values"), ((24, 24, 24), "spheres")]


@pytest.mark.parametrize("dims,kind", CASES)
@pytest.mark.parametrize("f32", [False, True])
def test_hierarchy_equals_oracle(dims, kind, f32):
    A, S = pair(dims, kind)
    H = O.Hierarchy(S, coarse_f32=f32, coarse_segment=64)
    M = la.MultiGridSolver(precision=np.float32 if f32 else np.float64).instantiate(A)
    assert M.level_num == H.level_num
    for l in range(H.level_num):
        assert M.level_size(l) == H.size(l)
    for l in range(H.level_num - 1):
        np.testing.assert_array_equal(M.aggregates(l), H.agg_map(l))
    for l in range(1, H.level_num):
        rp, cols, vals = M.matrix(l)
        ref = H.matrix(l)
        np.testing.assert_array_equal(rp, ref.rowptr)
        np.testing.assert_array_equal(cols, ref.cols)
        np.testing.assert_array_equal(vals, ref.vals)


@pytest.mark.parametrize("dims,kind", CASES)
@pytest.mark.parametrize("ultra", [False, True])
def test_vcycle_equals_oracle(dims, kind, ultra):
    A, S = pair(dims, kind)
    for f32 in (False, True):
        H = O.Hierarchy(S, coarse_f32=f32, coarse_segment=64)
        M = la.MultiGridSolver(precision=np.float32 if f32 else np.float64, ultra=ultra).instantiate(A)
        r = problems.uniform(S.N, 5)
        if f32:
            r32 = r.astype(np.float32)
            z = M.apply(la.DenseVector(r32), la.DenseVector(S.N, np.float32)).numpy()
            ref = H.vcycle(r32, ultra)
        else:
            z = M.apply(la.DenseVector(r), la.DenseVector(S.N)).numpy()
            ref = H.vcycle(r, ultra)
        np.testing.assert_array_equal(z, ref)


@pytest.mark.parametrize("dims,kind", [((24, 20, 16), "values"), ((24, 24, 24), "spheres"), ((40, 40, 40), "values")])
@pytest.mark.parametrize("segments", [(64, 0), (64, 4096), (200, 0), (1000, 300), (0, 0)])
@pytest.mark.parametrize("ultra", [False, True])
def test_vcycle_segment_policies(dims, kind, segments, ultra):
    """The coarse smoothers' segments: coarse_segment_rows rows per segment (64:
    one wavefront per segment; longer: the scheduled sweep) and one segment on
    the levels of at most sequential_rows rows; the cycle equals the oracle's
    with the same segments bit for bit, f64 and f32."""
    seg, seq = segments
    A, S = pair(dims, kind)
    for f32 in (False, True):
        prec = np.float32 if f32 else np.float64
        H = O.Hierarchy(S, coarse_f32=f32, coarse_segment=seg, sequential_rows=seq)
        M = la.MultiGridSolver(precision=prec, ultra=ultra, coarse_segment_rows=seg,
                               sequential_rows=seq).instantiate(A)
        r = problems.uniform(S.N, 9).astype(prec)
        z = M.apply(la.DenseVector(r), la.DenseVector(S.N, prec)).numpy()
        np.testing.assert_array_equal(z, H.vcycle(r, ultra))


@pytest.mark.parametrize("dims,kind", CASES[:2])
def test_vcycle_lexicographic_coarse_levels(dims, kind):
    """coarse_segment_rows = 0: every level swept in one segment (the oracle with one thread)."""
    A, S = pair(dims, kind)
    H = O.Hierarchy(S)
    M = la.MultiGridSolver(coarse_segment_rows=0).instantiate(A)
    r = problems.uniform(S.N, 5)
    z = M.apply(la.DenseVector(r), la.DenseVector(S.N)).numpy()
    np.testing.assert_array_equal(z, H.vcycle(r))


def test_reference_pcg_amg_thresholds():
    """test/shared-stencil.hpp:154-190: known solution 2, pcg_amg <= 1e-13 within 32 iterations."""
    dims = (4, 9, 7)
    A, S = pair(dims, "lookup")
    two = la.DenseVector(np.full(S.N, 2.0))
    b = A * two
    for ultra in (False, True):
        M = la.MultiGridSolver(ultra=ultra).instantiate(A)
        for solver in (la.ConjugateGradientsSolver(8, M), la.BiCgStabSolver(8, M)):
            x = la.DenseVector(S.N, fill=0.0)
            res = solver.solve(A, x, b, rel_tol=1e-15, max_it=32)
            xv = x.numpy()
            assert res.residuals[-1] <= 1e-13
            assert np.abs(xv - 2.0).max() <= 1e-13


@pytest.mark.parametrize("dims,kind", CASES[1:])
@pytest.mark.parametrize("precision", [np.float64, np.float32])
def test_pcg_amg_matches_oracle(dims, kind, precision):
    A, S = pair(dims, kind)
    f32 = precision == np.float32
    H = O.Hierarchy(S, coarse_f32=f32, coarse_segment=64)
    M = la.MultiGridSolver(precision=precision).instantiate(A)
    b = S.rhs()
    x_ref, it_ref, hist_ref = O.pcg(S, b, H, precon=2 if f32 else 1, rel_tol=1e-8, max_it=200)
    x = la.DenseVector(S.N, fill=0.0)
    res = la.ConjugateGradientsSolver(8, M).solve(A, x, la.DenseVector(b), rel_tol=1e-8, max_it=200)
    assert res.converged
    assert abs(res.iterations - it_ref) <= 1
    xv = x.numpy()
    assert np.abs(xv - x_ref).max() / np.abs(x_ref).max() <= 1e-7
    # the same recurrence: residual histories agree to rounding
    k = min(len(hist_ref), len(res.residuals))
    np.testing.assert_allclose(res.residuals[:k], hist_ref[:k], rtol=1e-6)


@pytest.mark.parametrize("dims,kind", CASES[1:3])
def test_bicgstab_amg_matches_oracle(dims, kind):
    A, S = pair(dims, kind)
    H = O.Hierarchy(S, coarse_segment=64)
    M = la.MultiGridSolver().instantiate(A)
    b = S.rhs()
    x_ref, it_ref, _ = O.pbicgstab(S, b, H, precon=1, rel_tol=1e-8, max_it=200)
    x = la.DenseVector(S.N, fill=0.0)
    res = la.BiCgStabSolver(8, M).solve(A, x, la.DenseVector(b), rel_tol=1e-8, max_it=200)
    assert res.converged
    assert abs(res.iterations - it_ref) <= 1
    assert np.abs(x.numpy() - x_ref).max() / np.abs(x_ref).max() <= 1e-7


def test_unpreconditioned_cg_exact_termination():
    dims = (3, 4, 2)
    A, S = pair(dims, "values")
    b = S.rhs()
    x = la.DenseVector(S.N, fill=0.0)
    res = la.ConjugateGradientsSolver(8).solve(A, x, la.DenseVector(b), rel_tol=1e-14, max_it=S.N + 5)
    assert res.converged
    dense = np.zeros((S.N, S.N))
    for i in range(S.N):
        c, v = S.row(i)
        dense[i, c] = v
    np.testing.assert_allclose(x.numpy(), np.linalg.solve(dense, b), rtol=1e-12)


def test_solver_status_zero_rhs():
    A, S = pair((8, 8, 8), "values")
    x = la.DenseVector(S.N, fill=0.0)
    res = la.ConjugateGradientsSolver(8).solve(A, x, la.DenseVector(S.N, fill=0.0), rel_tol=1e-8, max_it=10)
    assert res.converged and res.iterations == 0


def test_bicgstab_advection_closed_form():
    """Constant coefficient, Pe = v L_x / D = 10: BiCGStab on the upwind
    operator, preconditioned by the AMG of the diffusion part, reaches the
    closed-form discrete solution p_i = A + B rho^i, rho = (D + v) / D."""
    n, D = 32, 1.0
    v = 10.0 / n
    dims = (n, 8, 8)
    c = np.full(n * 64, D)
    A_adv = la.StencilMatrix.values(la.BoxInfo(dims, velocity_x=v), c)
    A_diff = la.StencilMatrix.values(la.BoxInfo(dims), c)
    M = la.MultiGridSolver().instantiate(A_diff)
    b = A_adv.rhs()
    x = la.DenseVector(A_adv.row_num, fill=0.0)
    res = la.BiCgStabSolver(8, M).solve(A_adv, x, b, rel_tol=1e-12, max_it=300)
    assert res.converged
    rho = (D + v) / D
    Mx = np.array([[2 * D + v, 2 * D], [2 * D, (2 * D + v) * rho ** (n - 1)]])
    a_, b_ = np.linalg.solve(Mx, [(2 * D + v) * 1.0, 0.0])
    exact = a_ + b_ * rho ** np.arange(n)
    np.testing.assert_allclose(x.numpy().reshape(n, 64), np.repeat(exact[:, None], 64, axis=1), rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("dims", [(24, 24, 24), (32, 20, 28)])
def test_bicgstab_advection_two_phase_matches_oracle(dims):
    """Two-phase lookup field (pore 1, solid 1e-6) with upwind advection at
    Pe = 10 (v = 10 / L_x): GPU BiCGStab with the diffusion AMG against the
    oracle's BiCGStab with the same hierarchy."""
    m = problems.spheres(dims, 3, rmin=2, rmax=4)
    mc = (1.0, 1e-6, 0.0, 0.0)
    v = 10.0 / dims[0]
    A_adv = la.StencilMatrix.lookup(la.BoxInfo(dims, velocity_x=v), m, mc)
    A_diff = la.StencilMatrix.lookup(la.BoxInfo(dims), m, mc)
    S_adv = O.Stencil(dims, materials=m, mat_coeffs=mc, velocity=v)
    S_diff = O.Stencil(dims, materials=m, mat_coeffs=mc)
    H = O.Hierarchy(S_diff, coarse_segment=64)
    M = la.MultiGridSolver().instantiate(A_diff)
    b = S_adv.rhs()
    x_ref, it_ref, _ = O.pbicgstab(S_adv, b, H, precon=1, rel_tol=1e-10, max_it=300)
    x = la.DenseVector(S_adv.N, fill=0.0)
    res = la.BiCgStabSolver(8, M).solve(A_adv, x, la.DenseVector(b), rel_tol=1e-10, max_it=300)
    assert res.converged
    assert abs(res.iterations - it_ref) <= 2
    assert np.abs(x.numpy() - x_ref).max() / np.abs(x_ref).max() <= 1e-6
    r = O.residual(S_adv, b, x.numpy())
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("dims", [(256, 256, 256), (256, 128, 256)])
def test_cg_residual_history_full_size(dims):
    """Unpreconditioned CG at full size against the oracle: every reduction
    (<p, A p> fused into the stencil pass, ||r||^2 fused into the update) must
    cover exactly the partial sums its own pass wrote.  At these shapes the
    f32 and f64 stencil passes use different block counts."""
    A, S = pair(dims, "values")
    b = S.rhs()
    x = la.DenseVector(S.N, fill=0.0)
    res = la.ConjugateGradientsSolver(8).solve(A, x, la.DenseVector(b), rel_tol=1e-30, max_it=10)
    _, it, hist = O.pcg(S, b, None, precon=0, rel_tol=1e-30, max_it=10)
    assert res.iterations == it == 10
    np.testing.assert_allclose(res.residuals, hist, rtol=1e-9)


@pytest.mark.timeout(600)
def test_hierarchy_and_vcycle_full_size_256():
    """The 256^3 lognormal hierarchy (aggregate maps) and one FP64 V-cycle
    equal the oracle's bit for bit at the benchmark's size."""
    dims = (256, 256, 256)
    A, S = pair(dims, "values")
    M = la.MultiGridSolver(precision=np.float64).instantiate(A)
    H = O.Hierarchy(S, coarse_segment=64)
    assert [M.level_size(l) for l in range(M.level_num)] == [H.size(l) for l in range(H.level_num)]
    for l in range(H.level_num - 1):
        np.testing.assert_array_equal(M.aggregates(l), H.agg_map(l))
    r = problems.uniform(S.N, 4)
    z = la.DenseVector(S.N)
    M.apply(la.DenseVector(r), z)
    np.testing.assert_array_equal(z.numpy(), H.vcycle(r))


@pytest.mark.parametrize("f32", [False, True])
def test_advection_hierarchy_and_vcycle_equal_oracle(f32):
    """A hierarchy built from the nonsymmetric upwind-advection operator
    (two-phase field, Pe = 10): aggregates, coarse matrices and one V-cycle
    with the LU coarsest solve equal the oracle's bit for bit."""
    dims = (24, 20, 16)
    m = problems.spheres(dims, 3, rmin=2, rmax=4)
    mc = (1.0, 1e-6, 0.0, 0.0)
    v = 10.0 / dims[0]
    A = la.StencilMatrix.lookup(la.BoxInfo(dims, velocity_x=v), m, mc)
    S = O.Stencil(dims, materials=m, mat_coeffs=mc, velocity=v)
    H = O.Hierarchy(S, coarse_f32=f32, coarse_segment=64)
    assert H.is_lu
    M = la.MultiGridSolver(precision=np.float32 if f32 else np.float64).instantiate(A)
    assert M.level_num == H.level_num
    for l in range(H.level_num - 1):
        np.testing.assert_array_equal(M.aggregates(l), H.agg_map(l))
    for l in range(1, H.level_num):
        rp, cols, vals = M.matrix(l)
        ref = H.matrix(l)
        np.testing.assert_array_equal(cols, ref.cols)
        np.testing.assert_array_equal(vals, ref.vals)
    r = problems.uniform(S.N, 6)
    if f32:
        r = r.astype(np.float32)
        z = M.apply(la.DenseVector(r), la.DenseVector(S.N, np.float32)).numpy()
    else:
        z = M.apply(la.DenseVector(r), la.DenseVector(S.N)).numpy()
    np.testing.assert_array_equal(z, H.vcycle(r))


def test_bicgstab_nonsymmetric_amg_matches_oracle():
    dims = (32, 24, 20)
    m = problems.spheres(dims, 3, rmin=2, rmax=4)
    mc = (1.0, 1e-6, 0.0, 0.0)
    v = 10.0 / dims[0]
    A = la.StencilMatrix.lookup(la.BoxInfo(dims, velocity_x=v), m, mc)
    S = O.Stencil(dims, materials=m, mat_coeffs=mc, velocity=v)
    H = O.Hierarchy(S, coarse_segment=64)
    M = la.MultiGridSolver().instantiate(A)
    b = S.rhs()
    x_ref, it_ref, _ = O.pbicgstab(S, b, H, precon=1, rel_tol=1e-10, max_it=300)
    x = la.DenseVector(S.N, fill=0.0)
    res = la.BiCgStabSolver(8, M).solve(A, x, la.DenseVector(b), rel_tol=1e-10, max_it=300)
    assert res.converged and abs(res.iterations - it_ref) <= 1
    assert np.abs(x.numpy() - x_ref).max() <= 1e-8 * np.abs(x_ref).max()
    assert np.linalg.norm(O.residual(S, b, x.numpy())) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("case", [c[0] for c in __import__("aggregation_cases").CASES])
def test_hierarchy_known_answers(case):
    """The library's host aggregation on the hand-traced graphs of
    aggregation_cases.py (the same known answers that pin the oracle)."""
    import aggregation_cases as AC

    _, build, aggregates, _ = next(c for c in AC.CASES if c[0] == case)
    rp, cols, vals = build()
    A = la.CsrMatrix.from_arrays(rp, cols, vals)
    wright = la.HierarchyWrightParams(max_level_num=2, min_coarsen_size=1, min_coarsen_factor=1.0)
    M = la.MultiGridSolver(wright).instantiate(A)
    assert M.level_size(1) == len(aggregates)
    np.testing.assert_array_equal(M.aggregates(0), AC.expected_map(rp.size - 1, aggregates))


@pytest.mark.parametrize("case", [c[0] for c in __import__("aggregation_cases").NONSYMMETRIC_CASES])
def test_hierarchy_nonsymmetric_known_answers(case):
    """The library's aggregation on the hand-traced nonsymmetric graphs of
    aggregation_cases.py (one-way and influences-only couplings)."""
    import aggregation_cases as AC

    _, build, aggregates, _ = next(c for c in AC.NONSYMMETRIC_CASES if c[0] == case)
    rp, cols, vals = build()
    A = la.CsrMatrix.from_arrays(rp, cols, vals, symmetric=False)
    wright = la.HierarchyWrightParams(max_level_num=2, min_coarsen_size=1, min_coarsen_factor=1.0)
    M = la.MultiGridSolver(wright).instantiate(A)
    assert M.level_size(1) == len(aggregates)
    np.testing.assert_array_equal(M.aggregates(0), AC.expected_map(rp.size - 1, aggregates))


@pytest.mark.parametrize("relax,cycle", [(0.7, (la.CYCLE_REGULAR, 1)), (1.6, (la.CYCLE_REGULAR, 2)),
                                         (1.0, (la.CYCLE_DECREASING, 3))])
@pytest.mark.parametrize("f32", [False, True])
def test_cycle_relax_and_kinds_equal_oracle(relax, cycle, f32):
    """MultiGridSolver{relax, DefaultCycle{kind, iterations}} against the
    oracle's restatement of the reference's cycle stack, bit for bit."""
    dims = (24, 20, 16)
    A, S = pair(dims)
    H = O.Hierarchy(S, coarse_f32=f32, coarse_segment=64, relax=relax, cycle_kind=cycle[0],
                    cycle_iterations=cycle[1])
    M = la.MultiGridSolver(precision=np.float32 if f32 else np.float64, relax=relax,
                           cycle=cycle).instantiate(A)
    r = problems.uniform(S.N, 8)
    for _ in range(2):  # the level vectors persist between applications
        if f32:
            r32 = r.astype(np.float32)
            z = M.apply(la.DenseVector(r32), la.DenseVector(S.N, np.float32)).numpy()
            np.testing.assert_array_equal(z, H.vcycle(r32))
        else:
            z = M.apply(la.DenseVector(r), la.DenseVector(S.N)).numpy()
            np.testing.assert_array_equal(z, H.vcycle(r))
        r = problems.uniform(S.N, 9)


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("kind", ["values", "spheres"])
def test_hierarchy_import_round_trip_and_oracle_levels(f32, kind):
    """lgpu_amg_create_from_levels: the library's own hierarchy exported and
    imported again gives a bit-identical cycle; the oracle's hierarchy (the
    host setup restated from the reference) imported with its coarse -> fine
    lists gives the oracle's cycle bit for bit."""
    dims = (24, 20, 16) if kind == "values" else (24, 24, 24)
    A, S = pair(dims, kind)
    prec = np.float32 if f32 else np.float64
    solver = la.MultiGridSolver(precision=prec)
    M = solver.instantiate(A)
    M2 = la.MultiGridInstance.from_levels(A, M.export_levels(), solver)
    assert M2.level_num == M.level_num
    H = O.Hierarchy(S, coarse_f32=f32, coarse_segment=64)
    levels = []
    for l in range(H.level_num - 1):
        ptr, idx = H.c2f(l)
        m = H.matrix(l + 1)
        levels.append(la.AmgLevel(H.agg_map(l), m.rowptr, m.cols, m.vals, ptr, idx))
    M3 = la.MultiGridInstance.from_levels(A, levels, solver)
    r = problems.uniform(S.N, 12)
    rv = r.astype(prec)
    zs = [X.apply(la.DenseVector(rv), la.DenseVector(S.N, prec)).numpy() for X in (M, M2, M3)]
    np.testing.assert_array_equal(zs[1], zs[0])
    np.testing.assert_array_equal(zs[2], H.vcycle(rv))


def test_hierarchy_import_rejects_inconsistent_levels():
    dims = (12, 10, 9)
    A, S = pair(dims)
    M = la.MultiGridSolver().instantiate(A)
    levels = M.export_levels()
    bad = levels[0]
    agg = bad.agg.copy()
    agg[0] = bad.rowptr.size + 5  # out of range
    with pytest.raises(la.LinealError):
        la.MultiGridInstance.from_levels(A, [la.AmgLevel(agg, bad.rowptr, bad.cols, bad.vals)] + levels[1:])
    idx = bad.c2f_idx.copy()
    idx[[0, -1]] = idx[[-1, 0]]  # lists no longer match the map
    if bad.agg[idx[0]] != bad.agg[idx[-1]]:
        with pytest.raises(la.LinealError):
            la.MultiGridInstance.from_levels(A, [la.AmgLevel(bad.agg, bad.rowptr, bad.cols, bad.vals,
                                                             bad.c2f_ptr, idx)] + levels[1:])


@pytest.mark.parametrize("case", ["values", "spheres", "advection", "f32", "csr"])
def test_device_setup_equals_host_setup(case):
    """The strength criterion and the Galerkin products on the GPU give the
    host setup's hierarchy bit for bit: the same aggregate maps (the
    aggregation reads the device's strength flags) and coarse matrices."""
    dims = (40, 36, 32) if case in ("values", "f32") else (32, 32, 32)
    if case == "spheres":
        A, _ = pair(dims, "spheres")
    elif case == "advection":
        m = problems.spheres(dims, 3, rmin=2, rmax=4)
        A = la.StencilMatrix.lookup(la.BoxInfo(dims, velocity_x=10.0 / dims[0]), m, (1.0, 1e-6, 0.0, 0.0))
    else:
        A, _ = pair(dims)
        if case == "csr":
            A = A.to_csr()
    prec = np.float32 if case == "f32" else np.float64
    host = la.MultiGridSolver(la.HierarchyWrightParams(setup_on_device=False), precision=prec).instantiate(A)
    dev = la.MultiGridSolver(la.HierarchyWrightParams(setup_on_device=True), precision=prec).instantiate(A)
    assert dev.level_num == host.level_num >= 3
    for l in range(host.level_num - 1):
        np.testing.assert_array_equal(dev.aggregates(l), host.aggregates(l))
        np.testing.assert_array_equal(dev.transfer(l)[1], host.transfer(l)[1])
    for l in range(1, host.level_num):
        for a, b in zip(dev.matrix(l), host.matrix(l)):
            np.testing.assert_array_equal(a, b)
