"""GPU parity of the compressed stencil->CSR system and the effective
diffusivity (tensor/dynamic/csr/create-from.hpp:37-239,
component-wise/sink/effective-diffusion.hpp:229-327) against the oracle."""

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


def zero_pair(dims, kind, zero_to_one=False, lengths=None):
    """Lookup operators whose material 0 (diagonal lines) or the solid spheres
    have coefficient 0, so ZeroToZero leaves rows with a zero diagonal."""
    box = la.BoxInfo(dims, lengths=lengths, zero_to_one=zero_to_one)
    if kind == "lines":
        m, mc = problems.diagonal_lines(dims), (0.0, 1.0, 2.0, 3.0)
    else:
        m, mc = problems.spheres(dims, 4, rmin=1, rmax=3), (1.0, 0.0, 0.0, 0.0)
    return (la.StencilMatrix.lookup(box, m, mc),
            O.Stencil(dims, materials=m, mat_coeffs=mc, lengths=lengths, zero_to_one=zero_to_one), m)


@pytest.mark.parametrize("dims", [(4, 7, 6), (13, 17, 5), (33, 17, 65)])
@pytest.mark.parametrize("kind", ["lines", "spheres"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_compressed_system_equals_oracle(dims, kind, dtype):
    A, S, _ = zero_pair(dims, kind)
    sys = la.csr_from_adjacent_stencil(A, dtype, compress=True)
    C, rhs, cmap = S.to_csr(compress=True)
    np.testing.assert_array_equal(sys.compress_map, cmap)
    assert sys.lhs.row_num == C.N == sys.rhs.size
    assert C.N < S.N
    keep = np.flatnonzero(cmap >= 0)
    if dtype == np.float64:
        same(sys.rhs.numpy(), rhs)
    else:
        same(sys.rhs.numpy(), S.rhs(np.float32)[keep])
        C = O.Csr(C.rowptr, C.cols, C.vals.astype(np.float32).astype(np.float64), is_f32=True)
    x = problems.uniform(C.N, 6, dtype)
    same(sys.lhs.spmv(la.DenseVector(x), la.DenseVector(C.N, dtype)).numpy(), O.spmv(C, x))


@pytest.mark.parametrize("dims", [(4, 9, 7), (13, 17, 5)])
def test_uncompressed_system_and_zero_to_one(dims):
    """Without compression (or with ZeroToOne, where no diagonal is zero) the
    system is the plain CSR copy: identity map, full rhs."""
    for zero_to_one, compress in ((False, False), (True, True)):
        A, S, _ = zero_pair(dims, "lines", zero_to_one=zero_to_one)
        sys = la.csr_from_adjacent_stencil(A, np.float64, compress=compress)
        assert sys.lhs.row_num == S.N
        if compress:
            np.testing.assert_array_equal(sys.compress_map, np.arange(S.N))
        else:
            assert sys.compress_map is None
        same(sys.rhs.numpy(), S.rhs())
        C, _, _ = S.to_csr()
        x = problems.uniform(S.N, 7)
        same(sys.lhs.spmv(la.DenseVector(x), la.DenseVector(S.N)).numpy(), O.spmv(C, x))


# Tolerances: the area part is an integer count and must be equal; the flux
# sum is reduced in a different order from the oracle's sequential sum, so it
# agrees to 1e-12 relative in f64.  An f32 solution's differences are taken
# in f32 (the oracle takes them in f64), so 1e-6 relative there.
@pytest.mark.parametrize("dims", [(3, 1, 1), (4, 9, 7), (13, 17, 5), (33, 17, 65), (6, 21, 130)])
@pytest.mark.parametrize("kind", ["values", "lookup", "spheres"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_eff_diff_matches_oracle(dims, kind, dtype):
    lengths = (2.0, 3.5, 0.75)
    box = la.BoxInfo(dims, lengths=lengths, p_start=0.25, p_end=1.5)
    kw = dict(lengths=lengths, p_start=0.25, p_end=1.5)
    none = None
    if kind == "values":
        c = problems.lognormal(dims, 1.0, 2)
        A, S = la.StencilMatrix.values(box, c), O.Stencil(dims, coeff=c, **kw)
    else:
        rng = np.random.default_rng(11)
        m = (rng.integers(0, 4, size=int(np.prod(dims)), dtype=np.uint8) if kind == "lookup"
             else problems.spheres(dims, 5, rmin=1, rmax=3))
        mc = (0.5, 1.0, 2.0, 3.0) if kind == "lookup" else (1.0, 1e-6, 0.0, 0.0)
        none = (False, True, False, True) if kind == "lookup" else (False, True, False, False)
        A, S = la.StencilMatrix.lookup(box, m, mc), O.Stencil(dims, materials=m, mat_coeffs=mc, **kw)
    calc = la.EffDiffCalc(A, none)
    area = O.eff_diff_area_part(S, none)
    assert calc.area_part == area
    assert la.EffDiffCalc(A).area_part == O.eff_diff_area_part(S) == 1.0
    sol = problems.uniform(S.N, 9, dtype)
    got = calc(A, la.DenseVector(sol))
    ref = O.eff_diff(S, sol.astype(np.float64), area)
    if not np.isfinite(ref):
        # no filled cell in the mid planes: area part 0, the quotient is
        # +-inf (or NaN) on both sides, as in the reference
        assert area == 0.0 and (got == ref or (np.isnan(got) and np.isnan(ref)))
        return
    rtol = 1e-12 if dtype == np.float64 else 1e-6
    assert abs(got - ref) <= rtol * max(abs(ref), 1e-300) + 1e-300


@pytest.mark.parametrize("dims", [(4, 7, 6), (33, 17, 65)])
def test_eff_diff_through_the_compress_map(dims):
    """The reference's compressed case (test/shared-csr.cpp:193-252): solve the
    compressed system with CG on the GPU, then the effective diffusivity read
    through the compress map is the known answer, and equals the oracle's on
    the same solution."""
    A, S, _ = zero_pair(dims, "lines")
    sys = la.csr_from_adjacent_stencil(A, np.float64, compress=True)
    sol = la.DenseVector(sys.lhs.row_num, np.float64, fill=0.0)
    res = la.ConjugateGradientsSolver(8).solve(sys.lhs, sol, sys.rhs, rel_tol=1e-13, max_it=4000)
    assert res.converged
    calc = la.EffDiffCalc(A)
    got = calc(A, sol, sys)
    ref = problems.diagonal_lines_eff_diff(dims, (0.0, 1.0, 2.0, 3.0))
    assert abs(got - ref) <= 1e-8
    x = sol.numpy()
    oref = O.eff_diff(S, x, O.eff_diff_area_part(S), sys.compress_map)
    assert abs(got - oref) <= 1e-12 * abs(oref)


def test_eff_diff_full_size_known_answer():
    """256^3 diagonal lines with the exact linear field: the mean coefficient
    over the cross-section (test/aux/materials.hpp:235-248), a
    size-independent property; and the area part of the 256^3 two-phase
    spheres with the solid marked "none" equals a numpy count."""
    dims = (256, 256, 256)
    mc = (0.5, 1.0, 2.0, 3.0)
    m = problems.diagonal_lines(dims)
    A = la.StencilMatrix.lookup(la.BoxInfo(dims, lengths=(1.0, 2.0, 3.0)), m, mc)
    x = np.arange(256, dtype=np.float64)[:, None, None]
    p = np.ascontiguousarray(np.broadcast_to(1.0 - (x + 0.5) / 256, dims)).reshape(-1)
    got = la.EffDiffCalc(A)(A, la.DenseVector(p))
    ref = problems.diagonal_lines_eff_diff(dims, mc)
    assert abs(got - ref) <= 1e-12 * ref
    sp = problems.spheres(dims, 3)
    B = la.StencilMatrix.lookup(la.BoxInfo(dims), sp, (1.0, 1e-6, 0.0, 0.0))
    g = sp.reshape(dims)
    filled = int(np.count_nonzero(g[128] == 0) + np.count_nonzero(g[129] == 0))
    assert la.EffDiffCalc(B, (False, True, False, False)).area_part == filled / 256 ** 2 * 0.5


def test_eff_diff_argument_errors():
    A, _, _ = zero_pair((2, 4, 4), "lines")
    with pytest.raises(la.LinealError):
        la.EffDiffCalc(A)
    A, _, _ = zero_pair((4, 4, 4), "lines")
    B, _, _ = zero_pair((5, 4, 4), "lines")
    sys = la.csr_from_adjacent_stencil(B, np.float64, compress=True)
    calc = la.EffDiffCalc(A)
    with pytest.raises(la.LinealError):
        calc(A, la.DenseVector(sys.lhs.row_num), sys)  # map of another stencil
    with pytest.raises(la.LinealError):
        calc(A, la.DenseVector(7), None)  # wrong length
