"""GPU parity of the matrix and vector files (io/tensor.hpp:25-263): the
reference's UFZ_CT_02 fixtures (test/tensor-io.cpp:58-100) read into GPU
operators and vectors, written back byte for byte, and solved."""

import filecmp
import os

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
LHS = os.path.join(GOLDEN, "UFZ_CT_02_LHS.smf")
RHS = os.path.join(GOLDEN, "UFZ_CT_02_RHS.dvf")


def same(a, b):
    assert a.shape == b.shape
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


def oracle_csr(path, dtype=np.float64):
    _, _, _, rp, cols, vals = O.smf_read(path)
    if dtype == np.float32:
        vals = vals.astype(np.float32).astype(np.float64)
    return O.Csr(rp, cols, vals, symmetric=False, is_f32=dtype == np.float32)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_smf_read_spmv_and_write_round_trip(dtype, tmp_path):
    r = la.SmfReader(LHS)
    assert (r.row_num(), r.col_num(), r.nnz_num()) == (539, 539, 3367)
    A = r.read(dtype)
    assert A.row_num == 539
    C = oracle_csr(LHS, dtype)
    x = problems.uniform(539, 3, dtype)
    same(A.spmv(la.DenseVector(x), la.DenseVector(539, dtype)).numpy(), O.spmv(C, x))
    if dtype == np.float64:
        la.SmfWriter(tmp_path / "w.smf").write(A)
        assert filecmp.cmp(tmp_path / "w.smf", LHS, shallow=False)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dvf_read_and_write_round_trip(dtype, tmp_path):
    r = la.DvfReader(RHS)
    assert r.size() == 539
    v = r.read(dtype)
    ref = O.dvf_read(RHS)
    same(v.numpy(), ref.astype(dtype))
    la.DvfWriter(tmp_path / "w.dvf").write(v)
    if dtype == np.float64:
        assert filecmp.cmp(tmp_path / "w.dvf", RHS, shallow=False)
    else:
        same(O.dvf_read(tmp_path / "w.dvf"), ref.astype(np.float32).astype(np.float64))


def test_smf_duplicates_and_unsorted_rows(tmp_path):
    """A file with duplicate and unsorted columns reads as the oracle's
    row-wise builder does (tensor/dynamic/csr/shared.hpp:730-764)."""
    rng = np.random.default_rng(5)
    n = 70
    rows = []
    for i in range(n):
        cols = list(rng.integers(0, n, size=8)) + [i, i]
        rng.shuffle(cols)
        rows.append([(int(c), float(rng.uniform(-1, 1))) for c in cols])
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    ent = np.array([e for r in rows for e in r], dtype=[("c", "<u8"), ("v", "<f8")])
    with open(tmp_path / "d.smf", "wb") as f:
        f.write(np.array([n, n, off[-1]], "<u8").tobytes() + off.astype("<u8").tobytes() + ent.tobytes())
    A = la.SmfReader(tmp_path / "d.smf").read()
    C = oracle_csr(tmp_path / "d.smf")
    x = problems.uniform(n, 8)
    same(A.spmv(la.DenseVector(x), la.DenseVector(n)).numpy(), O.spmv(C, x))
    # written back: the oracle's writer on the oracle's reading, byte for byte
    la.SmfWriter(tmp_path / "g.smf").write(A)
    _, _, _, rp, cols, vals = O.smf_read(tmp_path / "d.smf")
    O.smf_write(tmp_path / "o.smf", n, n, rp, cols, vals)
    assert filecmp.cmp(tmp_path / "g.smf", tmp_path / "o.smf", shallow=False)


@pytest.mark.parametrize("kind", ["values", "lookup"])
def test_smf_write_of_a_stencil(kind, tmp_path):
    """SmfWriter over a stencil matrix: rows in column order, zero entries
    skipped, the header counting the written entries (a stencil matrix has no
    non_zero_num)."""
    dims = (5, 6, 7)
    if kind == "values":
        c = problems.lognormal(dims, 1.0, 4)
        A, S = la.StencilMatrix.values(la.BoxInfo(dims), c), O.Stencil(dims, coeff=c)
    else:
        m = problems.diagonal_lines(dims)
        A = la.StencilMatrix.lookup(la.BoxInfo(dims), m)
        S = O.Stencil(dims, materials=m)
    la.SmfWriter(tmp_path / "s.smf").write(A)
    n, ncols, nnz, rp, cols, vals = O.smf_read(tmp_path / "s.smf")
    C, _, _ = S.to_csr()
    assert n == ncols == S.N and nnz == C.rowptr[-1]
    same(rp, C.rowptr)
    same(cols, C.cols)
    same(vals, C.vals)


@pytest.mark.parametrize("symmetric", [True, False])
def test_ufz_system_solves(symmetric):
    """UFZ_CT_02 (SPD) with AMG-preconditioned CG (declared symmetric:
    Cholesky coarsest) and BiCGStab (the reference's nonsymmetric reading: LU
    coarsest) to 1e-10; the solution matches a dense solve of the oracle's
    reading."""
    A = la.SmfReader(LHS).read(np.float64, symmetric=symmetric)
    b = la.DvfReader(RHS).read()
    M = la.MultiGridSolver(precision=np.float64).instantiate(A)
    x = la.DenseVector(539, fill=0.0)
    solver = la.ConjugateGradientsSolver(8, M) if symmetric else la.BiCgStabSolver(8, M)
    res = solver.solve(A, x, b, rel_tol=1e-10, max_it=200)
    assert res.converged, res
    C = oracle_csr(LHS)
    ref = np.linalg.solve(C.dense(), O.dvf_read(RHS))
    # relative residual 1e-10 with a condition number of about 200
    assert np.abs(x.numpy() - ref).max() <= 1e-7 * np.abs(ref).max()


def test_file_errors(tmp_path):
    with pytest.raises(la.LinealError):
        la.SmfReader(tmp_path / "missing.smf")
    with pytest.raises(la.LinealError):
        la.DvfReader(tmp_path / "missing.dvf")
    data = open(LHS, "rb").read()
    (tmp_path / "short.smf").write_bytes(data[:1000])
    with pytest.raises(la.LinealError):
        la.SmfReader(tmp_path / "short.smf").read()
    (tmp_path / "short.dvf").write_bytes(open(RHS, "rb").read()[:100])
    with pytest.raises(la.LinealError):
        la.DvfReader(tmp_path / "short.dvf").read()
    # a non-square matrix
    with open(tmp_path / "rect.smf", "wb") as f:
        f.write(np.array([2, 3, 2], "<u8").tobytes() + np.array([0, 1, 2], "<u8").tobytes()
                + np.array([(0, 1.0), (1, 1.0)], dtype=[("c", "<u8"), ("v", "<f8")]).tobytes())
    with pytest.raises(la.LinealError):
        la.SmfReader(tmp_path / "rect.smf").read()
