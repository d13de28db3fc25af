"""GPU parity of the single kernels against the oracle (bit-exact).

Every kernel performs the oracle's operations in the oracle's order (explicit
fma, no contraction), so results must be identical, not merely close.
"""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu

SHAPES = [(4, 9, 7), (1, 1, 1), (3, 1, 70), (13, 17, 5), (33, 17, 65), (40, 40, 40), (6, 21, 130)]


def make_pair(dims, kind, seed=1, sigma=1.0, p_start=1.0, p_end=0.0, mat_coeffs=(0.0, 1.0, 2.0, 3.0)):
    box = la.BoxInfo(dims, p_start=p_start, p_end=p_end)
    if kind == "values":
        c = problems.lognormal(dims, sigma, seed)
        return la.StencilMatrix.values(box, c), O.Stencil(dims, coeff=c, p_start=p_start, p_end=p_end)
    m = problems.diagonal_lines(dims) if kind == "lookup" else problems.spheres(dims, seed, rmin=1, rmax=3)
    mc = mat_coeffs if kind == "lookup" else (1.0, 1e-6, 0.0, 0.0)
    return (la.StencilMatrix.lookup(box, m, mc),
            O.Stencil(dims, materials=m, mat_coeffs=mc, p_start=p_start, p_end=p_end))


def same(a, b, case=""):
    assert a.shape == b.shape
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{case}{bad.size} mismatches, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("kind", ["values", "lookup", "spheres"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_residual_rhs(dims, kind, dtype):
    A, S = make_pair(dims, kind)
    n = S.N
    x = problems.uniform(n, 2, dtype)
    xd = la.DenseVector(x)
    same(A.spmv(xd, la.DenseVector(n, dtype)).numpy(), O.spmv(S, x))
    b = S.rhs(dtype)
    same(A.rhs(dtype).numpy(), b)
    r = A.residual(la.DenseVector(b), xd, la.DenseVector(n, dtype)).numpy()
    same(r, O.residual(S, b, x))


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("kind", ["values", "lookup"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("with_x", [True, False])
def test_jacobi(dims, kind, dtype, with_x):
    A, S = make_pair(dims, kind)
    n = S.N
    x = problems.uniform(n, 2, dtype)
    b = problems.uniform(n, 3, dtype)
    omega = 2.0 / 3.0
    out = A.jacobi(la.DenseVector(x) if with_x else None, la.DenseVector(n, dtype), la.DenseVector(b), omega)
    same(out.numpy(), O.jacobi(S, x if with_x else None, b, dtype(omega)))


@pytest.mark.parametrize("dims", SHAPES + [(17, 64, 64)])
@pytest.mark.parametrize("kind", ["values", "lookup", "spheres"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("forward", [True, False])
@pytest.mark.parametrize("ultra", [False, True])
@pytest.mark.parametrize("cache", [True, False])
def test_sor_wavefront(dims, kind, dtype, forward, ultra, cache):
    """Both sweep kernels: cached per-cell face weights and diagonal (the
    default) and the weights evaluated on the fly."""
    A, S = make_pair(dims, kind)
    A.set_sor_cache(cache)
    n = S.N
    x = problems.uniform(n, 2, dtype)
    b = problems.uniform(n, 3, dtype)
    for omega, with_x in ((1.0, True), (1.3, True), (1.0, False)):
        out = A.sor(la.DenseVector(x) if with_x else None, la.DenseVector(n, dtype), la.DenseVector(b),
                    omega, forward, ultra)
        ref = O.sor(S, x if with_x else None, b, dtype(omega), forward, ultra)
        same(out.numpy(), ref)


def special_coeffs(dims, which, seed=5):
    """Lognormal coefficients with 30% of the cells replaced by range edges.

    "edges": +0 and the ends of the scaling-free division range (f64 and f32);
    "wide": values outside that range (exact IEEE division path);
    "negzero": -0 and +0 cells (a -0 coefficient also selects the exact path).
    """
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    c = np.exp(rng.normal(0.0, 1.0, n))
    vals = {"edges": [0.0, 2.0 ** -100, 2.0 ** 100, 2.0 ** -24, 2.0 ** 24],
            "wide": [0.0, 2.0 ** -200, 2.0 ** 60, 1e-300],
            "negzero": [-0.0, 0.0]}[which]
    m = rng.random(n) < 0.3
    c[m] = np.asarray(vals)[rng.integers(0, len(vals), int(m.sum()))]
    return c


@pytest.mark.parametrize("dims", [(13, 17, 5), (33, 17, 65), (5, 22, 130)])
@pytest.mark.parametrize("which", ["edges", "wide", "negzero"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_values_coefficient_ranges(dims, which, dtype):
    c = special_coeffs(dims, which)
    for zero_to_one in (True, False):
        box = la.BoxInfo(dims, p_start=1.0, p_end=0.0, zero_to_one=zero_to_one)
        A = la.StencilMatrix.values(box, c)
        S = O.Stencil(dims, coeff=c, p_start=1.0, p_end=0.0, zero_to_one=zero_to_one)
        n = S.N
        x = problems.uniform(n, 2, dtype)
        b = problems.uniform(n, 3, dtype)
        xd = la.DenseVector(x)
        same(A.spmv(xd, la.DenseVector(n, dtype)).numpy(), O.spmv(S, x))
        same(A.residual(la.DenseVector(b), xd, la.DenseVector(n, dtype)).numpy(), O.residual(S, b, x))
        same(A.jacobi(xd, la.DenseVector(n, dtype), la.DenseVector(b), 0.7).numpy(),
             O.jacobi(S, x, b, dtype(0.7)))
        for cache in (True, False):
            A.set_sor_cache(cache)
            for forward in (True, False):
                out = A.sor(xd, la.DenseVector(n, dtype), la.DenseVector(b), 1.1, forward)
                same(out.numpy(), O.sor(S, x, b, dtype(1.1), forward, False))


@pytest.mark.parametrize("dims", [(4, 9, 7), (13, 17, 5), (33, 17, 65)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_kernels(dims, dtype):
    A, S = make_pair(dims, "values")
    C, _, _ = S.to_csr()
    if dtype == np.float32:
        C = O.Csr(C.rowptr, C.cols, C.vals.astype(np.float32).astype(np.float64), is_f32=True)
    G = A.to_csr(dtype)
    n = S.N
    x = problems.uniform(n, 2, dtype)
    b = problems.uniform(n, 3, dtype)
    same(G.spmv(la.DenseVector(x), la.DenseVector(n, dtype)).numpy(), O.spmv(C, x))
    same(G.residual(la.DenseVector(b), la.DenseVector(x), la.DenseVector(n, dtype)).numpy(), O.residual(C, b, x))
    same(G.jacobi(la.DenseVector(x), la.DenseVector(n, dtype), la.DenseVector(b), 0.7).numpy(),
         O.jacobi(C, x, b, dtype(0.7)))
    for forward in (True, False):
        for ultra in (False, True):
            for with_x in (True, False):
                out = G.sor(la.DenseVector(x) if with_x else None, la.DenseVector(n, dtype), la.DenseVector(b),
                            1.0, forward, ultra)
                same(out.numpy(), O.sor(C, x if with_x else None, b, dtype(1.0), forward, ultra))
                for seg in (1, 7, 32, 64):
                    out = G.sor(la.DenseVector(x) if with_x else None, la.DenseVector(n, dtype),
                                la.DenseVector(b), 1.1, forward, ultra, segment_rows=seg)
                    bounds = list(range(0, n, seg)) + [n]
                    same(out.numpy(), O.sor(C, x if with_x else None, b, dtype(1.1), forward, ultra, bounds))


def long_row_csr(n, seed, long_rows):
    """Random diagonally dominant CSR with sorted columns: short rows (at most 8
    neighbours, mostly near the diagonal) and, at the listed rows, rows of up to
    60 entries with more than 16 new-value entries inside their 64-row segment
    and more than 24 trailing ones."""
    rng = np.random.default_rng(seed)
    rp, cols, vals = [0], [], []
    for i in range(n):
        if i in long_rows:
            s0 = (i // 64) * 64
            before = np.arange(s0, i)[-20:]
            after = rng.choice(np.arange(i + 1, n), size=min(30, n - i - 1), replace=False) if i + 1 < n else []
            c = np.union1d(np.union1d(before, after), [i])
        else:
            near = i + rng.integers(-5, 6, size=6)
            far = rng.integers(0, n, size=2)
            c = np.union1d(np.clip(np.concatenate([near, far]), 0, n - 1), [i])
        v = rng.uniform(-1.0, 1.0, size=c.size)
        v[c == i] = np.abs(v).sum() + 1.0
        cols.extend(c.tolist())
        vals.extend(v.tolist())
        rp.append(len(cols))
    return np.array(rp, np.int64), np.array(cols, np.int64), np.array(vals)


@pytest.mark.parametrize("long_rows", [(), (70, 130), (3, 64, 100, 190, 250)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_segment_sor_register_and_fallback_paths(long_rows, dtype):
    """64-row segment SOR: waves whose rows fit the register lists take the
    register path, waves with a long row the LDS path; both equal the oracle
    bit for bit, forward and backward, regular and ultra, with and without
    x_in."""
    n = 260
    rp, cols, vals = long_row_csr(n, 17 + len(long_rows), set(long_rows))
    if dtype == np.float32:
        vals = vals.astype(np.float32).astype(np.float64)
    C = O.Csr(rp, cols, vals, symmetric=False, is_f32=dtype == np.float32)
    G = la.CsrMatrix.from_arrays(rp, cols, vals, dtype, symmetric=False)
    x = problems.uniform(n, 4, dtype)
    b = problems.uniform(n, 5, dtype)
    bounds = list(range(0, n, 64)) + [n]
    for forward in (True, False):
        for ultra in (False, True):
            for with_x in (True, False):
                out = G.sor(la.DenseVector(x) if with_x else None, la.DenseVector(n, dtype),
                            la.DenseVector(b), 1.1, forward, ultra, segment_rows=64)
                same(out.numpy(), O.sor(C, x if with_x else None, b, dtype(1.1), forward, ultra, bounds))


def tiered_csr(seed):
    """Five 64-row segments whose rows need, in turn, at most (new-value,
    trailing) entries of (1, 1), (4, 8), (8, 16), (16, 24) and more: each
    wave of the segment sweep runs a different chain length (the last one the
    LDS walk).  Rows have old-value entries before the segment too."""
    rng = np.random.default_rng(seed)
    limits = [(1, 1), (4, 8), (8, 16), (16, 24), (20, 30)]
    n = 64 * len(limits)
    rp, cols, vals = [0], [], []
    for i in range(n):
        seg = i // 64
        s0 = seg * 64
        nb, na = limits[seg]
        before = np.arange(s0, i)[-nb:] if nb else np.array([], np.int64)
        pool = np.arange(i + 1, n)
        after = rng.choice(pool, size=min(na, pool.size), replace=False) if pool.size else np.array([], np.int64)
        lead = rng.integers(0, s0, size=3) if s0 else np.array([], np.int64)
        c = np.union1d(np.union1d(np.union1d(before, after), lead), [i]).astype(np.int64)
        v = rng.uniform(-1.0, 1.0, size=c.size)
        v[c == i] = np.abs(v).sum() + 1.0
        cols.extend(c.tolist())
        vals.extend(v.tolist())
        rp.append(len(cols))
    return np.array(rp, np.int64), np.array(cols, np.int64), np.array(vals)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_segment_sor_chain_tiers(dtype):
    """Every chain length of the 64-row segment sweep's register path (and
    the LDS walk) against the oracle, bit for bit, both directions, regular
    and ultra."""
    rp, cols, vals = tiered_csr(29)
    n = rp.size - 1
    if dtype == np.float32:
        vals = vals.astype(np.float32).astype(np.float64)
    C = O.Csr(rp, cols, vals, symmetric=False, is_f32=dtype == np.float32)
    G = la.CsrMatrix.from_arrays(rp, cols, vals, dtype, symmetric=False)
    x = problems.uniform(n, 6, dtype)
    b = problems.uniform(n, 7, dtype)
    bounds = list(range(0, n, 64)) + [n]
    for forward in (True, False):
        for ultra in (False, True):
            out = G.sor(la.DenseVector(x), la.DenseVector(n, dtype), la.DenseVector(b), 0.9, forward, ultra,
                        segment_rows=64)
            same(out.numpy(), O.sor(C, x, b, dtype(0.9), forward, ultra, bounds),
                 f"forward={forward} ultra={ultra}: ")


@pytest.mark.parametrize("n,long_rows", [(65, ()), (300, ()), (300, (3, 64, 100, 190, 250)), (512, (7, 300, 511))])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_round_program_sor(n, long_rows, dtype):
    """One-segment SOR of a small level by its round program (rows grouped by
    dependency level, chains of 8 to 64 entries): bit for bit the oracle's
    one-segment sweep, forward and backward, regular and ultra, with and
    without x_in."""
    rp, cols, vals = long_row_csr(n, 41 + len(long_rows), set(long_rows))
    if dtype == np.float32:
        vals = vals.astype(np.float32).astype(np.float64)
    C = O.Csr(rp, cols, vals, symmetric=False, is_f32=dtype == np.float32)
    G = la.CsrMatrix.from_arrays(rp, cols, vals, dtype, symmetric=False)
    x = problems.uniform(n, 8, dtype)
    b = problems.uniform(n, 9, dtype)
    for forward in (True, False):
        for ultra in (False, True):
            for with_x in (True, False):
                out = G.sor(la.DenseVector(x) if with_x else None, la.DenseVector(n, dtype),
                            la.DenseVector(b), 1.2, forward, ultra, segment_rows=n)
                same(out.numpy(), O.sor(C, x if with_x else None, b, dtype(1.2), forward, ultra, [0, n]),
                     f"forward={forward} ultra={ultra} x_in={with_x}: ")


def banded_csr(n, per_row, width, seed):
    """Random diagonally dominant CSR: about per_row entries per row within
    `width` of the diagonal, sorted columns."""
    rng = np.random.default_rng(seed)
    rp, cols, vals = [0], [], []
    for i in range(n):
        c = np.union1d(np.clip(i + rng.integers(-width, width + 1, size=per_row), 0, n - 1), [i])
        v = rng.uniform(-1.0, 1.0, size=c.size)
        v[c == i] = np.abs(v).sum() + 1.0
        cols.extend(c.tolist())
        vals.extend(v.tolist())
        rp.append(len(cols))
    return np.array(rp, np.int64), np.array(cols, np.int64), np.array(vals)


@pytest.mark.parametrize("per_row", [12, 22, 30, 40])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_round_program_register_slots(per_row, dtype):
    """Round programs whose levels hold 16, 24 or 32 entries per row in
    registers (f32; f64 always 16), the rest of longer rows read from LDS:
    bit for bit the oracle's one-segment sweep."""
    n = 377
    rp, cols, vals = banded_csr(n, per_row, 40, 50 + per_row)
    if dtype == np.float32:
        vals = vals.astype(np.float32).astype(np.float64)
    C = O.Csr(rp, cols, vals, symmetric=False, is_f32=dtype == np.float32)
    G = la.CsrMatrix.from_arrays(rp, cols, vals, dtype, symmetric=False)
    x = problems.uniform(n, 10, dtype)
    b = problems.uniform(n, 11, dtype)
    for forward in (True, False):
        for ultra in (False, True):
            out = G.sor(la.DenseVector(x), la.DenseVector(n, dtype), la.DenseVector(b), 0.8, forward, ultra,
                        segment_rows=n)
            same(out.numpy(), O.sor(C, x, b, dtype(0.8), forward, ultra, [0, n]),
                 f"forward={forward} ultra={ultra}: ")


@pytest.mark.parametrize("long_rows", [(), (70, 130, 400), (3, 64, 100, 190, 250, 599)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_scheduled_segment_sor(long_rows, dtype):
    """Segments longer than a wavefront (the scheduled sweep: rows grouped by
    dependency depth, one workgroup per segment), one segment for the whole
    matrix included: bit for bit the oracle's thread-segment SOR, forward and
    backward, regular and ultra, with and without x_in; rows of up to 60
    entries take the path that reads their entries from memory."""
    n = 600
    rp, cols, vals = long_row_csr(n, 23 + len(long_rows), set(long_rows))
    if dtype == np.float32:
        vals = vals.astype(np.float32).astype(np.float64)
    C = O.Csr(rp, cols, vals, symmetric=False, is_f32=dtype == np.float32)
    G = la.CsrMatrix.from_arrays(rp, cols, vals, dtype, symmetric=False)
    x = problems.uniform(n, 6, dtype)
    b = problems.uniform(n, 7, dtype)
    for seg in (65, 100, 256, 600, 5000):
        bounds = list(range(0, n, seg)) + [n]
        for forward in (True, False):
            for ultra in (False, True):
                for with_x in (True, False):
                    out = G.sor(la.DenseVector(x) if with_x else None, la.DenseVector(n, dtype),
                                la.DenseVector(b), 0.9, forward, ultra, segment_rows=seg)
                    same(out.numpy(), O.sor(C, x if with_x else None, b, dtype(0.9), forward, ultra, bounds),
                         f"segment {seg}, forward {forward}, ultra {ultra}, x_in {with_x}: ")


@pytest.mark.parametrize("dims", [(13, 17, 5), (33, 17, 65), (6, 21, 130)])
@pytest.mark.parametrize("kind", ["values", "lookup"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_advection_kernels(dims, kind, dtype):
    """First-order upwind advection along +x: every pass stays exact."""
    v = 0.37
    box = la.BoxInfo(dims, p_start=1.0, p_end=0.0, velocity_x=v)
    if kind == "values":
        c = problems.lognormal(dims, 1.0, 1)
        A = la.StencilMatrix.values(box, c)
        S = O.Stencil(dims, coeff=c, p_start=1.0, p_end=0.0, velocity=v)
    else:
        m = problems.diagonal_lines(dims)
        mc = (0.0, 1.0, 2.0, 3.0)
        A = la.StencilMatrix.lookup(box, m, mc)
        S = O.Stencil(dims, materials=m, mat_coeffs=mc, p_start=1.0, p_end=0.0, velocity=v)
    n = S.N
    x = problems.uniform(n, 2, dtype)
    b = problems.uniform(n, 3, dtype)
    xd = la.DenseVector(x)
    same(A.spmv(xd, la.DenseVector(n, dtype)).numpy(), O.spmv(S, x))
    same(A.rhs(dtype).numpy(), S.rhs(dtype))
    same(A.residual(la.DenseVector(b), xd, la.DenseVector(n, dtype)).numpy(), O.residual(S, b, x))
    same(A.jacobi(xd, la.DenseVector(n, dtype), la.DenseVector(b), 0.7).numpy(), O.jacobi(S, x, b, dtype(0.7)))
    for cache in (True, False):
        A.set_sor_cache(cache)
        for forward in (True, False):
            for ultra in (False, True):
                for with_x in (True, False):
                    out = A.sor(xd if with_x else None, la.DenseVector(n, dtype), la.DenseVector(b), 1.1,
                                forward, ultra)
                    same(out.numpy(), O.sor(S, x if with_x else None, b, dtype(1.1), forward, ultra))


@pytest.mark.parametrize("dims", [(33, 17, 65), (64, 64, 64)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("velocity", [0.0, 0.3])
def test_kernels_within_rounding_of_the_vectorised_order(dims, dtype, velocity):
    """The reference's vectorised build accumulates each row in the order
    -x,+x,-y,+y,-z,+z, diagonal last (its SOR through a recurrence along z);
    the kernels follow its scalar order.  The two differ by rounding only:
    SpMV by at most 16 eps of the row's magnitude sum |a_ij x_j| + |y_i|,
    one SOR sweep by at most 64 eps of the largest value."""
    box = la.BoxInfo(dims, p_start=1.0, p_end=0.0, velocity_x=velocity)
    c = problems.lognormal(dims, 1.0, 7)
    A = la.StencilMatrix.values(box, c)
    S = O.Stencil(dims, coeff=c, p_start=1.0, p_end=0.0, velocity=velocity)
    n = S.N
    eps = np.finfo(dtype).eps
    x = problems.uniform(n, 8, dtype)
    b = problems.uniform(n, 9, dtype)
    y = A.spmv(la.DenseVector(x), la.DenseVector(n, dtype)).numpy().astype(np.float64)
    C = S.to_csr()[0]
    mag = O.spmv(O.Csr(C.rowptr, C.cols, np.abs(C.vals), symmetric=False), np.abs(x).astype(np.float64))
    for fused in (True, False):
        d = np.abs(O.spmv_banded(S, x, fused).astype(np.float64) - y)
        assert np.all(d <= 16 * eps * (mag + np.abs(y))), f"fused={fused}: {d.max()}"
    for forward in (True, False):
        for ultra in (False, True):
            out = A.sor(la.DenseVector(x), la.DenseVector(n, dtype), la.DenseVector(b), 1.0, forward,
                        ultra).numpy().astype(np.float64)
            ref = O.sor_banded(S, x, b, dtype(1.0), forward, ultra).astype(np.float64)
            assert np.abs(out - ref).max() <= 64 * eps * np.abs(ref).max(), f"forward={forward} ultra={ultra}"
