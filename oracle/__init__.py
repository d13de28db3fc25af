"""ctypes front end of the CPU oracle (liblineal_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg as the checker.  The product package
``lineal_amd`` never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblineal_oracle.so")

i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)


class StencilStruct(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("len", C.c_double * 3),
        ("kind", C.c_int),
        ("zero_to_one", C.c_int),
        ("coeff", f64p),
        ("mat", u8p),
        ("mat_coeffs", C.c_double * 4),
        ("p_start", C.c_double),
        ("p_end", C.c_double),
        ("velocity", C.c_double),
    ]


class CsrStruct(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("rowptr", i64p),
        ("cols", i64p),
        ("vals", f64p),
        ("is_f32", C.c_int),
        ("symmetric", C.c_int),
    ]


class MatStruct(C.Structure):
    _fields_ = [
        ("is_csr", C.c_int),
        ("st", C.POINTER(StencilStruct)),
        ("csr", C.POINTER(CsrStruct)),
    ]


class AmgParams(C.Structure):
    _fields_ = [
        ("max_level_num", C.c_int),
        ("min_coarsen_size", C.c_int64),
        ("min_coarsen_factor", C.c_double),
        ("strong_threshold", C.c_double),
        ("aggressive_strong_threshold", C.c_double),
        ("min_aggregate_size", C.c_int64),
        ("max_aggregate_size", C.c_int64),
        ("coarse_f32", C.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    cmd = ["make", "-s", "-C", _HERE]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sp = C.POINTER(StencilStruct)
        mp = C.POINTER(MatStruct)
        cp = C.POINTER(CsrStruct)
        hp = C.c_void_p
        sigs = {
            "or_axis_quotients": (None, [sp, f64p]),
            "or_harmonic_mean_f64": (C.c_double, [C.c_double, C.c_double]),
            "or_harmonic_mean_f32": (C.c_float, [C.c_float, C.c_float]),
            "or_stencil_row_f64": (C.c_int, [sp, C.c_int64, i64p, f64p]),
            "or_stencil_row_f32": (C.c_int, [sp, C.c_int64, i64p, f32p]),
            "or_stencil_rhs_f64": (None, [sp, f64p]),
            "or_stencil_rhs_f32": (None, [sp, f32p]),
            "or_spmv_banded_f64": (None, [sp, f64p, f64p, C.c_int]),
            "or_spmv_banded_f32": (None, [sp, f32p, f32p, C.c_int]),
            "or_sor_banded_f64": (None, [sp, f64p, f64p, f64p, C.c_double, C.c_int, C.c_int]),
            "or_sor_banded_f32": (None, [sp, f32p, f32p, f32p, C.c_float, C.c_int, C.c_int]),
            "or_spmv_f64": (None, [mp, f64p, f64p]),
            "or_spmv_f32": (None, [mp, f32p, f32p]),
            "or_residual_f64": (None, [mp, f64p, f64p, f64p]),
            "or_residual_f32": (None, [mp, f32p, f32p, f32p]),
            "or_jacobi_f64": (None, [mp, f64p, f64p, f64p, C.c_double]),
            "or_jacobi_f32": (None, [mp, f32p, f32p, f32p, C.c_float]),
            "or_sor_f64": (None, [mp, f64p, f64p, f64p, C.c_double, C.c_int, C.c_int, C.c_int, i64p]),
            "or_sor_f32": (None, [mp, f32p, f32p, f32p, C.c_float, C.c_int, C.c_int, C.c_int, i64p]),
            "or_dot_f64": (C.c_double, [f64p, f64p, C.c_int64]),
            "or_csr_from_stencil": (C.c_int64, [sp, C.c_int, i64p, i64p, i64p, f64p, f64p, i64p]),
            "or_amg_setup": (hp, [mp, C.POINTER(AmgParams)]),
            "or_strength_graph": (C.c_int64, [mp, C.c_double, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
            "or_amg_free": (None, [hp]),
            "or_amg_level_num": (C.c_int, [hp]),
            "or_amg_level_size": (C.c_int64, [hp, C.c_int]),
            "or_amg_agg_map": (u32p, [hp, C.c_int]),
            "or_amg_c2f_ptr": (i64p, [hp, C.c_int]),
            "or_amg_c2f_idx": (i64p, [hp, C.c_int]),
            "or_amg_matrix": (cp, [hp, C.c_int]),
            "or_amg_cholesky": (cp, [hp]),
            "or_amg_stats": (None, [hp, C.c_int, i64p]),
            "or_amg_set_coarse_segment": (None, [hp, C.c_int64]),
            "or_amg_set_level_segment": (None, [hp, C.c_int, C.c_int64]),
            "or_amg_set_cycle": (None, [hp, C.c_double, C.c_int, C.c_int]),
            "or_cholesky_solve": (None, [cp, f64p, f64p]),
            "or_cholesky_decompose": (C.c_int64, [cp, i64p, i64p, f64p]),
            "or_vcycle_f64": (None, [hp, f64p, f64p, C.c_int]),
            "or_vcycle_f32": (None, [hp, f32p, f32p, C.c_int]),
            "or_cycle_f64": (None, [hp, f64p, f64p, f64p, C.c_int]),
            "or_cycle_f32": (None, [hp, f32p, f32p, f32p, C.c_int]),
            "or_pcg": (C.c_int, [mp, hp, C.c_int, C.c_int, f64p, f64p, C.c_double, C.c_int, f64p]),
            "or_pbicgstab": (C.c_int, [mp, hp, C.c_int, C.c_int, f64p, f64p, C.c_double, C.c_int, f64p]),
            "or_amg_is_lu": (C.c_int, [hp]),
            "or_lu_decompose_dense": (C.c_int, [cp, f64p, f64p]),
            "or_lu_solve_dense": (None, [C.c_int64, f64p, f64p, f64p, f64p]),
            "or_outlet_flux": (C.c_double, [sp, f64p]),
            "or_outlet_flux_adjoint_rhs": (None, [sp, f64p]),
            "or_outlet_flux_gradient": (None, [sp, f64p, f64p, f64p]),
            "or_ssor_f64": (None, [mp, f64p, f64p, f64p, C.c_double, C.c_int, C.c_int]),
            "or_eff_diff_area_part": (C.c_double, [sp, C.POINTER(C.c_uint8)]),
            "or_eff_diff": (C.c_double, [sp, i64p, f64p, C.c_double]),
            "or_bcsr_spmv_f64": (None, [C.c_int64, C.c_int, i64p, i64p, f64p, f64p, f64p]),
            "or_bcsr_spmv_f32": (None, [C.c_int64, C.c_int, i64p, i64p, f64p, f32p, f32p]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray | None, t):
    if a is None:
        return C.cast(None, C.POINTER(t))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(t))


def _pf(a: np.ndarray | None):
    if a is None:
        return None
    if a.dtype == np.float64:
        return _p(a, C.c_double)
    assert a.dtype == np.float32
    return _p(a, C.c_float)


class Stencil:
    """Matrix-free diffusion stencil (values or lookup valuator)."""

    def __init__(self, dims, coeff=None, materials=None, mat_coeffs=(0.0, 1.0, 2.0, 3.0),
                 lengths=None, zero_to_one=True, p_start=1.0, p_end=0.0, velocity=0.0):
        self.dims = tuple(int(d) for d in dims)
        self.N = int(np.prod(self.dims))
        s = StencilStruct()
        for d in range(3):
            s.n[d] = self.dims[d]
            s.len[d] = float(lengths[d]) if lengths is not None else float(self.dims[d])
        s.zero_to_one = int(bool(zero_to_one))
        s.p_start, s.p_end = float(p_start), float(p_end)
        s.velocity = float(velocity)
        self.velocity = float(velocity)
        self.coeff = self.materials = None
        if coeff is not None:
            self.coeff = np.ascontiguousarray(coeff, dtype=np.float64).reshape(-1)
            assert self.coeff.size == self.N
            s.kind = 0
            s.coeff = _p(self.coeff, C.c_double)
        else:
            self.materials = np.ascontiguousarray(materials, dtype=np.uint8).reshape(-1)
            assert self.materials.size == self.N
            s.kind = 1
            s.mat = _p(self.materials, C.c_uint8)
            for k in range(4):
                s.mat_coeffs[k] = float(mat_coeffs[k])
        self.mat_coeffs = tuple(float(c) for c in mat_coeffs)
        self.lengths = tuple(s.len[d] for d in range(3))
        self.zero_to_one = bool(zero_to_one)
        self.p_start, self.p_end = float(p_start), float(p_end)
        self.s = s
        self.m = MatStruct(0, C.pointer(self.s), None)

    @property
    def mat(self):
        return self.m

    def quotients(self):
        q = np.zeros(3)
        lib().or_axis_quotients(C.byref(self.s), _p(q, C.c_double))
        return q

    def row(self, i, dtype=np.float64):
        cols = np.zeros(8, np.int64)
        vals = np.zeros(8, dtype)
        fn = lib().or_stencil_row_f64 if dtype == np.float64 else lib().or_stencil_row_f32
        k = fn(C.byref(self.s), int(i), _p(cols, C.c_int64), _pf(vals))
        return cols[:k].copy(), vals[:k].copy()

    def rhs(self, dtype=np.float64):
        b = np.zeros(self.N, dtype)
        fn = lib().or_stencil_rhs_f64 if dtype == np.float64 else lib().or_stencil_rhs_f32
        fn(C.byref(self.s), _pf(b))
        return b

    def to_csr(self, compress=False):
        L = lib()
        rows = C.c_int64()
        nnz = L.or_csr_from_stencil(C.byref(self.s), int(compress), C.byref(rows), None, None,
                                    None, None, None)
        rp = np.zeros(rows.value + 1, np.int64)
        cols = np.zeros(max(nnz, 1), np.int64)
        vals = np.zeros(max(nnz, 1), np.float64)
        rhs = np.zeros(rows.value, np.float64)
        cmap = np.zeros(self.N, np.int64)
        L.or_csr_from_stencil(C.byref(self.s), int(compress), C.byref(rows), _p(rp, C.c_int64),
                              _p(cols, C.c_int64), _p(vals, C.c_double), _p(rhs, C.c_double),
                              _p(cmap, C.c_int64) if compress else None)
        m = Csr(rp, cols[:nnz], vals[:nnz], symmetric=self.velocity == 0.0)
        return m, rhs, (cmap if compress else None)


class Csr:
    def __init__(self, rowptr, cols, vals, symmetric=True, is_f32=False):
        self.rowptr = np.ascontiguousarray(rowptr, np.int64)
        self.cols = np.ascontiguousarray(cols, np.int64)
        self.vals = np.ascontiguousarray(vals, np.float64)
        self.N = self.rowptr.size - 1
        self.c = CsrStruct(self.N, _p(self.rowptr, C.c_int64), _p(self.cols, C.c_int64),
                           _p(self.vals, C.c_double), int(is_f32), int(symmetric))
        self.m = MatStruct(1, None, C.pointer(self.c))

    @property
    def mat(self):
        return self.m

    def dense(self):
        A = np.zeros((self.N, self.N))
        for i in range(self.N):
            for k in range(self.rowptr[i], self.rowptr[i + 1]):
                A[i, self.cols[k]] += self.vals[k]
        return A

    @staticmethod
    def from_struct(cs: CsrStruct):
        n = cs.n
        rp = np.ctypeslib.as_array(cs.rowptr, shape=(n + 1,)).copy()
        nnz = int(rp[-1])
        cols = np.ctypeslib.as_array(cs.cols, shape=(max(nnz, 1),))[:nnz].copy()
        vals = np.ctypeslib.as_array(cs.vals, shape=(max(nnz, 1),))[:nnz].copy()
        return Csr(rp, cols, vals, symmetric=bool(cs.symmetric), is_f32=bool(cs.is_f32))


def _fn(name, dtype):
    return getattr(lib(), name + ("_f64" if dtype == np.float64 else "_f32"))


def spmv(A, x):
    y = np.zeros_like(x)
    _fn("or_spmv", x.dtype)(C.byref(A.mat), _pf(x), _pf(y))
    return y


def spmv_banded(S, x, fused=True):
    """y = A x in the reference's vectorised row order (banded_iterate):
    -x, +x, -y, +y, -z, +z, the diagonal last; fused or separately rounded
    multiply-adds (a compiler may contract `sum += a * x` or not)."""
    y = np.zeros_like(x)
    _fn("or_spmv_banded", x.dtype)(C.byref(S.s), _pf(x), _pf(y), int(fused))
    return y


def sor_banded(S, x_in, b, omega=1.0, forward=True, ultra=False):
    """One lexicographic SOR sweep in the reference's vectorised form
    (compute_base_simd: w / d formed first, the z-recurrence x = add - factor
    x_prev)."""
    x_out = np.zeros_like(b)
    xin = None if x_in is None else np.ascontiguousarray(x_in, b.dtype)
    _fn("or_sor_banded", b.dtype)(C.byref(S.s), _pf(xin), _pf(x_out), _pf(b), omega, int(forward),
                                  int(ultra))
    return x_out


def residual(A, b, x):
    r = np.zeros_like(x)
    _fn("or_residual", x.dtype)(C.byref(A.mat), _pf(b), _pf(x), _pf(r))
    return r


def jacobi(A, x_in, b, omega):
    x_out = np.zeros_like(b)
    _fn("or_jacobi", b.dtype)(C.byref(A.mat), _pf(x_in), _pf(x_out), _pf(b), omega)
    return x_out


def sor(A, x_in, b, omega=1.0, forward=True, ultra=False, segments=None):
    n = b.size
    seg = np.asarray(segments if segments is not None else [0, n], np.int64)
    x_out = np.zeros_like(b)
    _fn("or_sor", b.dtype)(C.byref(A.mat), _pf(x_in), _pf(x_out), _pf(b), omega,
                           int(forward), int(ultra), seg.size - 1, _p(seg, C.c_int64))
    return x_out


def ssor(A, x_in, b, omega=1.0, ultra=False, iterations=1):
    """SsorSolver steps (gauss-seidel-variant/solver.hpp:45-96): forward then
    backward lexicographic sweeps, `iterations` times; f64."""
    b = np.ascontiguousarray(b, np.float64)
    x_out = np.zeros_like(b)
    xin = None if x_in is None else np.ascontiguousarray(x_in, np.float64)
    lib().or_ssor_f64(C.byref(A.mat), _pf(xin), _pf(x_out), _pf(b), float(omega), int(ultra),
                      int(iterations))
    return x_out


def dot(a, b):
    return lib().or_dot_f64(_pf(a), _pf(b), a.size)


def strength_graph(A, theta=1.0 / 3.0):
    """Edge flags (bit 0 depends, bit 1 influences) of the off-diagonal
    entries in row order and the isolated flag of every vertex."""
    L = lib()
    ne = L.or_strength_graph(C.byref(A.mat), theta, None, None)
    props = np.zeros(max(ne, 1), np.uint8)
    iso = np.zeros(max(A.N, 1), np.uint8)
    u8 = C.POINTER(C.c_uint8)
    L.or_strength_graph(C.byref(A.mat), theta, props.ctypes.data_as(u8), iso.ctypes.data_as(u8))
    return props[:ne], iso[: A.N]


# coarse levels of at most this many rows are swept in one segment (the
# library's default lgpu_amg_params.sequential_rows; 0: none)
SEQUENTIAL_ROWS = 512


def default_params(coarse_f32=False):
    """Parameters of test/aux/default.hpp:170-188."""
    return AmgParams(16, 32, 1.2, 1.0 / 3.0, 1.0 / 3.0, 6, 9, int(coarse_f32))


class Hierarchy:
    def __init__(self, A, params: AmgParams | None = None, coarse_f32=False, coarse_segment=0,
                 relax=1.0, cycle_kind=0, cycle_iterations=1, level_segments=None,
                 sequential_rows=SEQUENTIAL_ROWS):
        """relax, cycle_kind (0 REGULAR, 1 DECREASING) and cycle_iterations:
        MultiGridSolver{relax, DefaultCycle{kind, iterations}}.  coarse_segment:
        rows per smoother segment on the coarse levels (0: one segment);
        coarse levels of at most sequential_rows rows are swept in one segment;
        level_segments: {level: rows} overriding both on single levels."""
        self.A = A  # keep the finest operator alive
        self.params = params or default_params(coarse_f32)
        self.h = lib().or_amg_setup(C.byref(A.mat), C.byref(self.params))
        if not self.h:
            raise ValueError("the coarsest factorisation failed")
        lib().or_amg_set_coarse_segment(self.h, int(coarse_segment))
        for lvl in range(1, self.level_num):
            if self.size(lvl) <= sequential_rows:
                lib().or_amg_set_level_segment(self.h, lvl, 0)
        for lvl, rows in (level_segments or {}).items():
            lib().or_amg_set_level_segment(self.h, int(lvl), int(rows))
        lib().or_amg_set_cycle(self.h, float(relax), int(cycle_kind), int(cycle_iterations))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_amg_free(self.h)
            self.h = None

    @property
    def level_num(self):
        return lib().or_amg_level_num(self.h)

    def size(self, level):
        return lib().or_amg_level_size(self.h, level)

    def agg_map(self, level):
        n = self.size(level)
        return np.ctypeslib.as_array(lib().or_amg_agg_map(self.h, level), shape=(n,)).copy()

    def c2f(self, level):
        nc = self.size(level + 1)
        ptr = np.ctypeslib.as_array(lib().or_amg_c2f_ptr(self.h, level), shape=(nc + 1,)).copy()
        m = int(ptr[-1])
        idx = np.ctypeslib.as_array(lib().or_amg_c2f_idx(self.h, level), shape=(max(m, 1),))[:m].copy()
        return ptr, idx

    def matrix(self, level):
        return Csr.from_struct(lib().or_amg_matrix(self.h, level).contents)

    def cholesky(self):
        return Csr.from_struct(lib().or_amg_cholesky(self.h).contents)

    def stats(self, level):
        out = np.zeros(5, np.int64)
        lib().or_amg_stats(self.h, level, _p(out, C.c_int64))
        return out

    @property
    def is_lu(self):
        """True if the coarsest level is solved by LU (nonsymmetric hierarchy)."""
        return bool(lib().or_amg_is_lu(self.h))

    def vcycle(self, r, ultra=False, x_in=None):
        """One cycle z = M r from zero, or from the initial guess x_in."""
        z = np.zeros_like(r)
        if x_in is None:
            _fn("or_vcycle", r.dtype)(self.h, _pf(r), _pf(z), int(ultra))
        else:
            x_in = np.ascontiguousarray(x_in, dtype=r.dtype)
            _fn("or_cycle", r.dtype)(self.h, _pf(x_in), _pf(r), _pf(z), int(ultra))
        return z


def cholesky_solve(L: Csr, b):
    x = np.zeros_like(b)
    lib().or_cholesky_solve(C.byref(L.c), _pf(b), _pf(x))
    return x


def lu_dense(A: Csr):
    """Dense row-wise LU without pivoting: (unit lower L, upper U)."""
    n = A.N
    lo = np.zeros((n, n))
    up = np.zeros((n, n))
    if lib().or_lu_decompose_dense(C.byref(A.c), _pf(lo), _pf(up)) != 0:
        raise ValueError("zero pivot")
    return lo, up


def lu_solve_dense(lo, up, b):
    x = np.zeros_like(b)
    lib().or_lu_solve_dense(lo.shape[0], _pf(np.ascontiguousarray(lo)), _pf(np.ascontiguousarray(up)),
                            _pf(np.ascontiguousarray(b, dtype=np.float64)), _pf(x))
    return x


def cholesky_decompose(A: Csr):
    n = A.N
    cap = n * (n + 1) // 2 + 1
    rp = np.zeros(n + 1, np.int64)
    cols = np.zeros(cap, np.int64)
    vals = np.zeros(cap, np.float64)
    nnz = lib().or_cholesky_decompose(C.byref(A.c), _p(rp, C.c_int64), _p(cols, C.c_int64),
                                      _p(vals, C.c_double))
    if nnz < 0:
        raise ValueError("matrix is not SPD")
    return Csr(rp, cols[:nnz], vals[:nnz], symmetric=False)


_baseline_libs = {}


def baseline_lib(native=True):
    """The threaded CPU baseline (oracle/baseline.c): the build for this
    host's CPU when `native` (made on first use), else the portable one."""
    name = "liblineal_baseline_native.so" if native else "liblineal_baseline.so"
    if name not in _baseline_libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE, "baseline-native" if native else "baseline"],
                           check=True)
        L = C.CDLL(path)
        L.bl_pcg.restype = C.c_int
        L.bl_pcg.argtypes = [C.POINTER(MatStruct), C.c_void_p, f64p, f64p, C.c_double, C.c_int, C.c_int]
        _baseline_libs[name] = L
    return _baseline_libs[name]


class BaselineHierarchy:
    """The oracle's hierarchy built by the baseline library (the same code at
    -O3), for baseline_pcg."""

    def __init__(self, A, native=True):
        self.A, self.L = A, baseline_lib(native)
        self.L.or_amg_setup.restype = C.c_void_p
        self.L.or_amg_setup.argtypes = [C.POINTER(MatStruct), C.POINTER(AmgParams)]
        self.L.or_amg_free.argtypes = [C.c_void_p]
        self.params = default_params()
        self.h = self.L.or_amg_setup(C.byref(A.mat), C.byref(self.params))
        if not self.h:
            raise ValueError("the coarsest factorisation failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_amg_free(self.h)
            self.h = None


def baseline_pcg(A, b, hierarchy, rel_tol=1e-8, max_it=500, threads=0, native=True, x0=None):
    """PCG-AMG on `threads` host threads (0: all) with thread-segment SOR
    smoothing (oracle/baseline.c); `hierarchy` is a Hierarchy or a
    BaselineHierarchy; returns (x, iterations)."""
    x = np.zeros(A.N) if x0 is None else np.array(x0, dtype=np.float64)
    it = baseline_lib(native).bl_pcg(C.byref(A.mat), hierarchy.h, _pf(np.ascontiguousarray(b, np.float64)),
                                     _pf(x), float(rel_tol), int(max_it), int(threads))
    return x, it


def pcg(A, b, hierarchy=None, precon=0, ultra=False, rel_tol=1e-8, max_it=500, x0=None):
    x = np.zeros_like(b) if x0 is None else x0.astype(np.float64).copy()
    hist = np.zeros(max_it + 1)
    it = lib().or_pcg(C.byref(A.mat), hierarchy.h if hierarchy else None, precon, int(ultra),
                      _pf(b), _pf(x), rel_tol, max_it, _pf(hist))
    return x, it, hist[: it + 1]


def pbicgstab(A, b, hierarchy=None, precon=0, ultra=False, rel_tol=1e-8, max_it=500, x0=None):
    x = np.zeros_like(b) if x0 is None else x0.astype(np.float64).copy()
    hist = np.zeros(max_it + 1)
    it = lib().or_pbicgstab(C.byref(A.mat), hierarchy.h if hierarchy else None, precon,
                            int(ultra), _pf(b), _pf(x), rel_tol, max_it, _pf(hist))
    return x, it, hist[: it + 1]


def _require_values(S):
    if S.coeff is None or S.velocity != 0.0:
        raise ValueError("the outlet-flux adjoint needs per-cell coefficients and no advection")


def outlet_flux(S, p):
    """J = sum over the x = nx-1 cells of (2 q_0 c_i)(p_i - p_end)."""
    _require_values(S)
    p = np.ascontiguousarray(p, dtype=np.float64)
    return float(lib().or_outlet_flux(C.byref(S.s), _pf(p)))


def outlet_flux_adjoint_rhs(S):
    """dJ/dp: 2 q_0 c_i on the outlet cells, 0 elsewhere."""
    _require_values(S)
    g = np.zeros(S.N)
    lib().or_outlet_flux_adjoint_rhs(C.byref(S.s), _pf(g))
    return g


def outlet_flux_gradient(S, p, lam):
    """dJ/dlog c from the forward solution p and the adjoint solution lam."""
    _require_values(S)
    p = np.ascontiguousarray(p, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    g = np.zeros(S.N)
    lib().or_outlet_flux_gradient(C.byref(S.s), _pf(p), _pf(lam), _pf(g))
    return g


def adjoint_outlet_flux(S, hierarchy=None, precon=1, rel_tol=1e-10, max_it=1000):
    """FP64 forward and adjoint PCG solves plus the gradient: (J, grad, p, lam,
    forward iterations, adjoint iterations)."""
    _require_values(S)
    b = S.rhs()
    p, it_f, _ = pcg(S, b, hierarchy, precon=precon if hierarchy else 0, rel_tol=rel_tol,
                     max_it=max_it)
    g = outlet_flux_adjoint_rhs(S)
    lam, it_a, _ = pcg(S, g, hierarchy, precon=precon if hierarchy else 0, rel_tol=rel_tol,
                       max_it=max_it)
    return outlet_flux(S, p), outlet_flux_gradient(S, p, lam), p, lam, it_f, it_a


def eff_diff_area_part(S, none_mats=None):
    """EffDiffCalc's area part: filled cells of the mid plane x = nx // 2 and
    the plane after it, / (ny nz) * 0.5."""
    if S.dims[0] < 3:
        raise ValueError("the effective diffusivity needs at least three cells along x")
    flags = None
    if none_mats is not None:
        arr = np.ascontiguousarray([1 if f else 0 for f in none_mats], dtype=np.uint8)
        assert arr.size == 4
        flags = _p(arr, C.c_uint8)
    return float(lib().or_eff_diff_area_part(C.byref(S.s), flags))


def eff_diff(S, sol, area_part, cmap=None):
    """EffDiffCalc::operator(): the mid-plane flux normalised to an effective
    diffusivity; with cmap, sol is indexed by the compressed system's rows."""
    if S.dims[0] < 3:
        raise ValueError("the effective diffusivity needs at least three cells along x")
    sol = np.ascontiguousarray(sol, dtype=np.float64)
    cm = None
    if cmap is not None:
        cmap = np.ascontiguousarray(cmap, dtype=np.int64)
        assert cmap.size == S.N
        cm = _p(cmap, C.c_int64)
    return float(lib().or_eff_diff(C.byref(S.s), cm, _pf(sol), float(area_part)))


# ---------------------------------------------------------------------------
# matrix and vector files, io/tensor.hpp:25-263 (little-endian u64 / f64)


def smf_read(path):
    """SmfReader::read (io/tensor.hpp:47-75) with the row-wise builder's
    semantics (tensor/dynamic/csr/shared.hpp:730-764): per row, an entry whose
    column was seen before is added to the first one (f64, file order), then
    the row is sorted by column.  Returns (rows, columns, header nonzeros,
    rowptr, cols, vals)."""
    with open(path, "rb") as f:
        raw = f.read()
    rows, ncols, nnz = (int(v) for v in np.frombuffer(raw[:24], "<u8"))
    off = np.frombuffer(raw[24:24 + (rows + 1) * 8], "<u8").astype(np.int64)
    ent = np.frombuffer(raw[24 + (rows + 1) * 8:24 + (rows + 1) * 8 + int(off[-1]) * 16],
                        dtype=[("c", "<u8"), ("v", "<f8")])
    rp, cs, vs = [0], [], []
    for r in range(rows):
        acc = {}
        for c, v in zip(ent["c"][off[r]:off[r + 1]].tolist(), ent["v"][off[r]:off[r + 1]].tolist()):
            acc[c] = acc[c] + v if c in acc else v
        for c in sorted(acc):
            cs.append(c)
            vs.append(acc[c])
        rp.append(len(cs))
    return rows, ncols, nnz, np.array(rp, np.int64), np.array(cs, np.int64), np.array(vs, np.float64)


def smf_write(path, rows, ncols, rowptr, cols, vals, stored_count=True):
    """SmfWriter::write (io/tensor.hpp:158-213): nonzero entries only, in the
    given (column) order; the header counts the stored entries (stored_count,
    a CsrMatrix's non_zero_num) or else the written ones."""
    rowptr = np.asarray(rowptr, np.int64)
    vals = np.asarray(vals, np.float64)
    keep = vals != 0
    written = np.concatenate([[0], np.cumsum(keep)])[rowptr]
    head = np.array([rows, ncols, int(rowptr[-1]) if stored_count else int(written[-1])], "<u8")
    ent = np.empty(int(keep.sum()), dtype=[("c", "<u8"), ("v", "<f8")])
    ent["c"] = np.asarray(cols, np.int64)[keep]
    ent["v"] = vals[keep]
    with open(path, "wb") as f:
        f.write(head.tobytes() + written.astype("<u8").tobytes() + ent.tobytes())


def dvf_read(path):
    """DvfReader::read (io/tensor.hpp:226-232)."""
    with open(path, "rb") as f:
        raw = f.read()
    n = int(np.frombuffer(raw[:8], "<u8")[0])
    return np.frombuffer(raw[8:8 + 8 * n], "<f8").copy()


def dvf_write(path, v):
    """DvfWriter::write (io/tensor.hpp:253-263)."""
    v = np.asarray(v, np.float64)
    with open(path, "wb") as f:
        f.write(np.array([v.size], "<u8").tobytes() + v.astype("<f8").tobytes())


def bcsr_spmv(rowptr, cols, vals, x):
    """y = A x for a CSR matrix with B x B block values (vals: blocks x B x B);
    x f64 or f32 selects the variant."""
    rp = np.ascontiguousarray(rowptr, np.int64)
    ci = np.ascontiguousarray(cols, np.int64)
    va = np.ascontiguousarray(vals, np.float64)
    bsz = va.shape[1] if va.ndim == 3 else 1
    y = np.zeros_like(x)
    _fn("or_bcsr_spmv", x.dtype)(rp.size - 1, bsz, _p(rp, C.c_int64), _p(ci, C.c_int64),
                                 _p(va.reshape(-1), C.c_double), _pf(x), _pf(y))
    return y
