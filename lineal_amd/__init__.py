"""lineal_amd: MI355X (gfx950) implementation of lineal's solve-phase hot path.

This module loads ``liblineal_amd.so`` (hand-written HIP kernels behind a C
ABI, declared in ``include/lineal_amd.h``) through ctypes and mirrors the
reference's matrix / solver / preconditioner interface:

========================================  =============================================
reference (include/lineal/...)            here
========================================  =============================================
DenseVector                               DenseVector
AdjacentStencilMatrix + Values valuator   StencilMatrix.values
AdjacentStencilMatrix + Lookup valuator   StencilMatrix.lookup
AdjacentStencilVector                     StencilMatrix.rhs
CsrMatrix / csr_from_adjacent_stencil     CsrMatrix, StencilMatrix.to_csr
JacobiSolver / SorSolver                  JacobiSolver, SorSolver (+ .instantiate/.apply)
MultiGridSolver + HierarchyWright         MultiGridSolver(relax, HierarchyWrightParams)
ConjugateGradientsSolver                  ConjugateGradientsSolver
BiCgStabSolver                            BiCgStabSolver
run_iterative_solver                      run_iterative_solver
========================================  =============================================

There is no CPU fallback: importing works without a GPU, but every call that
reaches the library needs the built ``liblineal_amd.so`` and a GPU and raises
otherwise.
"""

from __future__ import annotations

import atexit
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LINEAL_AMD_LIBRARY selects another build of the same ABI (e.g. a profiling build)
LIB_PATH = os.environ.get("LINEAL_AMD_LIBRARY") or os.path.join(_HERE, "liblineal_amd.so")

F64, F32 = 0, 1
OK, INVALID_ARG, OUT_OF_MEMORY, HIP_ERROR, BREAKDOWN, NONFINITE, NOT_CONVERGED, NOT_SUPPORTED, SYNC_TIMEOUT = range(9)
# kernel classes of Context.kernel_time
KT_STENCIL_SOR_FWD, KT_STENCIL_SOR_BWD, KT_STENCIL_PASS, KT_CSR_SOR, KT_CSR_PASS = range(5)
KT_ALL = (1 << 5) - 1

_vp = C.c_void_p
_i64 = C.c_int64
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)


class LinealError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class _Box(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("len", C.c_double * 3),
        ("zero_to_one", C.c_int),
        ("p_start", C.c_double),
        ("p_end", C.c_double),
        ("velocity_x", C.c_double),
    ]


class _AmgParams(C.Structure):
    _fields_ = [
        ("max_level_num", C.c_int),
        ("min_coarsen_size", C.c_int64),
        ("min_coarsen_factor", C.c_double),
        ("strong_threshold", C.c_double),
        ("aggressive_strong_threshold", C.c_double),
        ("min_aggregate_size", C.c_int64),
        ("max_aggregate_size", C.c_int64),
        ("precision", C.c_int),
        ("ultra", C.c_int),
        ("coarse_segment_rows", C.c_int64),
        ("relax", C.c_double),
        ("cycle_kind", C.c_int),
        ("cycle_iterations", C.c_int),
        ("setup_on_device", C.c_int),
        ("sequential_rows", C.c_int64),
    ]


CYCLE_REGULAR, CYCLE_DECREASING = 0, 1
# lgpu_amg_params.sequential_rows by default: coarse levels of at most this
# many rows are swept in one segment (0: none)
SEQUENTIAL_ROWS = 512


class _AmgLevel(C.Structure):
    _fields_ = [
        ("agg", _u32p),
        ("c2f_ptr", _i64p),
        ("c2f_idx", _i32p),
        ("coarse_n", C.c_int64),
        ("rowptr", _i64p),
        ("cols", _i32p),
        ("vals", _f64p),
    ]


class _KrylovParams(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("max_it", C.c_int), ("rere_every", C.c_int)]


# exported symbol -> (restype, argtypes); the list mirrors include/lineal_amd.h
SIGNATURES = {
    "lgpu_status_string": (C.c_char_p, [C.c_int]),
    "lgpu_last_error": (C.c_char_p, []),
    "lgpu_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "lgpu_ctx_destroy": (C.c_int, [_vp]),
    "lgpu_ctx_synchronize": (C.c_int, [_vp]),
    "lgpu_timer_start": (C.c_int, [_vp]),
    "lgpu_timer_stop": (C.c_int, [_vp, _f64p]),
    "lgpu_vec_create": (C.c_int, [_vp, _i64, C.c_int, C.POINTER(_vp)]),
    "lgpu_vec_wrap": (C.c_int, [_vp, _vp, _i64, C.c_int, C.POINTER(_vp)]),
    "lgpu_vec_destroy": (C.c_int, [_vp]),
    "lgpu_vec_size": (C.c_int, [_vp, _i64p]),
    "lgpu_vec_data": (_vp, [_vp]),
    "lgpu_vec_upload": (C.c_int, [_vp, _vp, _i64]),
    "lgpu_vec_download": (C.c_int, [_vp, _vp, _i64]),
    "lgpu_vec_fill": (C.c_int, [_vp, C.c_double]),
    "lgpu_vec_copy": (C.c_int, [_vp, _vp]),
    "lgpu_vec_dot": (C.c_int, [_vp, _vp, _f64p]),
    "lgpu_stencil_values_create": (C.c_int, [_vp, C.POINTER(_Box), _f64p, C.POINTER(_vp)]),
    "lgpu_stencil_lookup_create": (C.c_int, [_vp, C.POINTER(_Box), _u8p, _f64p, C.POINTER(_vp)]),
    "lgpu_csr_create": (C.c_int, [_vp, _i64, _i64p, _i32p, _f64p, C.c_int, C.c_int, C.POINTER(_vp)]),
    "lgpu_csr_from_stencil": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "lgpu_csr_system_from_stencil": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_vp), C.POINTER(_vp), _i64p]),
    "lgpu_smf_info": (C.c_int, [C.c_char_p, _i64p, _i64p, _i64p]),
    "lgpu_smf_read": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int, C.POINTER(_vp)]),
    "lgpu_smf_write": (C.c_int, [_vp, C.c_char_p]),
    "lgpu_dvf_size": (C.c_int, [C.c_char_p, _i64p]),
    "lgpu_dvf_read": (C.c_int, [_vp, C.c_char_p, C.c_int, C.POINTER(_vp)]),
    "lgpu_dvf_write": (C.c_int, [_vp, C.c_char_p]),
    "lgpu_eff_diff_area_part": (C.c_int, [_vp, _u8p, _f64p]),
    "lgpu_eff_diff": (C.c_int, [_vp, _vp, _vp, C.c_double, _f64p]),
    "lgpu_op_destroy": (C.c_int, [_vp]),
    "lgpu_bcsr_create": (C.c_int, [_vp, _i64, C.c_int, _i64p, _i32p, _f64p, C.c_int, C.POINTER(_vp)]),
    "lgpu_bcsr_destroy": (C.c_int, [_vp]),
    "lgpu_bcsr_rows": (C.c_int, [_vp, _i64p, C.POINTER(C.c_int)]),
    "lgpu_bcsr_spmv": (C.c_int, [_vp, _vp, _vp]),
    "lgpu_op_rows": (C.c_int, [_vp, _i64p]),
    "lgpu_op_rhs": (C.c_int, [_vp, _vp]),
    "lgpu_op_spmv": (C.c_int, [_vp, _vp, _vp]),
    "lgpu_op_residual": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lgpu_op_jacobi": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double]),
    "lgpu_op_sor": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int]),
    "lgpu_op_set_sor_cache": (C.c_int, [_vp, C.c_int]),
    "lgpu_op_sor_segmented": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int, C.c_int, _i64]),
    "lgpu_op_ssor": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int]),
    "lgpu_op_ssor_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int]),
    "lgpu_ssor_precon_create": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int, C.POINTER(_vp)]),
    "lgpu_amg_default_params": (None, [C.POINTER(_AmgParams)]),
    "lgpu_amg_create": (C.c_int, [_vp, C.POINTER(_AmgParams), C.POINTER(_vp)]),
    "lgpu_amg_create_from_levels": (C.c_int, [_vp, C.POINTER(_AmgParams), C.c_int, _vp, C.POINTER(_vp)]),
    "lgpu_amg_transfer": (C.c_int, [_vp, C.c_int, _i64p, _i32p]),
    "lgpu_amg_destroy": (C.c_int, [_vp]),
    "lgpu_amg_level_num": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "lgpu_amg_level_size": (C.c_int, [_vp, C.c_int, _i64p]),
    "lgpu_amg_aggregates": (C.c_int, [_vp, C.c_int, _u32p]),
    "lgpu_amg_matrix": (C.c_int, [_vp, C.c_int, _i64p, _i64p, _i32p, _f64p]),
    "lgpu_amg_setup_seconds": (C.c_int, [_vp, _f64p]),
    "lgpu_amg_setup_split": (C.c_int, [_vp, _f64p]),
    "lgpu_amg_vcycle": (C.c_int, [_vp, _vp, _vp]),
    "lgpu_amg_cycle": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lgpu_cg_create": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "lgpu_bicgstab_create": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "lgpu_krylov_iterate": (C.c_int, [_vp]),
    "lgpu_krylov_residual_norm": (C.c_int, [_vp, _f64p]),
    "lgpu_krylov_rho": (C.c_int, [_vp, _f64p]),
    "lgpu_krylov_iteration": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "lgpu_krylov_destroy": (C.c_int, [_vp]),
    "lgpu_cg_solve": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(_KrylovParams), C.POINTER(C.c_int), _f64p]),
    "lgpu_bicgstab_solve": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(_KrylovParams), C.POINTER(C.c_int), _f64p]),
    "lgpu_ctx_kernel_timing": (C.c_int, [_vp, C.c_int]),
    "lgpu_ctx_set_wait_error": (C.c_int, [_vp]),
    "lgpu_ctx_kernel_time": (C.c_int, [_vp, C.c_int, _f64p, _i64p]),
    "lgpu_outlet_flux": (C.c_int, [_vp, _vp, _f64p]),
    "lgpu_outlet_flux_adjoint_rhs": (C.c_int, [_vp, _vp]),
    "lgpu_outlet_flux_gradient": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lgpu_adjoint_grad": (C.c_int, [_vp, _vp, C.POINTER(_KrylovParams), _vp, _vp, _vp, _f64p,
                                    C.POINTER(C.c_int)]),
}

_lib = None
_finalizing = False


@atexit.register
def _mark_finalizing():
    # objects collected during interpreter shutdown skip their destroy calls
    # (the process exit releases the device memory)
    global _finalizing
    _finalizing = True


def lib():
    """The loaded C ABI; raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LinealError(HIP_ERROR, f"{LIB_PATH} is missing: build it with lineal_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status: int, allow=()):
    if status != OK and status not in allow:
        msg = lib().lgpu_last_error().decode()
        raise LinealError(status, f"{lib().lgpu_status_string(status).decode()}: {msg}")
    return status


def _dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise ValueError(f"unsupported dtype {dt}")


class Context:
    """One HIP stream on one device (replaces the reference Env's execution policy)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().lgpu_ctx_create(device, C.byref(h)))
        self.h = h

    def synchronize(self):
        _check(lib().lgpu_ctx_synchronize(self.h))

    def timer_start(self):
        _check(lib().lgpu_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_double()
        _check(lib().lgpu_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def set_wait_error(self):
        """Set the sticky wait-error word: every bounded wait of the next
        wavefront sweeps gives up at once (tests of the timeout path); the
        entry point that ran them raises SYNC_TIMEOUT and clears the word."""
        _check(lib().lgpu_ctx_set_wait_error(self.h))

    def kernel_timing(self, enable):
        """Bracket every launch of the chosen kernel classes with HIP events
        (True: all, False: none, or a mask of 1 << KT_* bits); clears records."""
        mask = (KT_ALL if enable else 0) if isinstance(enable, bool) else int(enable)
        _check(lib().lgpu_ctx_kernel_timing(self.h, mask))

    def kernel_time(self, kernel_class: int):
        """(total milliseconds, launches) of a kernel class since timing was enabled."""
        ms, n = C.c_double(), C.c_int64()
        _check(lib().lgpu_ctx_kernel_time(self.h, int(kernel_class), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None) and not _finalizing:
            lib().lgpu_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DenseVector:
    """Device vector (tensor/dynamic/dense/vector/shared.hpp:35-362)."""

    def __init__(self, size_or_array, dtype=None, ctx: Context | None = None, fill=None):
        self.ctx = ctx or default_context()
        if isinstance(size_or_array, (int, np.integer)):
            self.n = int(size_or_array)
            self.dtype = np.dtype(dtype or np.float64)
            host = None
        else:
            if dtype is None:
                src = np.asarray(size_or_array)
                dtype = src.dtype if src.dtype in (np.float32, np.float64) else np.float64
            host = np.ascontiguousarray(size_or_array, dtype=dtype).reshape(-1)
            self.n = host.size
            self.dtype = host.dtype
        h = _vp()
        _check(lib().lgpu_vec_create(self.ctx.h, self.n, _dtype_code(self.dtype), C.byref(h)))
        self.h = h
        if host is not None:
            self.upload(host)
        elif fill is not None:
            self.fill(fill)

    @classmethod
    def _adopt(cls, h, n: int, dtype, ctx: "Context") -> "DenseVector":
        """Wrap a vector handle the library created."""
        v = cls.__new__(cls)
        v.ctx, v.n, v.dtype, v.h = ctx, int(n), np.dtype(dtype), h
        return v

    @property
    def size(self):
        return self.n

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype).reshape(-1)
        assert host.size == self.n
        _check(lib().lgpu_vec_upload(self.h, host.ctypes.data_as(_vp), self.n))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        _check(lib().lgpu_vec_download(self.h, out.ctypes.data_as(_vp), self.n))
        return out

    def fill(self, value: float):
        _check(lib().lgpu_vec_fill(self.h, float(value)))

    def assign(self, other: "DenseVector"):
        _check(lib().lgpu_vec_copy(other.h, self.h))

    def dot(self, other: "DenseVector") -> float:
        out = C.c_double()
        _check(lib().lgpu_vec_dot(self.h, other.h, C.byref(out)))
        return out.value

    def __del__(self):
        try:
            if getattr(self, "h", None) and not _finalizing:
                lib().lgpu_vec_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _vh(v):
    return None if v is None else v.h


class _Matrix:
    def __init__(self, h, ctx):
        self.h = h
        self.ctx = ctx

    @property
    def row_num(self) -> int:
        n = C.c_int64()
        _check(lib().lgpu_op_rows(self.h, C.byref(n)))
        return n.value

    column_num = row_num

    def spmv(self, x: DenseVector, y: DenseVector):
        """y = A x (MatrixVectorProductExpr)."""
        _check(lib().lgpu_op_spmv(self.h, x.h, y.h))
        return y

    def __mul__(self, x: DenseVector) -> DenseVector:
        y = DenseVector(x.n, x.dtype, self.ctx)
        return self.spmv(x, y)

    def residual(self, b: DenseVector, x: DenseVector, r: DenseVector):
        _check(lib().lgpu_op_residual(self.h, b.h, x.h, r.h))
        return r

    def jacobi(self, x_in, x_out, b, omega):
        _check(lib().lgpu_op_jacobi(self.h, _vh(x_in), x_out.h, b.h, float(omega)))
        return x_out

    def sor(self, x_in, x_out, b, omega=1.0, forward=True, ultra=False, segment_rows=0):
        """One SOR sweep; segment_rows > 0 (CSR only) uses thread segments of that many rows."""
        if segment_rows:
            _check(lib().lgpu_op_sor_segmented(self.h, _vh(x_in), x_out.h, b.h, float(omega), int(forward),
                                               int(ultra), int(segment_rows)))
        else:
            _check(lib().lgpu_op_sor(self.h, _vh(x_in), x_out.h, b.h, float(omega), int(forward), int(ultra)))
        return x_out

    def __del__(self):
        try:
            if getattr(self, "h", None) and not _finalizing:
                lib().lgpu_op_destroy(self.h)
                self.h = None
        except Exception:
            pass


@dataclass
class BoxInfo:
    """Equidistant box (discretization/box-equidistant.hpp:57-111), flow axis 0."""

    dims: tuple
    lengths: tuple | None = None
    zero_to_one: bool = True
    p_start: float = 1.0
    p_end: float = 0.0
    # first-order upwind advection along +x (not in the reference)
    velocity_x: float = 0.0

    def _struct(self):
        b = _Box()
        for d in range(3):
            b.n[d] = int(self.dims[d])
            b.len[d] = float(self.lengths[d]) if self.lengths is not None else float(self.dims[d])
        b.zero_to_one = int(self.zero_to_one)
        b.p_start, b.p_end = float(self.p_start), float(self.p_end)
        b.velocity_x = float(self.velocity_x)
        return b


class StencilMatrix(_Matrix):
    """AdjacentStencilMatrix (tensor/dynamic/stencil/matrix/shared.hpp:35-360)."""

    @classmethod
    def values(cls, box: BoxInfo, coeff, ctx: Context | None = None):
        ctx = ctx or default_context()
        c = np.ascontiguousarray(coeff, dtype=np.float64).reshape(-1)
        h = _vp()
        bs = box._struct()
        _check(lib().lgpu_stencil_values_create(ctx.h, C.byref(bs), c.ctypes.data_as(_f64p), C.byref(h)))
        m = cls(h, ctx)
        m.box = box
        return m

    @classmethod
    def lookup(cls, box: BoxInfo, materials, mat_coeffs=(0.0, 1.0, 2.0, 3.0), ctx: Context | None = None):
        ctx = ctx or default_context()
        mats = np.ascontiguousarray(materials, dtype=np.uint8).reshape(-1)
        mc = np.ascontiguousarray(mat_coeffs, dtype=np.float64)
        h = _vp()
        bs = box._struct()
        _check(lib().lgpu_stencil_lookup_create(ctx.h, C.byref(bs), mats.ctypes.data_as(_u8p),
                                                mc.ctypes.data_as(_f64p), C.byref(h)))
        m = cls(h, ctx)
        m.box = box
        return m

    def rhs(self, dtype=np.float64) -> DenseVector:
        """AdjacentStencilVector: the Dirichlet right-hand side."""
        b = DenseVector(self.row_num, dtype, self.ctx)
        _check(lib().lgpu_op_rhs(self.h, b.h))
        return b

    def set_sor_cache(self, enable: bool):
        """Cached per-cell face weights and diagonal for the SOR sweeps (the
        default), or evaluation on the fly (enable False frees the caches);
        the sweeps give the same bits either way."""
        _check(lib().lgpu_op_set_sor_cache(self.h, int(bool(enable))))
        return self

    def to_csr(self, dtype=np.float64) -> "CsrMatrix":
        h = _vp()
        _check(lib().lgpu_csr_from_stencil(self.h, _dtype_code(dtype), C.byref(h)))
        return CsrMatrix(h, self.ctx)


@dataclass
class CsrSystem:
    """The { lhs, rhs, compress_map } of csr_from_adjacent_stencil
    (tensor/dynamic/csr/create-from.hpp:37-239).  compress_map holds the row of
    every cell or -1 (None without compression)."""

    lhs: "CsrMatrix"
    rhs: DenseVector
    compress_map: np.ndarray | None


def csr_from_adjacent_stencil(A: StencilMatrix, dtype=np.float64, compress: bool = False) -> CsrSystem:
    """CSR copy of a stencil system; with compress, rows whose diagonal is zero
    (ZeroToZero) are dropped and the rest renumbered in order
    (create-from.hpp:54-55, 92-99, 197-229)."""
    lhs, rhs = _vp(), _vp()
    cmap = np.empty(A.row_num, np.int64)
    _check(lib().lgpu_csr_system_from_stencil(A.h, _dtype_code(dtype), int(compress), C.byref(lhs),
                                              C.byref(rhs), cmap.ctypes.data_as(_i64p)))
    M = CsrMatrix(lhs, A.ctx)
    b = DenseVector._adopt(rhs, M.row_num, dtype, A.ctx)
    return CsrSystem(M, b, cmap if compress else None)


class CsrMatrix(_Matrix):
    """CsrMatrix (tensor/dynamic/csr/shared.hpp:37-1103)."""

    @classmethod
    def from_arrays(cls, rowptr, cols, vals, dtype=np.float64, symmetric=True, ctx: Context | None = None):
        ctx = ctx or default_context()
        rp = np.ascontiguousarray(rowptr, np.int64)
        ci = np.ascontiguousarray(cols, np.int32)
        va = np.ascontiguousarray(vals, np.float64)
        h = _vp()
        _check(lib().lgpu_csr_create(ctx.h, rp.size - 1, rp.ctypes.data_as(_i64p), ci.ctypes.data_as(_i32p),
                                     va.ctypes.data_as(_f64p), _dtype_code(dtype), int(symmetric), C.byref(h)))
        return cls(h, ctx)


class BlockCsrMatrix:
    """CsrMatrix over B x B block values (tensor/dynamic/csr/shared.hpp:37-1103
    with fix::DenseMatrix<Real, B, B>, tensor/fixed/matrix.hpp:151); vectors
    hold block_rows * B entries."""

    def __init__(self, rowptr, cols, vals, dtype=np.float64, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        rp = np.ascontiguousarray(rowptr, np.int64)
        ci = np.ascontiguousarray(cols, np.int32)
        va = np.ascontiguousarray(vals, np.float64)
        self.block = 1 if va.ndim == 1 else va.shape[1]
        self.dtype = np.dtype(dtype)
        h = _vp()
        _check(lib().lgpu_bcsr_create(self.ctx.h, rp.size - 1, self.block, rp.ctypes.data_as(_i64p),
                                      ci.ctypes.data_as(_i32p), va.ctypes.data_as(_f64p),
                                      _dtype_code(dtype), C.byref(h)))
        self.h = h
        self.row_num = (rp.size - 1) * self.block

    def spmv(self, x: DenseVector, y: DenseVector):
        _check(lib().lgpu_bcsr_spmv(self.h, x.h, y.h))
        return y

    def __mul__(self, x: DenseVector) -> DenseVector:
        return self.spmv(x, DenseVector(x.n, x.dtype, self.ctx))

    def __del__(self):
        try:
            if getattr(self, "h", None) and not _finalizing:
                lib().lgpu_bcsr_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# stationary solvers (linear-solver/iterative-stationary)


class _SweepInstance:
    """A stationary solver instance whose apply is `iterations` sweeps:
    apply(sol_in, sol_out, rhs) ex-situ (sol_in None: the constant-zero
    guess) and apply_in_situ(sol, rhs) in-situ through an auxiliary vector
    (SingleSweepImpl, iterative-stationary/broom.hpp:27-123)."""

    def __init__(self, lhs, sweep, iterations):
        self.lhs, self._sweep, self.iterations = lhs, sweep, iterations
        self._aux = None

    def apply(self, sol_in, sol_out, rhs):
        self._sweep(sol_in, sol_out, rhs)
        for _ in range(self.iterations - 1):
            self.apply_in_situ_once(sol_out, rhs)
        return sol_out

    def apply_in_situ_once(self, sol, rhs):
        if self._aux is None or self._aux.dtype != sol.dtype:
            self._aux = DenseVector(sol.n, sol.dtype, sol.ctx)
        self._aux.assign(sol)
        self._sweep(self._aux, sol, rhs)

    def apply_in_situ(self, sol, rhs):
        for _ in range(self.iterations):
            self.apply_in_situ_once(sol, rhs)
        return sol


class JacobiSolver:
    """JacobiSolver (iterative-stationary/jacobi/solver.hpp:21-39); `iterations`
    sweeps per apply (FixedIterationManager)."""

    def __init__(self, relax: float, iterations: int = 1):
        self.relax, self.iterations = relax, iterations

    def instantiate(self, lhs: _Matrix):
        return _SweepInstance(lhs, lambda i, o, b: lhs.jacobi(i, o, b, self.relax), self.iterations)


class SorSolver:
    """SorSolver<Real, IterMan, dir, variant>{relax[, IterMan{iterations}]}
    (gauss-seidel-variant/solver.hpp:22-41) with one thread segment."""

    def __init__(self, relax: float, forward: bool = True, ultra: bool = False, iterations: int = 1):
        self.relax, self.forward, self.ultra, self.iterations = relax, forward, ultra, iterations

    def instantiate(self, lhs: _Matrix):
        return _SweepInstance(lhs, lambda i, o, b: lhs.sor(i, o, b, self.relax, self.forward, self.ultra),
                              self.iterations)


class SsorSolver:
    """SsorSolver<Real, IterMan, variant>{relax, IterMan{iterations}}
    (gauss-seidel-variant/solver.hpp:45-96): each apply runs `iterations`
    SSOR steps (a forward and a backward lexicographic sweep).  An instance is
    also a Krylov preconditioner (applied from a zero initial guess)."""

    def __init__(self, relax: float = 1.0, ultra: bool = False, iterations: int = 1):
        self.relax, self.ultra, self.iterations = relax, ultra, iterations

    def instantiate(self, lhs: _Matrix) -> "SsorInstance":
        return SsorInstance(lhs, self)


class SsorInstance:
    def __init__(self, lhs: _Matrix, solver: SsorSolver):
        self.lhs, self.solver = lhs, solver
        h = _vp()
        _check(lib().lgpu_ssor_precon_create(lhs.h, float(solver.relax), int(solver.ultra),
                                             int(solver.iterations), C.byref(h)))
        self.h = h

    def apply(self, sol_in, sol_out, rhs):
        """sol_out = SSOR^iterations(sol_in); sol_in may be None (zero) or sol_out."""
        s = self.solver
        _check(lib().lgpu_op_ssor(self.lhs.h, _vh(sol_in), sol_out.h, rhs.h, float(s.relax), int(s.ultra)))
        for _ in range(s.iterations - 1):
            _check(lib().lgpu_op_ssor(self.lhs.h, sol_out.h, sol_out.h, rhs.h, float(s.relax), int(s.ultra)))
        return sol_out

    def apply_in_situ(self, sol, rhs):
        return self.apply(sol, sol, rhs)

    def __del__(self):
        try:
            if getattr(self, "h", None) and not _finalizing:
                lib().lgpu_amg_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# AMG


@dataclass
class HierarchyWrightParams:
    """HierarchyWright + aggregation + criterion parameters (test/aux/default.hpp:170-188)."""

    max_level_num: int = 16
    min_coarsen_size: int = 32
    min_coarsen_factor: float = 1.2
    strong_threshold: float = 1.0 / 3.0
    aggressive_strong_threshold: float = 1.0 / 3.0
    min_aggregate_size: int = 6
    max_aggregate_size: int = 9
    # the strength criterion and Galerkin products on the GPU (True) or on
    # host threads; the hierarchy is the same either way
    setup_on_device: bool = True

    def _struct(self, precision: int, ultra: bool, coarse_segment_rows: int, relax=1.0,
                cycle=(CYCLE_REGULAR, 1), sequential_rows=None):
        seq = SEQUENTIAL_ROWS if sequential_rows is None else sequential_rows
        return _AmgParams(self.max_level_num, self.min_coarsen_size, self.min_coarsen_factor,
                          self.strong_threshold, self.aggressive_strong_threshold,
                          self.min_aggregate_size, self.max_aggregate_size, precision, int(ultra),
                          int(coarse_segment_rows), float(relax), int(cycle[0]), int(cycle[1]),
                          int(self.setup_on_device), int(seq))


class MultiGridSolver:
    """MultiGridSolver{relax, DefaultCycle{kind, iterations}, wright}
    (multigrid/solver.hpp:183-196, multigrid/cycle.hpp:24-79).

    precision: np.float64 for an FP64 cycle, np.float32 for the FP32 hierarchy,
    smoothers and level operators with an FP64 coarsest solve.
    coarse_segment_rows: rows per thread segment of the coarse-level smoothers
    (0: one lexicographic segment per level).
    sequential_rows: coarse levels with at most this many rows are swept in one
    segment whatever coarse_segment_rows says (0: none).
    relax: the coarse correction is added scaled by 1/relax.
    cycle: (CYCLE_REGULAR or CYCLE_DECREASING, iterations); (REGULAR, 1) is the V-cycle.
    """

    def __init__(self, wright: HierarchyWrightParams | None = None, precision=np.float64, ultra=False,
                 coarse_segment_rows=64, relax=1.0, cycle=(CYCLE_REGULAR, 1), sequential_rows=None):
        self.wright = wright or HierarchyWrightParams()
        self.precision = precision
        self.ultra = ultra
        self.coarse_segment_rows = coarse_segment_rows
        self.sequential_rows = SEQUENTIAL_ROWS if sequential_rows is None else int(sequential_rows)
        self.relax = relax
        self.cycle = tuple(cycle)

    def instantiate(self, lhs: _Matrix) -> "MultiGridInstance":
        return MultiGridInstance(lhs, self)


@dataclass
class AmgLevel:
    """One transfer of a hierarchy built elsewhere: the aggregate map of the
    finer level (0xFFFFFFFE: isolated), optional coarse -> fine lists in
    summation order, and the coarse matrix as CSR arrays."""

    agg: np.ndarray
    rowptr: np.ndarray
    cols: np.ndarray
    vals: np.ndarray
    c2f_ptr: np.ndarray | None = None
    c2f_idx: np.ndarray | None = None


class MultiGridInstance:
    def __init__(self, lhs: _Matrix, solver: MultiGridSolver, levels: list[AmgLevel] | None = None):
        self.lhs = lhs
        p = solver.wright._struct(_dtype_code(solver.precision), solver.ultra, solver.coarse_segment_rows,
                                  solver.relax, solver.cycle, solver.sequential_rows)
        h = _vp()
        if levels is None:
            _check(lib().lgpu_amg_create(lhs.h, C.byref(p), C.byref(h)))
        else:
            keep, arr = [], (_AmgLevel * len(levels))()
            for k, lv in enumerate(levels):
                agg = np.ascontiguousarray(lv.agg, np.uint32)
                rp = np.ascontiguousarray(lv.rowptr, np.int64)
                cols = np.ascontiguousarray(lv.cols, np.int32)
                vals = np.ascontiguousarray(lv.vals, np.float64)
                keep += [agg, rp, cols, vals]
                arr[k].agg, arr[k].rowptr = agg.ctypes.data_as(_u32p), rp.ctypes.data_as(_i64p)
                arr[k].cols, arr[k].vals = cols.ctypes.data_as(_i32p), vals.ctypes.data_as(_f64p)
                arr[k].coarse_n = rp.size - 1
                if lv.c2f_ptr is not None:
                    cp = np.ascontiguousarray(lv.c2f_ptr, np.int64)
                    ci = np.ascontiguousarray(lv.c2f_idx, np.int32)
                    keep += [cp, ci]
                    arr[k].c2f_ptr, arr[k].c2f_idx = cp.ctypes.data_as(_i64p), ci.ctypes.data_as(_i32p)
            _check(lib().lgpu_amg_create_from_levels(lhs.h, C.byref(p), len(levels), arr, C.byref(h)))
        self.h = h

    @classmethod
    def from_levels(cls, lhs: _Matrix, levels: list[AmgLevel], solver: MultiGridSolver | None = None):
        """A hierarchy built elsewhere (lgpu_amg_create_from_levels)."""
        return cls(lhs, solver or MultiGridSolver(), levels)

    def transfer(self, level: int):
        """(c2f_ptr, c2f_idx) of transfer `level`."""
        nc = self.level_size(level + 1)
        ptr = np.empty(nc + 1, np.int64)
        _check(lib().lgpu_amg_transfer(self.h, level, ptr.ctypes.data_as(_i64p), None))
        idx = np.empty(max(int(ptr[-1]), 1), np.int32)
        _check(lib().lgpu_amg_transfer(self.h, level, ptr.ctypes.data_as(_i64p), idx.ctypes.data_as(_i32p)))
        return ptr, idx[: int(ptr[-1])]

    def export_levels(self) -> list[AmgLevel]:
        """This hierarchy's transfers, as lgpu_amg_create_from_levels takes them."""
        out = []
        for l in range(self.level_num - 1):
            rp, cols, vals = self.matrix(l + 1)
            ptr, idx = self.transfer(l)
            out.append(AmgLevel(self.aggregates(l), rp, cols, vals, ptr, idx))
        return out

    @property
    def level_num(self) -> int:
        n = C.c_int()
        _check(lib().lgpu_amg_level_num(self.h, C.byref(n)))
        return n.value

    def level_size(self, level: int) -> int:
        n = C.c_int64()
        _check(lib().lgpu_amg_level_size(self.h, level, C.byref(n)))
        return n.value

    def level_nnz(self, level: int) -> int:
        """Stored entries of level >= 1's matrix."""
        nnz = C.c_int64()
        _check(lib().lgpu_amg_matrix(self.h, level, C.byref(nnz), None, None, None))
        return nnz.value

    def aggregates(self, level: int) -> np.ndarray:
        out = np.empty(self.level_size(level), np.uint32)
        _check(lib().lgpu_amg_aggregates(self.h, level, out.ctypes.data_as(_u32p)))
        return out

    def matrix(self, level: int):
        nnz = C.c_int64()
        _check(lib().lgpu_amg_matrix(self.h, level, C.byref(nnz), None, None, None))
        n = self.level_size(level)
        rp = np.empty(n + 1, np.int64)
        cols = np.empty(max(nnz.value, 1), np.int32)
        vals = np.empty(max(nnz.value, 1), np.float64)
        _check(lib().lgpu_amg_matrix(self.h, level, C.byref(nnz), rp.ctypes.data_as(_i64p),
                                     cols.ctypes.data_as(_i32p), vals.ctypes.data_as(_f64p)))
        return rp, cols[: nnz.value], vals[: nnz.value]

    @property
    def setup_seconds(self) -> float:
        s = C.c_double()
        _check(lib().lgpu_amg_setup_seconds(self.h, C.byref(s)))
        return s.value

    @property
    def setup_split(self) -> dict:
        """Host setup seconds by phase."""
        out = np.zeros(5)
        _check(lib().lgpu_amg_setup_split(self.h, out.ctypes.data_as(_f64p)))
        return dict(zip(("criterion", "aggregation", "galerkin", "coarsest", "rest"), out.tolist()))

    def apply(self, r: DenseVector, z: DenseVector):
        """z = M r: one cycle from a zero initial guess (the preconditioner)."""
        _check(lib().lgpu_amg_vcycle(self.h, r.h, z.h))
        return z

    def apply_ex_situ(self, sol_in: DenseVector | None, sol_out: DenseVector, rhs: DenseVector):
        """MultiGridSolver::Instance::apply(in_sol, out_sol, rhs): one cycle from
        sol_in (None: zero) into sol_out (multigrid/solver.hpp:47-53)."""
        _check(lib().lgpu_amg_cycle(self.h, _vh(sol_in), rhs.h, sol_out.h))
        return sol_out

    def apply_in_situ(self, sol: DenseVector, rhs: DenseVector):
        """MultiGridSolver::Instance::apply(sol, rhs): one cycle from sol
        (multigrid/solver.hpp:41-46)."""
        aux = getattr(self, "_aux", None)
        if aux is None or aux.dtype != sol.dtype:
            aux = self._aux = DenseVector(sol.n, sol.dtype, sol.ctx)
        aux.assign(sol)
        return self.apply_ex_situ(aux, sol, rhs)

    def __del__(self):
        try:
            if getattr(self, "h", None) and not _finalizing:
                lib().lgpu_amg_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Krylov solvers


@dataclass
class SolveResult:
    status: int
    iterations: int
    residuals: np.ndarray

    @property
    def converged(self):
        return self.status == OK


class KrylovSystemInstance:
    """SystemInstance of ConjugateGradientsSolver / BiCgStabSolver
    (conjugate-gradients/solver.hpp:38-198, bicgstab/solver.hpp:46-170):
    constructed from (sol, rhs), then iterate(), residual_norm(), rho(),
    solution().  Each instance owns its device work vectors and scalars."""

    def __init__(self, create_fn, lhs, precon, sol: DenseVector, rhs: DenseVector, rere_every: int):
        self.sol, self.rhs, self.lhs, self.precon = sol, rhs, lhs, precon
        h = _vp()
        _check(getattr(lib(), create_fn)(lhs.h, _vh(precon), rhs.h, sol.h, int(rere_every), C.byref(h)))
        self.h = h

    def iterate(self):
        _check(lib().lgpu_krylov_iterate(self.h))

    def residual_norm(self) -> float:
        out = C.c_double()
        _check(lib().lgpu_krylov_residual_norm(self.h, C.byref(out)))
        return out.value

    def rho(self) -> float:
        out = C.c_double()
        _check(lib().lgpu_krylov_rho(self.h, C.byref(out)))
        return out.value

    @property
    def iteration(self) -> int:
        out = C.c_int()
        _check(lib().lgpu_krylov_iteration(self.h, C.byref(out)))
        return out.value

    def solution(self) -> DenseVector:
        return self.sol

    def __del__(self):
        try:
            if getattr(self, "h", None) and not _finalizing:
                lib().lgpu_krylov_destroy(self.h)
                self.h = None
        except Exception:
            pass


class _Krylov:
    _fn = ""
    _create = ""

    def __init__(self, rere_every: int = 8, precon: MultiGridInstance | None = None):
        self.rere_every = rere_every
        self.precon = precon

    def solve(self, lhs: _Matrix, sol: DenseVector, rhs: DenseVector, rel_tol=1e-8, max_it=1000) -> SolveResult:
        """The fused loop: iterations until ||r_k|| <= rel_tol ||r_0|| or max_it."""
        p = _KrylovParams(float(rel_tol), int(max_it), int(self.rere_every))
        it = C.c_int()
        hist = np.zeros(max_it + 1)
        st = getattr(lib(), self._fn)(lhs.h, _vh(self.precon), rhs.h, sol.h, C.byref(p), C.byref(it),
                                      hist.ctypes.data_as(_f64p))
        _check(st, allow=(NOT_CONVERGED, BREAKDOWN, NONFINITE))
        return SolveResult(st, it.value, hist[: it.value + 1])

    def instantiate(self, lhs: _Matrix) -> "_KrylovInstance":
        return _KrylovInstance(self, lhs)


class _KrylovInstance:
    def __init__(self, solver: _Krylov, lhs: _Matrix):
        self.solver, self.lhs = solver, lhs

    def instantiate(self, sol: DenseVector, rhs: DenseVector) -> KrylovSystemInstance:
        s = self.solver
        return KrylovSystemInstance(s._create, self.lhs, s.precon, sol, rhs, s.rere_every)


class ConjugateGradientsSolver(_Krylov):
    """ConjugateGradientsSolver (conjugate-gradients/solver.hpp:25-221)."""

    _fn = "lgpu_cg_solve"
    _create = "lgpu_cg_create"


class BiCgStabSolver(_Krylov):
    """BiCgStabSolver (bicgstab/solver.hpp:26-259)."""

    _fn = "lgpu_bicgstab_solve"
    _create = "lgpu_bicgstab_create"


def run_iterative_solver(solver, lhs, sol, rhs, iter_op=None, rel_tol=1e-8, max_it=1000):
    """run_iterative_solver(solver, lhs, sol, rhs, aux, env, iter_op)
    (linear-solver/runner.hpp:17-84).

    With iter_op (called as iter_op(iteration, instance), returning None to
    continue or the result to stop): a Krylov solver's system instance is
    built from (sol, rhs) and iter_op sees it at iteration 0 and after every
    iterate(); a stationary solver (JacobiSolver, SorSolver, SsorSolver,
    MultiGridSolver) is applied in-situ to sol and iter_op sees its instance
    after every apply, from iteration 1.  Without iter_op a Krylov solver runs
    the fused loop with the relative-residual stop rule and returns a
    SolveResult."""
    if iter_op is None:
        return solver.solve(lhs, sol, rhs, rel_tol, max_it)
    inst = solver.instantiate(lhs)
    if isinstance(solver, _Krylov):
        system = inst.instantiate(sol, rhs)
        res = iter_op(0, system)
        i = 1
        while res is None:
            system.iterate()
            res = iter_op(i, system)
            i += 1
        return res
    i = 1
    while True:
        inst.apply_in_situ(sol, rhs)
        res = iter_op(i, inst)
        if res is not None:
            return res
        i += 1


# ---------------------------------------------------------------------------
# adjoint of the solve (not in the reference)


class SmfReader:
    """SmfReader (io/tensor.hpp:26-156): the header on construction, the CSR
    matrix on read().  The reference builds it as nonsymmetric; symmetric=True
    declares symmetry."""

    def __init__(self, path):
        self.path = os.fspath(path)
        r, c, z = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().lgpu_smf_info(self.path.encode(), C.byref(r), C.byref(c), C.byref(z)))
        self._dims = (r.value, c.value, z.value)

    def row_num(self) -> int:
        return self._dims[0]

    def col_num(self) -> int:
        return self._dims[1]

    def nnz_num(self) -> int:
        return self._dims[2]

    def read(self, dtype=np.float64, symmetric=False, ctx: Context | None = None) -> "CsrMatrix":
        ctx = ctx or default_context()
        h = _vp()
        _check(lib().lgpu_smf_read(ctx.h, self.path.encode(), _dtype_code(dtype), int(symmetric), C.byref(h)))
        return CsrMatrix(h, ctx)


class SmfWriter:
    """SmfWriter (io/tensor.hpp:158-213) of a CSR or stencil matrix."""

    def __init__(self, path):
        self.path = os.fspath(path)

    def write(self, matrix: _Matrix):
        _check(lib().lgpu_smf_write(matrix.h, self.path.encode()))


class DvfReader:
    """DvfReader (io/tensor.hpp:215-251)."""

    def __init__(self, path):
        self.path = os.fspath(path)
        n = C.c_int64()
        _check(lib().lgpu_dvf_size(self.path.encode(), C.byref(n)))
        self._size = n.value

    def size(self) -> int:
        return self._size

    def read(self, dtype=np.float64, ctx: Context | None = None) -> DenseVector:
        ctx = ctx or default_context()
        h = _vp()
        _check(lib().lgpu_dvf_read(ctx.h, self.path.encode(), _dtype_code(dtype), C.byref(h)))
        return DenseVector._adopt(h, self._size, dtype, ctx)


class DvfWriter:
    """DvfWriter (io/tensor.hpp:253-263): values written as f64."""

    def __init__(self, path):
        self.path = os.fspath(path)

    def write(self, vec: DenseVector):
        _check(lib().lgpu_dvf_write(vec.h, self.path.encode()))


class EffDiffCalc:
    """Effective diffusivity through the mid plane x = nx // 2
    (component-wise/sink/effective-diffusion.hpp:229-327).  The area part is
    computed once, at construction (:292-295); none_mats flags the "none"
    materials of a lookup operator."""

    def __init__(self, valuator: StencilMatrix, none_mats=None):
        flags = None
        if none_mats is not None:
            self._none = np.ascontiguousarray([1 if f else 0 for f in none_mats], np.uint8)
            assert self._none.size == 4
            flags = self._none.ctypes.data_as(_u8p)
        out = C.c_double()
        _check(lib().lgpu_eff_diff_area_part(valuator.h, flags, C.byref(out)))
        self.area_part = out.value

    def __call__(self, valuator: StencilMatrix, sol: DenseVector, system: CsrSystem | None = None) -> float:
        """sol is indexed by cells, or by the rows of a compressed system."""
        m = None if system is None or system.compress_map is None else system.lhs.h
        out = C.c_double()
        _check(lib().lgpu_eff_diff(valuator.h, m, sol.h, self.area_part, C.byref(out)))
        return out.value


def outlet_flux(A: StencilMatrix, p: DenseVector) -> float:
    """J = sum over the x = nx-1 cells of (2 q_0 c_i)(p_i - p_end)."""
    out = C.c_double()
    _check(lib().lgpu_outlet_flux(A.h, p.h, C.byref(out)))
    return out.value


def outlet_flux_adjoint_rhs(A: StencilMatrix, g: DenseVector) -> DenseVector:
    _check(lib().lgpu_outlet_flux_adjoint_rhs(A.h, g.h))
    return g


def outlet_flux_gradient(A: StencilMatrix, p: DenseVector, lam: DenseVector,
                         grad: DenseVector) -> DenseVector:
    """grad = dJ/dlog c from the forward solution p and the adjoint solution lam."""
    _check(lib().lgpu_outlet_flux_gradient(A.h, p.h, lam.h, grad.h))
    return grad


@dataclass
class AdjointResult:
    status: int
    flux: float
    forward_iterations: int
    adjoint_iterations: int

    @property
    def converged(self):
        return self.status == OK


def adjoint_grad(A: StencilMatrix, precon: MultiGridInstance | None, p: DenseVector,
                 lam: DenseVector, grad: DenseVector, rel_tol=1e-10, max_it=1000,
                 rere_every=8) -> AdjointResult:
    """Forward CG solve A p = b, adjoint CG solve A lam = dJ/dp with the same
    preconditioner, and the fused gradient pass (lgpu_adjoint_grad)."""
    prm = _KrylovParams(float(rel_tol), int(max_it), int(rere_every))
    flux = C.c_double()
    its = (C.c_int * 2)()
    st = lib().lgpu_adjoint_grad(A.h, _vh(precon), C.byref(prm), p.h, lam.h, grad.h,
                                 C.byref(flux), its)
    _check(st, allow=(NOT_CONVERGED, BREAKDOWN, NONFINITE))
    return AdjointResult(st, flux.value, its[0], its[1])
