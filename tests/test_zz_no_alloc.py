"""lgpu_op_ssor_noalloc (the SSOR step whose intermediate vector lives in a
static per-GPU scratch) against lgpu_op_ssor through the C ABI: bit for bit,
on two calls in a row, the second on a smaller operator; and the rejection of
operators larger than the scratch."""

import ctypes as C

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu


def _ssor(fn, A, x_in, b, dtype, omega, ultra):
    out = la.DenseVector(A.row_num, dtype)
    st = getattr(la.lib(), fn)(A.h, None if x_in is None else x_in.h, out.h, b.h, C.c_double(omega),
                               int(ultra))
    la._check(st)
    return out.numpy()


@pytest.mark.parametrize("kind,dtype", [("stencil", np.float64), ("stencil", np.float32),
                                        ("csr", np.float64), ("lookup", np.float32)])
def test_noalloc_ssor_equals_ssor(kind, dtype):
    for call, dims in enumerate(((20, 18, 17), (9, 7, 6))):
        c = problems.lognormal(dims, 1.0, 11 + call)
        box = la.BoxInfo(dims)
        if kind == "lookup":
            m = problems.spheres(dims, 3 + call, rmin=2, rmax=3)
            A = la.StencilMatrix.lookup(box, m, (1.0, 1e-6, 0.0, 0.0))
        else:
            A = la.StencilMatrix.values(box, c)
            if kind == "csr":
                A = A.to_csr(dtype)
        n = A.row_num
        b = la.DenseVector(problems.uniform(n, 21 + call).astype(dtype))
        x = la.DenseVector(problems.uniform(n, 31 + call).astype(dtype))
        if dtype == np.float32 and kind == "stencil":
            A.spmv(x, la.DenseVector(n, dtype))  # the first f32 pass makes the f32 copy
        for x_in in (None, x):
            for ultra in (False, True):
                ref = _ssor("lgpu_op_ssor", A, x_in, b, dtype, 1.0, ultra)
                got = _ssor("lgpu_op_ssor_noalloc", A, x_in, b, dtype, 1.0, ultra)
                np.testing.assert_array_equal(got, ref)
                assert np.all(np.isfinite(got))
        if kind == "stencil" and dtype == np.float64:
            S = O.Stencil(dims, coeff=c)
            np.testing.assert_array_equal(_ssor("lgpu_op_ssor_noalloc", A, x, b, dtype, 1.0, False),
                                          O.ssor(S, x.numpy(), b.numpy(), 1.0, False, 1))


def test_noalloc_ssor_rejects_operators_larger_than_the_scratch():
    dims = (256, 128, 129)  # 4227072 rows > 2^22
    A = la.StencilMatrix.values(la.BoxInfo(dims), np.ones(int(np.prod(dims))))
    n = A.row_num
    b, out = la.DenseVector(n, fill=0.0), la.DenseVector(n)
    st = la.lib().lgpu_op_ssor_noalloc(A.h, None, out.h, b.h, C.c_double(1.0), 0)
    assert st == la.INVALID_ARG
    assert b"static scratch" in la.lib().lgpu_last_error()
