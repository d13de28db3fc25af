"""GPU parity of the outlet-flux adjoint: the fused gradient pass against the
oracle, the forward + adjoint solves against the oracle's, a closed form, and
central finite differences of the flux (h = 1e-6 in log c)."""

import numpy as np
import pytest

import lineal_amd as la
import oracle as O
from lineal_amd import problems

pytestmark = pytest.mark.gpu


def fd_gradient(box, c, p, cells, M, h=1e-6, rel_tol=1e-12):
    """Central differences (J(theta + h e_k) - J(theta - h e_k)) / 2h.

    J(theta') - J(theta) is evaluated without cancellation: the correction d
    solves A' d = b' - A' p for the converged p, and
    J(theta') - J(theta) = sum_out 2 q_0 c'_i d_i + [k on the outlet] 2 q_0 (c'_k - c_k)(p_k - p_end).
    Only the rows next to k differ between the two right-hand sides, so the
    forward solve's own residual cancels between the + and - evaluations."""
    N = c.size
    nyz = box.dims[1] * box.dims[2]
    q0 = box.lengths[1] / box.dims[1] * (box.lengths[2] / box.dims[2]) / (box.lengths[0] / box.dims[0]) \
        if box.lengths is not None else 1.0
    pv = la.DenseVector(p)
    out = np.zeros(len(cells))
    for n_, k in enumerate(cells):
        dj = []
        for s in (1.0, -1.0):
            cs = c.copy()
            cs[k] = c[k] * np.exp(s * h)
            As = la.StencilMatrix.values(box, cs)
            r = As.residual(As.rhs(), pv, la.DenseVector(N))
            d = la.DenseVector(N, fill=0.0)
            res = la.ConjugateGradientsSolver(8, M).solve(As, d, r, rel_tol=rel_tol, max_it=500)
            assert res.converged, res
            dv = d.numpy()
            v = np.sum((2 * q0 * cs[-nyz:]) * dv[-nyz:])
            if k >= N - nyz:
                v += 2 * q0 * (cs[k] - c[k]) * (p[k] - box.p_end)
            dj.append(v)
        out[n_] = (dj[0] - dj[1]) / (2 * h)
    return out


SHAPES = [(5, 7, 9), (17, 3, 130), (33, 20, 16), (40, 1, 1)]


@pytest.mark.parametrize("dims", SHAPES)
def test_gradient_pass_equals_oracle(dims):
    lengths = (2.0, 3.5, 1.25)
    c = problems.lognormal(dims, 1.0, 7)
    box = la.BoxInfo(dims, lengths=lengths, p_start=1.5, p_end=0.25)
    A = la.StencilMatrix.values(box, c)
    S = O.Stencil(dims, coeff=c, lengths=lengths, p_start=1.5, p_end=0.25)
    p = problems.uniform(S.N, 8)
    lam = problems.uniform(S.N, 9)
    g = la.outlet_flux_gradient(A, la.DenseVector(p), la.DenseVector(lam), la.DenseVector(S.N)).numpy()
    np.testing.assert_array_equal(g, O.outlet_flux_gradient(S, p, lam))
    rhs = la.outlet_flux_adjoint_rhs(A, la.DenseVector(S.N)).numpy()
    np.testing.assert_array_equal(rhs, O.outlet_flux_adjoint_rhs(S))
    # the flux is a reduction over the outlet plane in another order
    assert la.outlet_flux(A, la.DenseVector(p)) == pytest.approx(O.outlet_flux(S, p), rel=1e-13, abs=1e-300)


def test_adjoint_rejects_unsupported_operators():
    dims = (6, 5, 4)
    m = problems.spheres(dims, 3, rmin=1, rmax=2)
    A = la.StencilMatrix.lookup(la.BoxInfo(dims), m, (1.0, 1e-6, 0.0, 0.0))
    with pytest.raises(la.LinealError):
        la.outlet_flux(A, la.DenseVector(A.row_num, fill=0.0))
    c = problems.lognormal(dims, 1.0, 1)
    Aa = la.StencilMatrix.values(la.BoxInfo(dims, velocity_x=0.5), c)
    with pytest.raises(la.LinealError):
        la.outlet_flux_adjoint_rhs(Aa, la.DenseVector(Aa.row_num))


@pytest.mark.parametrize("dims", [(24, 20, 16), (32, 32, 32)])
def test_adjoint_solve_matches_oracle(dims):
    c = problems.lognormal(dims, 1.0, 5)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    S = O.Stencil(dims, coeff=c)
    M = la.MultiGridSolver(precision=np.float64).instantiate(A)
    H = O.Hierarchy(S, coarse_segment=64)
    N = S.N
    p, lam, g = la.DenseVector(N, fill=0.0), la.DenseVector(N, fill=0.0), la.DenseVector(N)
    res = la.adjoint_grad(A, M, p, lam, g, rel_tol=1e-12)
    assert res.converged
    J, g_ref, p_ref, lam_ref, itf, ita = O.adjoint_outlet_flux(S, H, rel_tol=1e-12)
    assert abs(res.forward_iterations - itf) <= 1 and abs(res.adjoint_iterations - ita) <= 1
    assert np.abs(p.numpy() - p_ref).max() <= 1e-10
    assert res.flux == pytest.approx(J, rel=1e-10)
    gg = g.numpy()
    assert np.abs(gg - g_ref).max() <= 1e-8 * np.abs(g_ref).max()


def test_adjoint_closed_form_columns():
    """A field constant across y and z: J = n_y n_z q_0 dp / R, R = sum 1/c_i,
    and dJ/dlog c_k = J_col^2 / (q_0 dp c_k)."""
    dims = (48, 6, 5)
    col = problems.lognormal((dims[0], 1, 1), 1.0, 11)
    c = np.repeat(col, dims[1] * dims[2])
    A = la.StencilMatrix.values(la.BoxInfo(dims), c)
    M = la.MultiGridSolver().instantiate(A)
    N = c.size
    p, lam, g = la.DenseVector(N, fill=0.0), la.DenseVector(N, fill=0.0), la.DenseVector(N)
    res = la.adjoint_grad(A, M, p, lam, g, rel_tol=1e-13)
    assert res.converged
    j_col = 1.0 / np.sum(1.0 / col)
    assert res.flux == pytest.approx(dims[1] * dims[2] * j_col, rel=1e-11)
    np.testing.assert_allclose(g.numpy(), np.repeat(j_col ** 2 / col, dims[1] * dims[2]), rtol=1e-9)


@pytest.mark.parametrize("dims,lengths", [((32, 24, 20), None), ((20, 16, 24), (1.0, 2.0, 0.5))])
def test_adjoint_matches_central_differences(dims, lengths):
    c = problems.lognormal(dims, 1.0, 13)
    box = la.BoxInfo(dims, lengths=lengths)
    A = la.StencilMatrix.values(box, c)
    M = la.MultiGridSolver().instantiate(A)
    N = c.size
    p, lam, g = la.DenseVector(N, fill=0.0), la.DenseVector(N, fill=0.0), la.DenseVector(N)
    assert la.adjoint_grad(A, M, p, lam, g, rel_tol=1e-12).converged
    rng = np.random.default_rng(5)
    plane = dims[1] * dims[2]
    cells = np.concatenate([rng.choice(N, 12, replace=False), [0, N - 1, plane + 3, N - plane + 1]])
    fd = fd_gradient(box, c, p.numpy(), cells, M)
    ga = g.numpy()[cells]
    assert np.abs(ga - fd).max() <= 1e-6 * np.abs(fd).max()
    assert np.all(np.abs(ga - fd) <= 1e-5 * np.abs(fd))


def test_autograd_function_returns_the_adjoint_gradient():
    torch = pytest.importorskip("torch")
    from lineal_amd.autograd import outlet_flux

    dims = (20, 12, 10)
    c = problems.lognormal(dims, 1.0, 17)
    box = la.BoxInfo(dims)
    log_c = torch.tensor(np.log(c).reshape(dims), dtype=torch.float64, requires_grad=True)
    J = outlet_flux(log_c, box, rel_tol=1e-12)
    (3.0 * J).backward()
    # the wrapper sees c = exp(log c), which need not round-trip bit for bit
    A = la.StencilMatrix.values(box, np.exp(log_c.detach().numpy().reshape(-1)))
    M = la.MultiGridSolver().instantiate(A)
    N = c.size
    p, lam, g = la.DenseVector(N, fill=0.0), la.DenseVector(N, fill=0.0), la.DenseVector(N)
    res = la.adjoint_grad(A, M, p, lam, g, rel_tol=1e-12)
    assert float(J.detach()) == res.flux
    np.testing.assert_array_equal(log_c.grad.numpy().reshape(-1), 3.0 * g.numpy())


@pytest.mark.timeout(600)
def test_adjoint_outlet_flux_128_cubed():
    """Outlet-flux gradient w.r.t. 128^3 per-cell log coefficients (lognormal
    sigma = 1), FP64 forward + adjoint AMG-CG at rel. tol 1e-10: central
    differences (h = 1e-6) at 64 seeded cells and the oracle's FP64 gradient
    from its own solves."""
    import concurrent.futures as cf

    dims = (128, 128, 128)
    c = problems.lognormal(dims, 1.0, 1)
    box = la.BoxInfo(dims)
    A = la.StencilMatrix.values(box, c)
    M = la.MultiGridSolver(precision=np.float64).instantiate(A)
    N = c.size
    p, lam, g = la.DenseVector(N, fill=0.0), la.DenseVector(N, fill=0.0), la.DenseVector(N)
    res = la.adjoint_grad(A, M, p, lam, g, rel_tol=1e-10)
    assert res.converged
    gg, pp, ll = g.numpy(), p.numpy(), lam.numpy()
    S = O.Stencil(dims, coeff=c)
    # the fused pass equals the oracle's formula on the same fields
    np.testing.assert_array_equal(gg, O.outlet_flux_gradient(S, pp, ll))
    # central differences at 64 seeded cells
    cells = np.random.default_rng(5).choice(N, 64, replace=False)
    fd = fd_gradient(box, c, pp, cells, M)
    err_fd = np.abs(gg[cells] - fd).max() / np.abs(fd).max()
    print(f"\n128^3 adjoint: J = {res.flux:.12g}, CG iterations {res.forward_iterations} + "
          f"{res.adjoint_iterations}; max |g - fd| / max |fd| over 64 cells = {err_fd:.3e}", flush=True)
    assert err_fd <= 1e-5
    # the oracle's gradient from its own FP64 solves (forward and adjoint in parallel)
    H = O.Hierarchy(S, coarse_segment=64)
    with cf.ThreadPoolExecutor(2) as ex:
        fwd = ex.submit(O.pcg, S, S.rhs(), H, 1, False, 1e-10, 1000)
        adj = ex.submit(O.pcg, S, O.outlet_flux_adjoint_rhs(S), H, 1, False, 1e-10, 1000)
        p_ref, lam_ref = fwd.result()[0], adj.result()[0]
    g_ref = O.outlet_flux_gradient(S, p_ref, lam_ref)
    err_ref = np.abs(gg - g_ref).max() / np.abs(g_ref).max()
    err_j = abs(res.flux - O.outlet_flux(S, p_ref)) / abs(res.flux)
    print(f"max |g - g_oracle| / max |g_oracle| = {err_ref:.3e}, flux rel. diff {err_j:.3e}", flush=True)
    assert err_ref <= 1e-5 and err_j <= 1e-9
