"""Adjoint of the preconditioned solve (config 4): A(theta) x = b with theta
the CSR values, L = 0.5 ||x - x_t||^2.  The transposed solve uses the ILU
factors of A transposed (U^T then L^T sweeps) inside the same FGMRES; the
gradient gather forms dL/dtheta_ij = -lambda_i x_j.  Checked against FP64
central differences (h = 1e-6) on seeded directions and against an
independent oracle solve with the explicitly transposed matrix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(gpu, A, rtol=1e-13, **kw):
    o = gpu.default_opts()
    o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = 50, 200, rtol, 3
    o.trsv, o.prec_precond = gpu.TRSV_LEVEL, gpu.LACE_F64
    for k, v in kw.items():
        setattr(o, k, v)
    S = gpu.Solver(gpu.HostCSR.from_csr(A), o)
    S.setup()
    return S


def test_transposed_solve_vs_oracle_on_explicit_transpose(gpu, orc):
    A = orc.convdiff3d(14, 12, 10)
    g = gpu.gen_normal(7, A.n_rows)
    S = _solver(gpu, A, rtol=1e-12)
    lam, st = S.solve(g, trans=1)
    assert st.converged == 1
    AT = orc.transpose(A)
    Lf, Uf = orc.parilu(AT, 3).factors()
    ref = orc.fgmres_householder(AT, g, m=50, restart_max=50, rtol=1e-12, tol_type=1, choice=0,
                                 Lf=Lf, Uf=Uf)
    assert ref.info == 0
    assert np.allclose(lam, ref.x, rtol=1e-9, atol=1e-9 * np.abs(ref.x).max())
    assert np.linalg.norm(g - AT.scipy() @ lam) <= 1.01e-12 * np.linalg.norm(g)
    # the forward direction on the same handle still solves A x = g
    x, st2 = S.solve(g, trans=0)
    assert np.linalg.norm(g - A.scipy() @ x) <= 1.01e-12 * np.linalg.norm(g)


@pytest.mark.parametrize("mixed", [False, True])
def test_transposed_mixed_precision(gpu, orc, mixed):
    A = orc.convdiff3d(16, 16, 16)
    g = gpu.gen_normal(9, A.n_rows)
    kw = dict(trsv=gpu.TRSV_JACOBI, jacobi_sweeps=5, prec_precond=gpu.LACE_F32)
    if mixed:
        kw.update(inner_iters=8, search_max=20)
    S = _solver(gpu, A, rtol=1e-10, **kw)
    lam, st = S.solve(g, trans=1)
    assert st.converged == 1
    assert np.linalg.norm(g - A.scipy().T @ lam) <= 1.05e-10 * np.linalg.norm(g)


def _solve(S, A_vals, b, x0=None):
    S.set_values(A_vals)
    x, st = S.solve(b, x0=x0)
    assert st.converged == 1
    return x


def _central_difference(xp, xm, xt, h):
    """(L(+) - L(-)) / 2h for L = 0.5 ||x - x_t||^2, evaluated through the
    identity |a|^2 - |b|^2 = (a - b).(a + b): the same quantity without the
    cancellation of two O(n) losses."""
    return 0.5 * float(np.dot(xp - xm, xp + xm - 2.0 * xt)) / (2.0 * h)


@pytest.mark.parametrize("size", [12, 64])
def test_gradient_vs_central_differences(gpu, orc, size):
    """Config 4 at 64^3: x_t ~ N(0,1) seed 3, b ~ N(0,1) seed 4, directions
    seeds 100-119 (the 12^3 case is a quick check of the same code).  The
    perturbed solves start from the unperturbed solution x, so their
    relative tolerance applies to the O(h) correction: the difference
    quotient then carries no solve error of the size of x (rtol * |x| / h
    would be 1e-7 relative here, at the level of the 1e-6 bar)."""
    import lace_amd as L
    row, col, val = L.gen_convdiff3d(size, size, size)
    A = orc.CSR(len(row) - 1, len(row) - 1, row, col, val)
    n, nnz = A.n_rows, A.nnz
    xt = L.gen_normal(3, n)
    b = L.gen_normal(4, n)
    S = _solver(gpu, A)
    x = _solve(S, val, b)
    lam, st = S.solve(x - xt, trans=1)
    assert st.converged == 1
    dl = L.DevVec(n, np.float64, lam)
    dx = L.DevVec(n, np.float64, x)
    g = S.grad_values(dl, dx, nnz)
    # the gather itself against numpy: -lambda_i x_j on the pattern
    rows = np.repeat(np.arange(n), np.diff(row))
    assert np.array_equal(g, -lam[rows] * x[col])
    h = 1e-6
    ndir = 20 if size == 64 else 6
    for seed in range(100, 100 + ndir):
        d = L.gen_normal(seed, nnz)
        fd = _central_difference(_solve(S, val + h * d, b, x), _solve(S, val - h * d, b, x), xt, h)
        an = float(np.dot(g, d))
        assert abs(fd - an) <= 1e-6 * abs(an), (seed, fd, an)
    # dL/db = lambda: directional check in b
    S.set_values(val)
    for seed in (200, 201):
        e = L.gen_normal(seed, n)
        fd = _central_difference(_solve(S, val, b + h * e, x), _solve(S, val, b - h * e, x), xt, h)
        an = float(np.dot(lam, e))
        assert abs(fd - an) <= 1e-6 * abs(an)


def test_autograd_function(gpu, orc):
    torch = pytest.importorskip("torch")
    from lace_amd.autograd import sparse_solve
    import lace_amd as L
    A = orc.convdiff3d(10, 9, 8)
    S = _solver(gpu, A)
    theta = torch.tensor(A.val, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(L.gen_normal(4, A.n_rows), requires_grad=True)
    xt = torch.tensor(L.gen_normal(3, A.n_rows))
    x = sparse_solve(theta, b, S)
    loss = 0.5 * torch.sum((x - xt) ** 2)
    loss.backward()
    h = 1e-6
    for seed in (300, 301, 302):
        d = torch.tensor(L.gen_normal(seed, A.nnz))
        with torch.no_grad():
            xp = sparse_solve(theta + h * d, b, S).numpy()
            xm = sparse_solve(theta - h * d, b, S).numpy()
        fd = _central_difference(xp, xm, xt.numpy(), h)
        an = float(torch.dot(theta.grad, d))
        assert abs(fd - an) <= 1e-6 * abs(an)
    e = torch.tensor(L.gen_normal(303, A.n_rows))
    with torch.no_grad():
        xp = sparse_solve(theta, b + h * e, S).numpy()
        xm = sparse_solve(theta, b - h * e, S).numpy()
    fd = _central_difference(xp, xm, xt.numpy(), h)
    assert abs(fd - float(torch.dot(b.grad, e))) <= 1e-6 * abs(float(torch.dot(b.grad, e)))
