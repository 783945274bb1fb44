"""Config 5 at its own size: 32 right-hand sides (N(0,1), seeds 0-31) on the
128^3 convection-diffusion operator (n = 2,097,152), each with the quadratic
loss 0.5 ||x_c - x_t,c||^2 (x_t,c ~ N(0,1), seeds 1000-1031).

(a) The bench's configuration -- block Householder FGMRES(30) in FP64 with
    10 batched FP32 GMRES steps per column as the preconditioner, FP32 ILU
    (3 sweeps), Jacobi-5 -- solves the 32 forward systems and the 32
    adjoint systems A^T lambda_c = x_c - x_t,c to an FP64 true relative
    residual <= 1e-8 per column, recomputed on the host.
(b) The gradient of the summed loss with respect to A's stored values,
    dL/dtheta_k = -sum_c lambda_c[row(k)] x_c[col(k)], matches central
    differences (h = 1e-6) on two seeded directions to 1e-6 relative.  The
    gradient comes from forward and adjoint solves to 1e-11; the perturbed
    solves start from that solution, so their 1e-8 tolerance applies to the
    O(h) correction.  The gradient of the 1e-8 solves of (a) is compared
    with it as well.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 128
NRHS = 32


def _solver(gpu, A, rtol, restart_max=20):
    o = gpu.default_opts()
    o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = 30, restart_max, rtol, 3
    o.trsv, o.jacobi_sweeps, o.prec_precond, o.inner_iters = gpu.TRSV_JACOBI, 5, gpu.LACE_F32, 10
    S = gpu.Solver(gpu.HostCSR.from_csr(A), o)
    S.setup()
    return S


@pytest.fixture(scope="module")
def c5(gpu, orc):
    row, col, val = gpu.gen_convdiff3d(N, N, N)
    A = orc.CSR(len(row) - 1, len(row) - 1, row, col, val)
    n = A.n_rows
    B = np.stack([gpu.gen_normal(c, n) for c in range(NRHS)], axis=1)
    Xt = np.stack([gpu.gen_normal(1000 + c, n) for c in range(NRHS)], axis=1)
    state = {"A": A, "B": B, "Xt": Xt, "As": A.scipy()}
    yield state
    state.clear()


def _colrel(As, B, X, trans=False):
    R = B - (As.T @ X if trans else As @ X)
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


def _grad(gpu, S, Lam, X, nnz):
    n = X.shape[0]
    dL = gpu.DevVec(n * NRHS, np.float64, np.ascontiguousarray(Lam).reshape(-1))
    dX = gpu.DevVec(n * NRHS, np.float64, np.ascontiguousarray(X).reshape(-1))
    return S.grad_values_block(dL, dX, nnz)


def test_config5_forward_adjoint_bench_config(gpu, c5):
    A, B, Xt, As = c5["A"], c5["B"], c5["Xt"], c5["As"]
    S = _solver(gpu, A, 1e-8)
    X, stf = S.solve_block(B)
    assert stf.converged == 1
    rel = _colrel(As, B, X)
    assert rel.max() <= 1e-8, rel.max()
    assert np.allclose(rel, np.array(stf.rel_true_residual), rtol=1e-3, atol=1e-16)
    G = X - Xt
    Lam, sta = S.solve_block(G, trans=1)
    assert sta.converged == 1
    rel_a = _colrel(As, G, Lam, trans=True)
    assert rel_a.max() <= 1e-8, rel_a.max()
    c5["g_1e8"] = _grad(gpu, S, Lam, X, A.nnz)
    c5["steps"] = (stf.iterations, sta.iterations)


def test_config5_gradient_vs_central_differences(gpu, c5):
    A, B, Xt, As = c5["A"], c5["B"], c5["Xt"], c5["As"]
    T = _solver(gpu, A, 1e-11, restart_max=40)
    X, stf = T.solve_block(B)
    assert stf.converged == 1
    assert _colrel(As, B, X).max() <= 1e-11
    Lam, sta = T.solve_block(X - Xt, trans=1)
    assert sta.converged == 1
    g = _grad(gpu, T, Lam, X, A.nnz)
    del T
    S = _solver(gpu, A, 1e-8)
    h = 1e-6
    for seed in (100, 101):
        d = gpu.gen_normal(seed, A.nnz)

        def solve(vals):
            S.set_values(vals)
            Xv, st = S.solve_block(B, X0=X)
            assert st.converged == 1
            return Xv

        Xp, Xm = solve(A.val + h * d), solve(A.val - h * d)
        # (L(+) - L(-)) / 2h through |a|^2 - |b|^2 = (a - b).(a + b)
        fd = 0.5 * float(np.sum((Xp - Xm) * (Xp + Xm - 2.0 * Xt))) / (2.0 * h)
        an = float(g @ d)
        assert abs(fd - an) <= 1e-6 * abs(an), (seed, fd, an)
        if "g_1e8" in c5:
            # the 1e-8 solves of the bench configuration give the same
            # directional derivative to within their own tolerance
            assert abs(float(c5["g_1e8"] @ d) - an) <= 1e-4 * abs(an)
    S.set_values(A.val)
