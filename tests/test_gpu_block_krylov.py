"""Multi-RHS (block) Krylov path on the GPU: the tall-skinny FP64 MFMA
products against numpy (this also pins the MFMA operand layout)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _blocks(gpu, arrs):
    devs = [gpu.DevVec(a.size, np.float64, a.reshape(-1)) for a in arrs]
    ptrs = (C.c_void_p * len(devs))(*[d.ptr.value for d in devs])
    return devs, ptrs


@pytest.mark.parametrize("n,nb", [(1000, 1), (4099, 5), (20000, 9)])
def test_tsgemm_tn(gpu, n, nb):
    rng = np.random.default_rng(n + nb)
    X = [rng.standard_normal((n, 32)) for _ in range(nb)]
    W = rng.standard_normal((n, 32))
    devs, ptrs = _blocks(gpu, X)
    dW = gpu.DevVec(n * 32, np.float64, W.reshape(-1))
    dC = gpu.DevVec(nb * 1024, np.float64)
    rs_main, rs_last = 37, 5
    gpu.check(gpu.lib().lace_tsgemm_tn(n, nb, ptrs, dW.ptr, rs_main, rs_last, dC.ptr))
    Cg = dC.download().reshape(nb, 32, 32)
    for b in range(nb):
        rs = rs_last if b == nb - 1 else rs_main
        want = X[b][rs:].T @ W[rs:]
        assert np.allclose(Cg[b], want, rtol=1e-12, atol=1e-10), b


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_tsgemm_nn(gpu, mode):
    n, nb, rlo, e0 = 5003, 4, 70, 96
    rng = np.random.default_rng(mode)
    X = [rng.standard_normal((n, 32)) for _ in range(nb)]
    D = rng.standard_normal((nb, 32, 32))
    base = rng.standard_normal((n, 32))
    devs, ptrs = _blocks(gpu, X)
    dD = gpu.DevVec(nb * 1024, np.float64, D.reshape(-1))
    dB = gpu.DevVec(n * 32, np.float64, base.reshape(-1))
    dO = gpu.DevVec(n * 32, np.float64)
    gpu.check(gpu.lib().lace_tsgemm_nn(n, nb, ptrs, dD.ptr, dB.ptr, mode, e0, rlo, dO.ptr))
    out = dO.download().reshape(n, 32)
    if mode == 0:
        b0 = np.zeros((n, 32))
    elif mode == 1:
        b0 = base.copy()
    else:
        b0 = np.zeros((n, 32)); b0[e0:e0 + 32] = np.eye(32)
    want = b0.copy()
    want[rlo:] -= sum(X[b][rlo:] @ D[b] for b in range(nb))
    assert np.allclose(out, want, rtol=1e-12, atol=1e-10)
    # in place on the base block
    gpu.check(gpu.lib().lace_tsgemm_nn(n, nb, ptrs, dD.ptr, dB.ptr, 1, 0, rlo, dB.ptr))
    if mode == 1:
        assert np.allclose(dB.download().reshape(n, 32), want, rtol=1e-12, atol=1e-10)


# ---------------------------------------------------------------- 32-RHS kernels
def _oracle_cols(orc, A, X):
    return np.stack([orc.spmv(A, X[:, c].astype(np.float64)) for c in range(X.shape[1])], axis=1)


@pytest.mark.parametrize("xdt", [np.float64, np.float32])
def test_spmm32_bitexact(gpu, orc, xdt):
    """Column c of the block product is the serial CSR sum of column c."""
    A = orc.convdiff3d(13, 11, 9)
    M = gpu.DevCSRMatrix(gpu.HostCSR.from_csr(A))
    rng = np.random.default_rng(11)
    X = rng.standard_normal((A.n_rows, 32)).astype(xdt)
    want = _oracle_cols(orc, A, X)
    assert np.array_equal(M.spmm32(X), want)
    B = rng.standard_normal((A.n_rows, 32))
    assert np.array_equal(M.spmm32(X, B), B - want)
    # and the single-vector device product of an FP64 column
    if xdt == np.float64:
        assert np.array_equal(M.spmv(X[:, 5]), want[:, 5])


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("trans", [0, 1])
def test_block_precond_bitexact(gpu, orc, prec, trans):
    """Block Jacobi ILU solves equal the single-vector solves per column."""
    A = orc.convdiff3d(12, 10, 9)
    lp = gpu.LACE_F64 if prec == "f64" else gpu.LACE_F32
    dt = np.float64 if prec == "f64" else np.float32
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), lp)
    P.sweeps(3)
    R = np.random.default_rng(5 + trans).standard_normal((A.n_rows, 32)).astype(dt)
    Z = P.apply_block(R, gpu.TRSV_JACOBI, sweeps=4, trans=trans)
    for c in (0, 13, 31):
        z = P.apply(R[:, c].copy(), gpu.TRSV_JACOBI, sweeps=4, trans=trans)
        assert np.array_equal(Z[:, c], z), c
    if trans == 0:
        # against the oracle's Jacobi sweeps on the oracle's factors
        st = orc.parilu(A, 3, dtype=dt)
        Lf = orc.CSR(A.n_rows, A.n_cols, st.L.row, st.L.col, st.Lval)
        Ur = orc.transpose(orc.CSR(A.n_rows, A.n_cols, st.U.row, st.U.col, st.Uval))
        y = orc.trsv_jacobi(Lf, R[:, 7], upper=False, sweeps=4, dtype=dt)
        z = orc.trsv_jacobi(Ur, y, upper=True, sweeps=4, dtype=dt)
        assert np.array_equal(Z[:, 7], z)


def _variants(gpu, fn, form_opt):
    """fn() on the diagonal form (form_opt 1) and on the ELLPACK form
    (form_opt 0), each with the plane-segment block order on and off."""
    lib = gpu.lib()
    outs = []
    try:
        for form in (1, 0):
            gpu.check(lib.lace_set_option(form_opt, form))
            for order in (1, 0):
                gpu.check(lib.lace_set_option(b"mrhs_plane_order", order))
                outs.append(fn())
    finally:
        gpu.check(lib.lace_set_option(form_opt, 1))
        gpu.check(lib.lace_set_option(b"mrhs_plane_order", 1))
    return outs


@pytest.mark.parametrize("xdt", [np.float64, np.float32])
def test_spmm32_plane_order(gpu, orc, xdt):
    """On a 96 x 96 x 4 grid (planes of 9216 rows: more than 1 MB of
    right-hand rows in either precision, and a short last segment) the block
    product runs plane segment by plane segment: the same bits as in natural
    row order and on the ELLPACK form, and column c is still the serial CSR
    sum of column c."""
    A = orc.convdiff3d(96, 96, 4)
    M = gpu.DevCSRMatrix(gpu.HostCSR.from_csr(A))
    assert M.forms()[0] == 7
    rng = np.random.default_rng(23)
    X = rng.standard_normal((A.n_rows, 32)).astype(xdt)
    B = rng.standard_normal((A.n_rows, 32))
    outs = _variants(gpu, lambda: (M.spmm32(X), M.spmm32(X, B)), b"spmv_dia")
    y1, r1 = outs[0]
    for y0, r0 in outs[1:]:
        assert np.array_equal(y1, y0) and np.array_equal(r1, r0)
    for c in (0, 17, 31):
        want = orc.spmv(A, X[:, c].astype(np.float64))
        assert np.array_equal(y1[:, c], want), c
        assert np.array_equal(r1[:, c], B[:, c] - want), c


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("trans", [0, 1])
def test_block_precond_plane_order(gpu, orc, prec, trans):
    """The 32-RHS Jacobi sweeps on the 96 x 96 x 4 factors in plane-segment
    order: the same bits as natural order, as the ELLPACK form and as the
    single-vector solves."""
    A = orc.convdiff3d(96, 96, 4)
    lp = gpu.LACE_F64 if prec == "f64" else gpu.LACE_F32
    dt = np.float64 if prec == "f64" else np.float32
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), lp)
    P.sweeps(3)
    R = np.random.default_rng(29 + trans).standard_normal((A.n_rows, 32)).astype(dt)
    outs = _variants(gpu, lambda: P.apply_block(R, gpu.TRSV_JACOBI, sweeps=4, trans=trans), b"jacobi_dia")
    Z1 = outs[0]
    for Z0 in outs[1:]:
        assert np.array_equal(Z1, Z0)
    z = P.apply(R[:, 9].copy(), gpu.TRSV_JACOBI, sweeps=4, trans=trans)
    assert np.array_equal(Z1[:, 9], z)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_block_precond_sweep_pairs(gpu, orc, prec):
    """Two Jacobi sweeps per launch (each first-sweep row formed by the rows
    that read it) give the bits of one sweep per launch, for even and odd
    sweep counts, both factors and both directions, and match the
    single-vector solves."""
    A = orc.convdiff3d(96, 96, 4)
    lp = gpu.LACE_F64 if prec == "f64" else gpu.LACE_F32
    dt = np.float64 if prec == "f64" else np.float32
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), lp)
    P.sweeps(3)
    R = np.random.default_rng(31).standard_normal((A.n_rows, 32)).astype(dt)
    lib = gpu.lib()
    for trans in (0, 1):
        for sweeps in (2, 3, 4, 5):
            outs = []
            try:
                for on in (1, 0):
                    gpu.check(lib.lace_set_option(b"jacobi_pair", on))
                    outs.append(P.apply_block(R, gpu.TRSV_JACOBI, sweeps=sweeps, trans=trans))
            finally:
                gpu.check(lib.lace_set_option(b"jacobi_pair", 1))
            assert np.array_equal(outs[0], outs[1]), (trans, sweeps)
            z = P.apply(R[:, 21].copy(), gpu.TRSV_JACOBI, sweeps=sweeps, trans=trans)
            assert np.array_equal(outs[0][:, 21], z), (trans, sweeps)


def _check_panel(gpu, W, r0, tol, unique_r=True):
    n = W.shape[0]
    Y, Tb, Rt, passes = gpu.block_panel_qr(W, r0)
    assert np.all(Y[:r0] == 0.0)
    top = Y[r0:r0 + 32]
    assert np.allclose(np.diag(top), 1.0) and np.all(np.triu(top, 1) == 0.0)
    assert np.all(np.tril(Rt, -1) == 0.0) and np.all(np.triu(Tb) == Tb)
    # (I - Y Tb Y^T) [Rt; 0] reproduces the panel
    E = np.zeros((n, 32)); E[r0:r0 + 32] = Rt
    HE = E - Y @ (Tb @ (Y.T @ E))
    scale = np.linalg.norm(W[r0:])
    assert np.linalg.norm(HE[r0:] - W[r0:]) <= tol * scale
    assert np.linalg.norm(HE[:r0]) <= tol * scale
    # orthogonality of the block reflector: Y^T Y = T^{-1} + T^{-T}
    Ti = np.linalg.inv(Tb)
    G = Y.T @ Y
    assert np.linalg.norm(G - Ti - Ti.T) <= tol * np.linalg.norm(G)
    if unique_r:
        # full column rank: the triangular factor is numpy's up to row signs
        Rn = np.linalg.qr(W[r0:], mode="r")
        assert np.allclose(np.abs(Rt), np.abs(Rn), rtol=1e-9, atol=tol * scale)
    return passes, Rt


@pytest.mark.parametrize("r0", [0, 96])
def test_panel_qr_well_conditioned(gpu, r0):
    W = np.random.default_rng(r0).standard_normal((3001, 32))
    W[:, 3] *= 1e-4      # widely scaled columns
    W[:, 17] *= 1e3
    assert _check_panel(gpu, W, r0, 1e-13)[0] == 2


def test_panel_qr_ill_conditioned_shifts(gpu):
    """kappa ~ 1e12: the unshifted Cholesky fails, the shifted first pass
    and two more passes still give an orthogonal reflector."""
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((4000, 32)))
    V, _ = np.linalg.qr(rng.standard_normal((32, 32)))
    W = U @ np.diag(np.logspace(0, -12, 32)) @ V.T
    passes, _ = _check_panel(gpu, W, 40, 1e-12)
    assert passes >= 3


def test_panel_qr_dependent_column(gpu):
    W = np.random.default_rng(4).standard_normal((2500, 32))
    W[:, 9] = W[:, 3] - 2.0 * W[:, 20]
    passes, Rt = _check_panel(gpu, W, 0, 1e-12, unique_r=False)
    assert passes >= 3
    assert abs(Rt[20, 20]) <= 1e-12 * np.linalg.norm(W)  # column 20 is the dependent one


def test_panel_qr_zero_column(gpu):
    """An exactly zero column: the QR of the panel with that column filled
    in, with the column zeroed in R, is still an exact factorization."""
    W = np.random.default_rng(6).standard_normal((2500, 32))
    W[:, 9] = 0.0
    W[:, 30] = 0.0
    passes, Rt = _check_panel(gpu, W, 64, 1e-13, unique_r=False)
    assert np.all(Rt[:, 9] == 0.0) and np.all(Rt[:, 30] == 0.0)


# ---------------------------------------------------------------- 32-RHS solver
def _bsolver(gpu, A, **kw):
    o = gpu.default_opts()
    o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = 30, 30, 1e-8, 3
    o.trsv, o.jacobi_sweeps, o.prec_precond = gpu.TRSV_JACOBI, 5, gpu.LACE_F32
    for k, v in kw.items():
        setattr(o, k, v)
    S = gpu.Solver(gpu.HostCSR.from_csr(A), o)
    S.setup()
    return S


def _rhs(gpu, n, seed0=0):
    return np.stack([gpu.gen_normal(seed0 + c, n) for c in range(32)], axis=1)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_block_solve_converges(gpu, orc, prec):
    A = orc.convdiff3d(16, 16, 16)
    lp = gpu.LACE_F32 if prec == "f32" else gpu.LACE_F64
    S = _bsolver(gpu, A, prec_precond=lp)
    B = _rhs(gpu, A.n_rows)
    X, st = S.solve_block(B)
    assert st.converged == 1
    As = A.scipy()
    rel = np.linalg.norm(B - As @ X, axis=0) / np.linalg.norm(B, axis=0)
    assert rel.max() <= 1.0e-8
    assert np.allclose(rel, np.array(st.rel_true_residual), rtol=1e-6, atol=1e-15)
    # against a direct solve (FP64 oracle error per column)
    import scipy.sparse.linalg as spla
    Xd = spla.spsolve(As.tocsc(), B)
    err = np.linalg.norm(X - Xd, axis=0) / np.linalg.norm(Xd, axis=0)
    assert err.max() <= 1e-6
    # each column agrees with the single-right-hand-side solver
    x5, st5 = S.solve(B[:, 5])
    assert st5.converged == 1
    assert np.linalg.norm(x5 - X[:, 5]) <= 1e-6 * np.linalg.norm(x5)


def test_block_solve_transposed_and_restart(gpu, orc):
    A = orc.convdiff3d(14, 13, 12)
    S = _bsolver(gpu, A, search_max=6, restart_max=200, rtol=1e-9)
    G = _rhs(gpu, A.n_rows, 100)
    L, st = S.solve_block(G, trans=1)
    assert st.converged == 1 and st.restarts > 1
    rel = np.linalg.norm(G - A.scipy().T @ L, axis=0) / np.linalg.norm(G, axis=0)
    assert rel.max() <= 1.0e-9
    # nonzero initial guess: the tolerance is relative to ||g_c - A^T x0_c||
    X0 = 0.9 * L
    L2, st2 = S.solve_block(G, X0=X0, trans=1)
    assert st2.converged == 1
    r0 = np.linalg.norm(G - A.scipy().T @ X0, axis=0)
    assert np.allclose(np.array(st2.initial_residual), r0, rtol=1e-10)
    r2 = np.linalg.norm(G - A.scipy().T @ L2, axis=0)
    assert np.all(r2 <= 1.0e-9 * r0)


@pytest.mark.parametrize("trans", [0, 1])
def test_block_solve_nested_inner(gpu, orc, trans):
    """Outer FP64 block FGMRES preconditioned by batched inner FP32 GMRES(8)
    per column: converges in far fewer outer block steps than with the ILU
    alone, to the same FP64 tolerance."""
    A = orc.convdiff3d(16, 16, 16)
    B = _rhs(gpu, A.n_rows, 400)
    As = A.scipy() if trans == 0 else A.scipy().T.tocsr()
    plain = _bsolver(gpu, A, rtol=1e-9)
    _, st0 = plain.solve_block(B, trans=trans)
    S = _bsolver(gpu, A, rtol=1e-9, inner_iters=8, search_max=10)
    X, st = S.solve_block(B, trans=trans)
    assert st.converged == 1 and st0.converged == 1
    rel = np.linalg.norm(B - As @ X, axis=0) / np.linalg.norm(B, axis=0)
    assert rel.max() <= 1e-9
    assert st.iterations * 3 <= st0.iterations


def test_block_solve_zero_column(gpu, orc):
    """A zero right-hand side among the 32 keeps x = 0 and the others
    converge as usual."""
    A = orc.convdiff3d(12, 12, 12)
    S = _bsolver(gpu, A)
    B = _rhs(gpu, A.n_rows, 50)
    B[:, 4] = 0.0
    X, st = S.solve_block(B)
    assert st.converged == 1
    assert np.all(X[:, 4] == 0.0)
    rel = np.linalg.norm(B - A.scipy() @ X, axis=0)
    nb = np.linalg.norm(B, axis=0)
    assert np.all(rel[nb > 0] <= 1e-8 * nb[nb > 0])


def test_block_gradient(gpu, orc):
    """Sum of 32 quadratic losses: dL/dtheta from one block adjoint solve,
    against the per-position formula and central differences."""
    A = orc.convdiff3d(11, 11, 11)
    n, nnz = A.n_rows, A.row[-1]
    S = _bsolver(gpu, A, prec_precond=gpu.LACE_F64, rtol=1e-13, restart_max=100)
    B = _rhs(gpu, n, 200)
    Xt = _rhs(gpu, n, 300)
    dB = gpu.DevVec(n * 32, np.float64, B.reshape(-1))
    dX = gpu.DevVec(n * 32, np.float64, np.zeros(n * 32))
    assert S.solve_block_dev(dB, dX).converged == 1
    X = dX.download().reshape(n, 32)
    dG = gpu.DevVec(n * 32, np.float64, (X - Xt).reshape(-1))
    dL = gpu.DevVec(n * 32, np.float64, np.zeros(n * 32))
    assert S.solve_block_dev(dG, dL, trans=1).converged == 1
    Lam = dL.download().reshape(n, 32)
    g = S.grad_values_block(dL, dX, nnz)
    rows = np.repeat(np.arange(n), np.diff(A.row))
    want = -np.einsum("kc,kc->k", Lam[rows], X[A.col])
    assert np.allclose(g, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())

    def loss(vals):
        S.set_values(vals)
        Xv, stv = S.solve_block(B)
        assert stv.converged == 1
        return 0.5 * np.sum((Xv - Xt) ** 2)

    h = 1e-6
    for seed in (100, 101):
        d = gpu.gen_normal(seed, nnz)
        fd = (loss(A.val + h * d) - loss(A.val - h * d)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * abs(g @ d)
    S.set_values(A.val)
