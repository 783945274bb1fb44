"""Degenerate and ragged inputs through the device path: a 1 x 1 matrix, a
diagonal matrix (no off-diagonal entries: no diagonal or ELLPACK form of the
factors), rows of very different lengths with empty off-diagonal parts, a
zero right-hand side, and the rejected inputs (a row without its diagonal,
unsorted rows).  Exact results are checked bit for bit where IEEE arithmetic
fixes them (r / d, d * x), against the oracle otherwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _opts(gpu, **kw):
    o = gpu.default_opts()
    o.rtol, o.parilu_sweeps, o.search_max, o.restart_max = 1e-10, 3, 20, 20
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_one_by_one(gpu, orc):
    A = orc.CSR(1, 1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([4.0]))
    H = gpu.HostCSR.from_csr(A)
    P = gpu.ParILU(H, gpu.LACE_F64)
    P.sweeps(2)
    (lr, lc, lv), (ur, uc, uv) = P.factors()
    assert lv.tolist() == [1.0] and uv.tolist() == [4.0]
    for kind in (gpu.TRSV_LEVEL, gpu.TRSV_JACOBI):
        assert P.apply(np.array([2.0]), kind, sweeps=5).tolist() == [0.5]
    M = gpu.DevCSRMatrix(H, gpu.LACE_F64)
    assert M.spmv(np.array([3.0])).tolist() == [12.0]
    S = gpu.Solver(H, _opts(gpu, trsv=gpu.TRSV_LEVEL, prec_precond=gpu.LACE_F64))
    S.setup()
    x, st = S.solve(np.array([2.0]))
    assert st.converged == 1 and abs(x[0] - 0.5) <= 1e-15


def test_diagonal_matrix(gpu, orc):
    n = 1000
    d = 1.0 + np.random.default_rng(1).random(n)
    A = orc.CSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    H = gpu.HostCSR.from_csr(A)
    P = gpu.ParILU(H, gpu.LACE_F64)
    assert P.forms(False) == (0, 0) and P.forms(True) == (0, 0)
    P.sweeps(3)
    r = np.random.default_rng(2).standard_normal(n)
    for kind in (gpu.TRSV_LEVEL, gpu.TRSV_JACOBI):
        assert np.array_equal(P.apply(r, kind, sweeps=5), r / d)
    M = gpu.DevCSRMatrix(H, gpu.LACE_F64)
    assert M.forms()[0] == 1
    assert np.array_equal(M.spmv(r), d * r)
    S = gpu.Solver(H, _opts(gpu, trsv=gpu.TRSV_JACOBI, jacobi_sweeps=5, prec_precond=gpu.LACE_F64))
    S.setup()
    x, st = S.solve(r)
    assert st.converged == 1 and st.iterations <= 1
    assert np.allclose(x, r / d, rtol=1e-14, atol=0)


def _ragged(orc, n=2000, seed=3, most=10):
    """Symmetric pattern: every row its diagonal; 20 % of the rows nothing
    else, the rest up to `most` random partners (each pair lands in both
    rows, so rows run from 1 to a few times `most` entries); diagonally
    dominant values."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        if rng.random() < 0.2:
            continue
        for j in rng.choice(n, size=rng.integers(1, most + 1), replace=False):
            if j != i:
                rows += [i, j]
                cols += [j, i]
    key = np.unique(np.array(rows, np.int64) * n + np.array(cols, np.int64))
    r, c = key // n, key % n
    v = rng.uniform(-1.0, 0.0, len(r))
    off = orc._coo_to_csr(n, r, c, v)
    dsum = np.zeros(n)
    np.add.at(dsum, r, np.abs(v))
    rr = np.concatenate([r, np.arange(n)])
    cc = np.concatenate([c, np.arange(n)])
    vv = np.concatenate([v, 1.0 + dsum])
    return orc._coo_to_csr(n, rr, cc, vv), off


def test_ragged_rows_vs_oracle(gpu, orc):
    A, _ = _ragged(orc)
    lens = np.diff(A.row)
    # no ELLPACK form (rows beyond 8), CSR-stream products (mean under 16)
    assert lens.min() == 1 and lens.max() > 8 and lens.mean() < 16
    H = gpu.HostCSR.from_csr(A)
    P = gpu.ParILU(H, gpu.LACE_F64)
    P.sweeps(3)
    st = orc.parilu(A, 3)
    (lr, lc, lv), (ur, uc, uv) = P.factors()
    assert np.array_equal(lv, st.Lval) and np.array_equal(uv, st.Uval)
    Lf = orc.CSR(A.n_rows, A.n_cols, st.L.row, st.L.col, st.Lval)
    Ur = orc.transpose(orc.CSR(A.n_rows, A.n_cols, st.U.row, st.U.col, st.Uval))
    r = np.random.default_rng(4).standard_normal(A.n_rows)
    y = orc.trsv_jacobi(Lf, r, upper=False, sweeps=4)
    ref = orc.trsv_jacobi(Ur, y, upper=True, sweeps=4)
    assert np.array_equal(P.apply(r, gpu.TRSV_JACOBI, sweeps=4), ref)
    y = orc.trsv(Lf, r, upper=False)
    assert np.array_equal(P.apply(r, gpu.TRSV_LEVEL), orc.trsv(Ur, y, upper=True))
    M = gpu.DevCSRMatrix(H, gpu.LACE_F64)
    assert M.forms() == (0, 0)
    assert np.array_equal(M.spmv(r), orc.spmv(A, r))
    S = gpu.Solver(H, _opts(gpu, trsv=gpu.TRSV_JACOBI, jacobi_sweeps=5, prec_precond=gpu.LACE_F32,
                            inner_iters=5, search_max=10))
    S.setup()
    x, stt = S.solve(r)
    assert stt.converged == 1 and np.linalg.norm(r - A.scipy() @ x) <= 1.01e-10 * np.linalg.norm(r)
    # long rows (mean beyond 16): the lane-group SpMV sums each row in a
    # shuffle tree, so it meets the CSR-order sum to rounding, not bitwise
    B, _ = _ragged(orc, n=1500, seed=5, most=40)
    assert np.diff(B.row).mean() > 16
    MB = gpu.DevCSRMatrix(gpu.HostCSR.from_csr(B), gpu.LACE_F64)
    xb = np.random.default_rng(6).standard_normal(B.n_rows)
    ref = orc.spmv(B, xb)
    scale = np.abs(B.scipy()) @ np.abs(xb)
    assert np.all(np.abs(MB.spmv(xb) - ref) <= 1e-14 * scale)


def test_zero_rhs(gpu, orc):
    A = orc.convdiff3d(8, 8, 8)
    S = gpu.Solver(gpu.HostCSR.from_csr(A), _opts(gpu, trsv=gpu.TRSV_JACOBI, jacobi_sweeps=5,
                                                   prec_precond=gpu.LACE_F32, inner_iters=4, search_max=10))
    S.setup()
    x, st = S.solve(np.zeros(A.n_rows))
    assert st.converged == 1 and st.iterations == 0 and not np.any(x)


def test_rejected_inputs(gpu, orc):
    # row 1 has no diagonal entry
    row = np.array([0, 2, 3, 5], np.int32)
    col = np.array([0, 1, 0, 1, 2], np.int32)
    A = orc.CSR(3, 3, row, col, np.ones(5))
    with pytest.raises(gpu.LaceError, match="no diagonal"):
        gpu.ParILU(gpu.HostCSR.from_csr(A))
    # unsorted row 0
    row = np.array([0, 2, 3, 4], np.int32)
    col = np.array([1, 0, 1, 2], np.int32)
    A = orc.CSR(3, 3, row, col, np.ones(4))
    with pytest.raises(gpu.LaceError, match="not sorted"):
        gpu.ParILU(gpu.HostCSR.from_csr(A))
