"""Pin the CPU oracle before trusting it (CPU only).

Each oracle function is checked against something independent of its own
formula: the golden arrays of the reference's unit tests, Random123's
known-answer vectors, the reference's Householder fixtures, closed forms,
dense numpy/scipy routines, and brute force on tiny inputs.  Several
checks plant a mistake in a copy of the oracle's output to show the check
would catch it.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import GOLDEN


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLDEN, "reference_goldens.json")) as f:
        return json.load(f)


def _hex(v):
    return int(v, 16) if isinstance(v, str) else int(v)


# ------------------------------------------------------------------ RNG
def test_philox_known_answers(orc, gold):
    for case in gold["philox4x32_10_kat"]["cases"]:
        out = orc.philox4x32_10([_hex(v) for v in case["ctr"]], [_hex(v) for v in case["key"]])
        assert [int(x) for x in out] == [_hex(v) for v in case["out"]]


# ------------------------------------------------------------------ formats
def test_reader_and_splits_io_test(orc, gold):
    g = gold["io_test"]
    A = orc.read_mtx(os.path.join(GOLDEN, "io_test.mtx"))
    assert A.row.tolist() == g["csr_row"]
    assert A.col.tolist() == g["csr_col"]
    assert A.val.tolist() == g["csr_val"]
    T = orc.transpose(A)
    assert T.row.tolist() == g["trans_row"]
    assert T.val.tolist() == g["trans_val"]
    # columns of the transpose follow from the dense transpose
    assert np.array_equal(T.dense(), A.dense().T)
    Lv = orc.csr_split(A, upper=False, diag_mode=orc.DIAG_VALUE)
    assert Lv.row.tolist() == g["csrl_value_row"]
    assert Lv.val.tolist() == g["csrl_value_val"]
    assert np.array_equal(Lv.dense(), np.tril(A.dense()))
    Uv = orc.csr_split(A, upper=True, diag_mode=orc.DIAG_VALUE)
    assert Uv.row.tolist() == g["csru_row"]
    assert Uv.val.tolist() == g["csru_val"]
    assert np.array_equal(Uv.dense(), np.triu(A.dense()))
    # diagonal placement: last in CSRL rows, first in CSRU rows
    for i in range(A.n_rows):
        assert Lv.col[Lv.row[i + 1] - 1] == i
        assert Uv.col[Uv.row[i]] == i


def test_trefethen_reader_closed_form(orc):
    """Trefethen_20: primes on the diagonal, 1 where |i-j| is a power of 2."""
    A = orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx"))
    D = A.dense()
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71]
    n = A.n_rows
    ref = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                ref[i, j] = primes[i + 1]
            elif abs(i - j) in (1, 2, 4, 8, 16):
                ref[i, j] = 1.0
    assert np.array_equal(D, ref)


def test_trefethen_dense_subblocks(orc, gold):
    g = gold["trefethen_dense_sub"]
    A = orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx"))
    sub = lambda M: M.dense()[:5, :5].ravel().tolist()
    assert sub(A) == g["A"]
    L = orc.csr_split(A, False, orc.DIAG_UNIT)
    U = orc.csr_split(A, True, orc.DIAG_VALUE)
    assert sub(L) == g["L_unity"]
    assert sub(U) == g["U_value"]
    prod = (L.dense()[:5, :5] @ U.dense()[:5, :5]).ravel().tolist()
    assert prod == g["LU_product"]


def test_spgemm_goldens(orc, gold):
    g = gold["trefethen_spgemm"]
    As = orc.CSR(5, 5, np.array(g["Asub_row"]), np.array(g["Asub_col"]),
                 np.array(g["Asub_val"], float))
    # the sub-block itself comes out of the reader
    A = orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx"))
    assert np.array_equal(As.dense(), A.dense()[:5, :5])
    C = orc.spgemm(As, As)
    assert C.row.tolist() == g["C_row"] and C.col.tolist() == g["C_col"]
    assert C.val.tolist() == g["C_val"]
    g2 = gold["trefethen_spgemm_lu"]
    Ls = orc.CSR(5, 5, np.array(g2["Lsub_row"]), np.array(g2["Lsub_col"]),
                 np.array(g2["Lsub_val"], float))
    Us = orc.CSR(5, 5, np.array(g2["Usub_row"]), np.array(g2["Usub_col"]),
                 np.array(g2["Usub_val"], float))
    C2 = orc.spgemm(Ls, Us)
    assert C2.row.tolist() == g2["C_row"] and C2.col.tolist() == g2["C_col"]
    assert C2.val.tolist() == g2["C_val"]


def test_spgemm_random_vs_dense(orc):
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    for _ in range(5):
        a = sp.random(30, 25, density=0.15, random_state=rng, format="csr")
        b = sp.random(25, 20, density=0.2, random_state=rng, format="csr")
        a.sort_indices(); b.sort_indices()
        A = orc.CSR(30, 25, a.indptr, a.indices, a.data)
        B = orc.CSR(25, 20, b.indptr, b.indices, b.data)
        C = orc.spgemm(A, B)
        assert np.allclose(C.dense(), (a @ b).toarray(), rtol=1e-14, atol=1e-15)


# ------------------------------------------------------------------ ILU(0)
def test_ilu0_equals_dense_lu_without_fill(orc):
    """Tridiagonal matrices have no fill: ILU(0) is the no-pivot LU."""
    n = 40
    rng = np.random.default_rng(1)
    D = np.diag(4 + rng.random(n)) + np.diag(rng.random(n - 1) - 0.5, 1) + \
        np.diag(rng.random(n - 1) - 0.5, -1)
    import scipy.sparse as sp
    a = sp.csr_matrix(D)
    A = orc.CSR(n, n, a.indptr, a.indices, a.data)
    LU, bad = orc.ilu0(A)
    assert bad < 0
    Lf, Uf = orc.split_merged(LU)
    # Doolittle LU without pivoting
    M = D.copy()
    for k in range(n):
        M[k + 1:, k] /= M[k, k]
        M[k + 1:, k + 1:] -= np.outer(M[k + 1:, k], M[k, k + 1:])
    assert np.allclose(np.tril(M, -1) + np.eye(n), Lf.dense(), rtol=1e-14, atol=1e-15)
    assert np.allclose(np.triu(M), Uf.dense(), rtol=1e-14, atol=1e-15)


def test_ilu0_pattern_invariant(orc):
    """(LU)_ij = a_ij on the pattern of A (zero nonlinear residual)."""
    A = orc.convdiff3d(7, 6, 5)
    LU, bad = orc.ilu0(A)
    Lf, Uf = orc.split_merged(LU)
    P = Lf.dense() @ Uf.dense()
    Ad = A.dense()
    mask = Ad != 0
    assert np.allclose(P[mask], Ad[mask], rtol=1e-13, atol=1e-13)
    # fill outside the pattern exists (7-point ILU(0) is inexact)
    assert np.abs(P[~mask]).max() > 1e-3
    res, nl = orc.ilures(A, Lf, Uf)
    assert nl < 1e-12
    assert abs(res - np.linalg.norm(P - Ad)) < 1e-10
    # a planted error in one factor entry shows up in nonlinres
    Lbad = orc.CSR(Lf.n_rows, Lf.n_cols, Lf.row, Lf.col, Lf.val.copy())
    Lbad.val[3] += 1e-3
    assert orc.ilures(A, Lbad, Uf)[1] > 1e-4


def test_ilu0_small_pivot(orc):
    A = orc.CSR(2, 2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([1.0, 1.0, 1.0, 1.0]))
    _, bad = orc.ilu0(A)
    assert bad == 1


# ------------------------------------------------------------------ ParILU
def test_parilu_first_sweep_closed_form(orc):
    """Sweep 1 from L0 = tril(A,-1)+I, U0 = triu(A):
    l_ij = (a_ij - sum_{k<j} a_ik a_kj)/a_jj, u_ij = a_ij - sum_{k<i} a_ik a_kj,
    on A's pattern (dense numpy)."""
    for A in (orc.convdiff3d(5, 4, 3), orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx")),
              orc.read_mtx(os.path.join(GOLDEN, "io_test.mtx"))):
        Ad = A.dense()
        n = A.n_rows
        Ls = np.tril(Ad, -1)
        U0 = np.triu(Ad)
        P = Ls @ U0   # sum over k < i of a_ik u_kj (k <= j via triu)
        ref = np.zeros_like(Ad)
        for i in range(n):
            for j in range(n):
                if Ad[i, j] == 0:
                    continue
                if i > j:
                    s = P[i, j] - Ad[i, j] * Ad[j, j]  # drop k = j
                    ref[i, j] = (Ad[i, j] - s) / Ad[j, j]
                else:
                    ref[i, j] = Ad[i, j] - P[i, j]
        st = orc.ParILUState(A)
        st.sweep("jacobi")
        Lf, Uf = st.factors()
        got = np.tril(Lf.dense(), -1) + Uf.dense()
        scale = np.abs(ref).max()
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13 * scale)


def test_parilu_finite_termination_equals_gauss_seidel(orc):
    """Jacobi sweeps reach the fixed point after DAG-depth sweeps and then
    equal one row-order (Gauss-Seidel) sweep bit for bit; that fixed point
    is ILU(0) up to rounding."""
    A = orc.convdiff3d(5, 5, 4)
    jac = orc.ParILUState(A)
    sweeps = 0
    while sweeps < 500:
        Lp, Up = jac.Lval.copy(), jac.Uval.copy()
        jac.sweep("jacobi")
        sweeps += 1
        if np.array_equal(Lp, jac.Lval) and np.array_equal(Up, jac.Uval):
            break
    assert sweeps < 500
    gs = orc.parilu(A, 1, "gs")
    assert np.array_equal(jac.Lval, gs.Lval) and np.array_equal(jac.Uval, gs.Uval)
    ex, _ = orc.ilu0(A)
    Le, Ue = orc.split_merged(ex)
    Lj, Uj = jac.factors()
    assert np.allclose(Lj.val, Le.val, rtol=1e-13, atol=1e-15)
    assert np.allclose(Uj.val, Ue.val, rtol=1e-13, atol=1e-15)
    # two sweeps short of the fixed point the factors still differ
    short = orc.parilu(A, sweeps - 3, "jacobi")
    assert not np.array_equal(short.Lval, gs.Lval) or not np.array_equal(short.Uval, gs.Uval)


def test_parilu_brute_force_tiny(orc):
    """Pure-Python dense Jacobi ParILU (reads from the previous iterate) on a
    random 8x8 pattern with diagonal."""
    rng = np.random.default_rng(11)
    n = 8
    M = (rng.random((n, n)) < 0.4) * rng.standard_normal((n, n))
    M += np.diag(5 + rng.random(n))
    import scipy.sparse as sp
    a = sp.csr_matrix(M)
    a.sort_indices()
    A = orc.CSR(n, n, a.indptr, a.indices, a.data)
    L = np.tril(M, -1) + np.eye(n)
    U = np.triu(M)
    st = orc.ParILUState(A)
    for _ in range(4):
        Ln, Un = L.copy(), U.copy()
        for i in range(n):
            for j in range(n):
                if M[i, j] == 0:
                    continue
                m = min(i, j)
                s = M[i, j] - sum(L[i, k] * U[k, j] for k in range(m))
                if i > j:
                    Ln[i, j] = s / U[j, j]
                else:
                    Un[i, j] = s
        L, U = Ln, Un
        st.sweep("jacobi")
    Lf, Uf = st.factors()
    assert np.allclose(Lf.dense(), L, rtol=1e-12, atol=1e-12)
    assert np.allclose(Uf.dense(), U, rtol=1e-12, atol=1e-12)


def test_parilu_v03_omp_one_thread_is_gauss_seidel(orc):
    A = orc.convdiff3d(4, 4, 4)
    orc.set_threads(1)
    try:
        st, it, _, _ = orc.parilu_v03_omp(A, 1)
    finally:
        orc.set_threads(0)
    gs = orc.parilu(A, 1, "gs")
    assert np.array_equal(st.Lval, gs.Lval) and np.array_equal(st.Uval, gs.Uval)


# ------------------------------------------------------------------ trsv
def test_trsv_vs_scipy(orc):
    A = orc.convdiff3d(6, 5, 4)
    st = orc.parilu(A, 3)
    Lf, Uf = st.factors()
    r = np.random.default_rng(0).standard_normal(A.n_rows)
    y = orc.trsv(Lf, r, upper=False)
    assert np.allclose(y, sla.solve_triangular(Lf.dense(), r, lower=True), rtol=1e-12)
    z = orc.trsv(Uf, r, upper=True)
    assert np.allclose(z, sla.solve_triangular(Uf.dense(), r, lower=False), rtol=1e-12)


def test_trsv_jacobi_neumann_closed_form(orc):
    """y_k = sum_{m<k} (-D^{-1} N)^m D^{-1} r from y0 = 0."""
    A = orc.convdiff3d(5, 4, 3)
    Lf, Uf = orc.parilu(A, 2).factors()
    r = np.random.default_rng(2).standard_normal(A.n_rows)
    for T, upper in ((Lf, False), (Uf, True)):
        Td = T.dense()
        D = np.diag(np.diag(Td))
        Nm = Td - D
        Di = np.linalg.inv(D)
        for k in (1, 2, 5):
            ref = np.zeros_like(r)
            term = Di @ r
            for _ in range(k):
                ref += term
                term = -Di @ Nm @ term
            got = orc.trsv_jacobi(T, r, upper, k)
            assert np.allclose(got, ref, rtol=1e-12, atol=1e-13)


def test_parcsrtrsv_one_thread_is_exact(orc):
    """One thread, natural order: exact substitution, then a confirming
    sweep (src/trisolve.cpp:274-295 with tol 1.0)."""
    A = orc.convdiff3d(5, 5, 5)
    Lf, Uf = orc.parilu(A, 3).factors()
    r = np.random.default_rng(4).standard_normal(A.n_rows)
    orc.set_threads(1)
    try:
        y, it = orc.parcsrtrsv(Lf, r, upper=False, tol=1.0)
        z, it2 = orc.parcsrtrsv(Uf, r, upper=True, tol=1.0)
    finally:
        orc.set_threads(0)
    assert it == 2 and it2 == 2
    assert np.array_equal(y, orc.trsv(Lf, r, upper=False))
    assert np.array_equal(z, orc.trsv(Uf, r, upper=True))


def test_levels_closed_form_and_brute_force(orc):
    nx, ny = 9, 7
    A = orc.laplace2d(nx, ny)
    L = orc.csr_split(A, False, orc.DIAG_UNIT)
    lev, nl = orc.levels(L, False)
    assert nl == nx + ny - 1
    ij = np.array([i + j for j in range(ny) for i in range(nx)])
    assert np.array_equal(lev, ij)
    U = orc.csr_split(A, True, orc.DIAG_VALUE)
    levu, nlu = orc.levels(U, True)
    assert nlu == nl
    assert np.array_equal(levu, (nx - 1 + ny - 1) - ij)
    B = orc.convdiff3d(4, 3, 5)
    levb, nlb = orc.levels(orc.csr_split(B, False, orc.DIAG_UNIT), False)
    ijk = np.array([i + j + k for k in range(5) for j in range(3) for i in range(4)])
    assert np.array_equal(levb, ijk) and nlb == 4 + 3 + 5 - 2
    # brute force longest path on a random lower pattern
    rng = np.random.default_rng(5)
    n = 30
    M = np.tril(rng.random((n, n)) < 0.15, -1) + np.eye(n)
    import scipy.sparse as sp
    a = sp.csr_matrix(M.astype(float))
    T = orc.CSR(n, n, a.indptr, a.indices, a.data)
    got, _ = orc.levels(T, False)
    depth = [0] * n
    for i in range(n):
        preds = [j for j in range(i) if M[i, j]]
        depth[i] = max([depth[j] + 1 for j in preds], default=0)
    assert got.tolist() == depth


def test_spmv_vs_scipy(orc):
    A = orc.convdiff3d(8, 7, 6)
    x = np.random.default_rng(6).standard_normal(A.n_rows)
    y0 = np.random.default_rng(7).standard_normal(A.n_rows)
    assert np.allclose(orc.spmv(A, x, 2.0, -1.5, y0), 2.0 * (A.scipy() @ x) - 1.5 * y0,
                       rtol=1e-14, atol=1e-13)


def test_givens_closed_form(orc):
    rng = np.random.default_rng(8)
    for a, b in list(rng.standard_normal((20, 2))) + [(0.0, 1.0), (3.0, 0.0), (1e-300, 2.0)]:
        c, s = orc.givens(a, b, np.nextafter(0, 1))
        assert abs(c * c + s * s - 1) < 1e-15
        assert abs(-s * a + c * b) <= 1e-15 * max(abs(a), abs(b))
        assert abs(abs(c * a + s * b) - np.hypot(a, b)) <= 1e-15 * np.hypot(a, b)
    assert orc.givens(2.0, 0.0, 1e-300) == (1.0, 0.0)


# ------------------------------------------------------------------ FGMRES
def _min_residual_history(A, b, Minv, steps):
    """||r_j|| = min over x in K_j(A M^{-1}, r0) of ||b - A M^{-1} y||, by
    dense least squares on an orthonormalised Krylov basis."""
    AM = A @ Minv
    hist = []
    V = [b / np.linalg.norm(b)]
    for j in range(1, steps + 1):
        Q, _ = np.linalg.qr(np.array(V).T)
        K = AM @ Q
        y, *_ = np.linalg.lstsq(K, b, rcond=None)
        hist.append(np.linalg.norm(b - K @ y))
        w = AM @ V[-1]
        for v in V:
            w -= (v @ w) * v
        for v in V:
            w -= (v @ w) * v
        V.append(w / np.linalg.norm(w))
    return np.array(hist)


@pytest.mark.parametrize("kind", ["householder", "mgs"])
@pytest.mark.parametrize("choice", [0, 2])
def test_fgmres_residual_history_is_minimal(orc, kind, choice):
    A = orc.convdiff3d(5, 4, 3)
    b = np.random.default_rng(9).standard_normal(A.n_rows)
    Lf, Uf = orc.parilu(A, 2).factors()
    fn = orc.fgmres_householder if kind == "householder" else orc.fgmres_mgs
    res = fn(A, b, m=25, restart_max=1, rtol=1e-14, tol_type=1, choice=choice, Lf=Lf, Uf=Uf)
    Ad = A.dense()
    Minv = np.linalg.inv(Lf.dense() @ Uf.dense()) if choice == 0 else np.eye(A.n_rows)
    ref = _min_residual_history(Ad, b, Minv, len(res.hist))
    k = min(len(res.hist), 15)
    assert np.allclose(res.hist[:k], ref[:k], rtol=1e-8, atol=1e-12 * np.linalg.norm(b))


@pytest.mark.parametrize("kind", ["householder", "mgs"])
def test_fgmres_identity_and_diagonal(orc, kind):
    fn = orc.fgmres_householder if kind == "householder" else orc.fgmres_mgs
    n = 12
    I = orc.CSR(n, n, np.arange(n + 1), np.arange(n), np.ones(n))
    b = np.random.default_rng(1).standard_normal(n)
    r = fn(I, b, m=5, restart_max=1, rtol=1e-12, tol_type=1, choice=2)
    assert r.searches == 1 and np.allclose(r.x, b, rtol=1e-14)
    d = np.repeat([1.0, 2.0, 5.0], 4)  # three distinct eigenvalues
    D = orc.CSR(n, n, np.arange(n + 1), np.arange(n), d)
    r = fn(D, b, m=10, restart_max=1, rtol=1e-12, tol_type=1, choice=2)
    assert r.info == 0 and r.searches <= 3
    assert np.allclose(r.x, b / d, rtol=1e-10)


def test_fgmres_true_residual_tolerance(orc):
    A = orc.laplace2d(20, 20)
    b = np.ones(A.n_rows)
    Lf, Uf = orc.parilu(A, 5).factors()
    for fn in (orc.fgmres_householder, orc.fgmres_mgs):
        r = fn(A, b, m=15, restart_max=20, rtol=1e-10, tol_type=1, choice=0, Lf=Lf, Uf=Uf)
        assert r.info == 0
        assert np.linalg.norm(b - A.scipy() @ r.x) <= 1.01e-10 * np.linalg.norm(b)


# ------------------------------------------------------------------ Householder QR
def test_hqrd_magic7_fixture(orc):
    X = orc.read_mtx_dense(os.path.join(GOLDEN, "magic7.mtx"))
    Uref = orc.read_mtx_dense(os.path.join(GOLDEN, "magic7_Householder_U.mtx"))
    Rref = orc.read_mtx_dense(os.path.join(GOLDEN, "magic7_Householder_R.mtx"))
    U, R = orc.hqrd(X)
    assert np.allclose(U, Uref, rtol=1e-12, atol=1e-12)
    assert np.allclose(R, Rref, rtol=1e-12, atol=1e-11)
    for k in range(7):
        assert abs(np.linalg.norm(U[k:, k]) - np.sqrt(2)) < 1e-13
    assert np.allclose(np.abs(R), np.abs(np.linalg.qr(X)[1]), atol=1e-10)


def _frexp_scale(m):
    import math
    if m == 0 or not np.isfinite(m):
        return 1.0
    _, e = math.frexp(m)
    return 2.0 ** max(-126, min(127, e - 1))


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_parilu_16bit_storage_model(orc, storage):
    """Scalar 16-bit factor storage: each value rounded relative to the
    power-of-two scale of its own row.  Rows span 2^-12..2^12 in magnitude,
    so a scale taken from the wrong row (or column) would lose the value.
    Per sweep, the oracle's state equals a dense float64 Jacobi sweep from
    the oracle's previous state, rounded with independently computed row
    scales (math.frexp) and numpy's / torch's 16-bit conversions, to within
    one 16-bit step of the row (the sweeps' FP32 rounding may move a value
    across a rounding boundary); every stored value sits on the 16-bit grid
    of its row, and L's unit diagonal stays exact."""
    rng = np.random.default_rng(12)
    n = 10
    M = (rng.random((n, n)) < 0.4) * rng.standard_normal((n, n))
    M += np.diag(5 + rng.random(n))
    M *= (2.0 ** rng.integers(-12, 13, size=n))[:, None]
    import scipy.sparse as sp
    a = sp.csr_matrix(M)
    a.sort_indices()
    A = orc.CSR(n, n, a.indptr, a.indices, a.data)
    s = np.array([_frexp_scale(np.abs(M[i]).max()) for i in range(n)])
    assert np.array_equal(orc.scalar_row_scales(A), s.astype(np.float32))

    def to16(x):
        x = np.asarray(x, np.float32)
        if storage == "f16":
            return x.astype(np.float16).astype(np.float32)
        torch = pytest.importorskip("torch")
        return torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()

    def q(D):  # round each entry with its row's scale
        return to16(D / s[:, None]).astype(np.float64) * s[:, None]

    step = (2.0 ** -7 if storage == "bf16" else 2.0 ** -10) * 2.0
    st = orc.ParILUState(A, storage=storage)
    mask = M != 0
    for k in range(4):
        Lf, Uf = st.factors()
        L, U = Lf.dense(), Uf.dense()
        assert np.all(np.diag(L) == 1.0)
        # every stored value on its row's 16-bit grid
        assert np.array_equal(q(np.where(mask, np.tril(L, -1) + np.triu(U), 0)),
                              np.where(mask, np.tril(L, -1) + np.triu(U), 0))
        Ln, Un = L.copy(), U.copy()
        for i in range(n):
            for j in range(n):
                if not mask[i, j]:
                    continue
                m = min(i, j)
                t = M[i, j] - sum(L[i, kk] * U[kk, j] for kk in range(m))
                if i > j:
                    Ln[i, j] = t / U[j, j]
                else:
                    Un[i, j] = t
        Lw = np.where(np.tril(mask, -1), q(Ln), np.eye(n))
        Uw = np.where(np.triu(mask), q(Un), 0.0)
        st.sweep("jacobi")
        Lf, Uf = st.factors()
        tol = step * np.maximum(np.abs(Lw), s[:, None] * 2.0 ** -20)
        assert np.all(np.abs(Lf.dense() - Lw) <= tol), k
        tol = step * np.maximum(np.abs(Uw), s[:, None] * 2.0 ** -20)
        assert np.all(np.abs(Uf.dense() - Uw) <= tol), k


def test_parilu_tiled_one_tile_first_sweep_closed_form(orc):
    """The tiled ParILU restatement with one tile covering A: every entry of
    the first sweep reads the initial values (strict L and U of A, D = 1 /
    diag A), so U1 = A_U - (A_L A_U), D1 = 1 / (diag A - diag(A_L A_U)) and
    L1 = (A_L - A_L A_U) D1 on A's pattern; with small integer entries the
    product is exact and the closed form is computed with numpy.  A wrong
    snapshot (an update read back inside the tile), the L update using the
    old D, or the product missing the diagonal's terms would all differ."""
    rng = np.random.default_rng(3)
    n = 7
    M = (rng.random((n, n)) < 0.5) * rng.integers(-3, 4, (n, n)).astype(float)
    M += np.diag(rng.integers(5, 9, n).astype(float))
    import scipy.sparse as sp
    a = sp.csr_matrix(M)
    a.sort_indices()
    A = orc.CSR(n, n, a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data)
    lv, uv, it = orc.parilu_tiled(A, 16, max_iter=1)
    assert it == 1
    AL, AU, d = np.tril(M, -1), np.triu(M, 1), np.diag(M)
    Cm = AL @ AU
    U1 = AU - Cm
    D1 = 1.0 / (d - np.diag(Cm))
    L1 = (AL - Cm) * D1[None, :]
    want_l, want_u = [], []
    for i in range(n):
        want_l += [L1[i, j] for j in range(i) if M[i, j] != 0] + [1.0]
        want_u += [1.0 / D1[i]] + [U1[i, j] for j in range(i + 1, n) if M[i, j] != 0]
    assert np.array_equal(lv, np.array(want_l)) and np.array_equal(uv, np.array(want_u))
