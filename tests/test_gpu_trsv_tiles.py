"""Exact triangular solves in box order (option "trsv_exact_tiles").

Where the factors' rows are the points of a grid and every dependency steps
backwards along each axis, the persistent exact solve takes its blocks box
by box instead of as runs of consecutive rows.  The renumbering moves whole
rows and keeps each row's entries in order, so the result must be
bit-identical to forward/backward substitution (the oracle) and to the
consecutive-row blocks -- for L and U, for the transposed solves, after the
factor values change, with 256-row blocks, and on grids whose sides are not
multiples of the box.  Patterns that only look like a grid (offsets 1, nx,
nx*ny with entries that wrap round a grid line) and general patterns keep
consecutive-row blocks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _wrap(orc):
    """7-point pattern plus the entries (i, i-1) and (i-1, i) that join
    grid lines (x = 0 to x = nx-1 of the line before): distances stay
    {1, nx, nx*ny} but those steps go forwards in x."""
    A = orc.convdiff3d(12, 10, 9)
    ri = np.repeat(np.arange(A.n_rows), np.diff(A.row))
    i = np.arange(12, A.n_rows, 12)
    rows = np.concatenate([ri, i, i - 1])
    cols = np.concatenate([A.col, i - 1, i])
    vals = np.concatenate([A.val, np.full(len(i), -0.25), np.full(len(i), -0.25)])
    return orc._coo_to_csr(A.n_rows, rows, cols, vals)


def _problem(orc, name):
    if name == "convdiff":
        return orc.convdiff3d(20, 13, 11)  # sides not multiples of 8
    if name == "tiny":
        return orc.convdiff3d(3, 3, 2)  # one box, clipped to the grid
    if name == "chain":
        return orc.laplace2d(50, 1)  # distances {1} only: no grid to tile
    if name == "laplace2d":
        return orc.laplace2d(40, 37)
    if name == "wrap":
        return _wrap(orc)
    import os
    from conftest import GOLDEN
    return orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx"))


WANT = {"convdiff": (8, 8, 8), "laplace2d": (32, 16, 1), "wrap": (0, 0, 0), "trefethen20": (0, 0, 0),
        "tiny": (3, 3, 2), "chain": (0, 0, 0)}


def _opt(gpu, name, value):
    gpu.check(gpu.lib().lace_set_option(name.encode(), value))


@pytest.mark.parametrize("name", ["convdiff", "laplace2d", "wrap", "trefethen20", "tiny", "chain"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_exact_solve_box_order(gpu, orc, name, prec):
    A = _problem(orc, name)
    lp, dt = (gpu.LACE_F64, np.float64) if prec == "f64" else (gpu.LACE_F32, np.float32)
    print(f"box order: {name} {prec} n={A.n_rows}", flush=True)
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), lp)
    for w in range(4):
        assert P.tiles(w) == WANT[name], (w, P.tiles(w))
    r = np.random.default_rng(11).standard_normal(A.n_rows).astype(dt)
    for sweeps, more in ((3, 3), (5, 2)):
        P.sweeps(more)  # the factor values change between the two rounds
        st = orc.parilu(A, sweeps, dtype=dt)
        Lf = orc.CSR(A.n_rows, A.n_cols, st.L.row, st.L.col, st.Lval)
        Uc = orc.CSR(A.n_rows, A.n_cols, st.U.row, st.U.col, st.Uval)  # U^T, diagonal last
        Ur = orc.transpose(Uc)
        want = orc.trsv(Ur, orc.trsv(Lf, r, upper=False, dtype=dt), upper=True, dtype=dt)
        got = P.apply(r, gpu.TRSV_LEVEL)
        assert np.array_equal(got, want), (sweeps, np.max(np.abs(got - want)))
        got_t = P.apply(r, gpu.TRSV_LEVEL, trans=1)
        want_t = orc.trsv(orc.transpose(Lf), orc.trsv(Uc, r, upper=False, dtype=dt), upper=True, dtype=dt)
        tol = 1e-12 if prec == "f64" else 1e-4
        assert np.allclose(got_t, want_t, rtol=tol, atol=tol * np.abs(want_t).max())
        _opt(gpu, "trsv_exact_tiles", 0)
        try:
            assert np.array_equal(P.apply(r, gpu.TRSV_LEVEL), got)
            assert np.array_equal(P.apply(r, gpu.TRSV_LEVEL, trans=1), got_t)
        finally:
            _opt(gpu, "trsv_exact_tiles", 1)
        _opt(gpu, "trsv_exact_rows", 256)
        try:
            assert np.array_equal(P.apply(r, gpu.TRSV_LEVEL), got)
        finally:
            _opt(gpu, "trsv_exact_rows", 512)


def test_exact_solve_box_order_solver(gpu, orc):
    """FGMRES with exact solves on box-ordered blocks takes the same steps
    and returns the same bits as with consecutive-row blocks."""
    A = orc.convdiff3d(24, 24, 24)
    b = gpu.gen_normal(3, A.n_rows)
    out = []
    for tiles in (1, 0):
        _opt(gpu, "trsv_exact_tiles", tiles)
        try:
            o = gpu.default_opts()
            o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = 30, 20, 1e-10, 3
            o.trsv, o.prec_precond = gpu.TRSV_LEVEL, gpu.LACE_F64
            S = gpu.Solver(gpu.HostCSR.from_csr(A), o)
            S.setup()
            x, st = S.solve(b)
            out.append((x, st.iterations, st.converged))
        finally:
            _opt(gpu, "trsv_exact_tiles", 1)
    (x1, it1, c1), (x0, it0, c0) = out
    assert c1 == 1 and c0 == 1 and it1 == it0
    assert np.array_equal(x1, x0)
    assert np.linalg.norm(b - A.scipy() @ x1) <= 1e-10 * np.linalg.norm(b)
