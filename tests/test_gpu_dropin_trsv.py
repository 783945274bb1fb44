"""The triangular-solve drop-ins through the C ABI: data_parcsrtrsv,
data_forward_solve and data_backward_solve (src/trisolve.cpp:20-119,
256-351) on CSRL / CSRU factors, against the oracle's restatement of the
reference loop run by one thread -- in-place sweeps in natural order (L) or
reverse order (U) until sum((y_new - y_old)^2) <= tol.  That loop
substitutes exactly in its first sweep and confirms in the second, so y is
compared bit for bit and iter exactly, for the tol the reference's FGMRES
passes (1.0, src/fgmres.cpp:129), a tol the first step already meets, an
initial guess equal to the solution, and nonzero initial guesses."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _factors(orc, A):
    """ParILU(3) factors in the drop-in formats: CSRL (unit diagonal last)
    and CSRU (diagonal first)."""
    Lf, Uf = orc.parilu(A, 3).factors()
    return Lf, Uf


def _parcsrtrsv(gpu, T, r, y0, tol, upper):
    y = np.array(y0, dtype=np.float64, copy=True)
    it = C.c_int(-7)
    row = np.ascontiguousarray(T.row, dtype=np.int32)
    col = np.ascontiguousarray(T.col, dtype=np.int32)
    val = np.ascontiguousarray(T.val, dtype=np.float64)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    uplo = gpu.MagmaUpper if upper else gpu.MagmaLower
    storage = gpu.Magma_CSRU if upper else gpu.Magma_CSRL
    diag = gpu.Magma_VALUE if upper else gpu.Magma_UNITY
    rc = gpu.lib().data_parcsrtrsv(uplo, storage, diag, T.n_rows, val.ctypes.data_as(dp),
                                   row.ctypes.data_as(ip), col.ctypes.data_as(ip),
                                   rr.ctypes.data_as(dp), y.ctypes.data_as(dp), tol, C.byref(it))
    return rc, y, it.value


CASES = ["laplace64", "convdiff16", "trefethen20"]


def _problem(orc, name):
    if name == "laplace64":
        return orc.laplace2d(64, 64)
    if name == "convdiff16":
        return orc.convdiff3d(16, 16, 16)
    return orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx"))


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("upper", [False, True])
def test_parcsrtrsv_matches_one_thread_reference(gpu, orc, name, upper):
    A = _problem(orc, name)
    Lf, Uf = _factors(orc, A)
    T = Uf if upper else Lf
    n = A.n_rows
    rng = np.random.default_rng(11)
    r = rng.standard_normal(n)
    orc.set_threads(1)
    try:
        exact = orc.trsv(T, r, upper)
        # the reference starts from step = 1e8: tol >= 1e8 runs no sweep
        for y0, tol in [(np.zeros(n), 1.0), (np.zeros(n), 1e7), (exact, 1.0),
                        (rng.standard_normal(n), 1e-3), (0.5 * exact, 1.0),
                        (rng.standard_normal(n), 1e8), (rng.standard_normal(n), 1e30)]:
            want, it_want = orc.parcsrtrsv(T, r, upper, tol, y0)
            rc, got, it = _parcsrtrsv(gpu, T, r, y0, tol, upper)
            assert rc == 0
            assert np.array_equal(got, want), (name, upper, tol)
            assert np.array_equal(got, exact if tol < 1e8 else y0)
            assert it == it_want, (it, it_want, tol)
    finally:
        orc.set_threads(0)


def test_parcsrtrsv_rejects_wrong_storage(gpu, orc):
    A = orc.laplace2d(8, 8)
    Lf, _ = _factors(orc, A)
    r = np.ones(A.n_rows)
    y = np.zeros(A.n_rows)
    it = C.c_int(0)
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    row = np.ascontiguousarray(Lf.row, dtype=np.int32)
    col = np.ascontiguousarray(Lf.col, dtype=np.int32)
    val = np.ascontiguousarray(Lf.val)
    rc = gpu.lib().data_parcsrtrsv(gpu.MagmaUpper, gpu.Magma_CSRL, gpu.Magma_UNITY, A.n_rows,
                                   val.ctypes.data_as(dp), row.ctypes.data_as(ip),
                                   col.ctypes.data_as(ip), r.ctypes.data_as(dp),
                                   y.ctypes.data_as(dp), 1.0, C.byref(it))
    assert rc == -1


def _dm(gpu, T, storage, fill, diag):
    H = gpu.HostCSR(T.row, T.col, T.val, storage=storage)
    H.m.fill_mode = fill
    H.m.diagorder_type = diag
    return H


@pytest.mark.parametrize("name", CASES)
def test_forward_backward_solve_wrappers(gpu, orc, name, capfd):
    """data_forward_solve / data_backward_solve: the L \\ U pair applied the
    way the reference's drivers do, with the per-sweep log lines."""
    A = _problem(orc, name)
    Lf, Uf = _factors(orc, A)
    n = A.n_rows
    r = np.random.default_rng(12).standard_normal(n)
    HL = _dm(gpu, Lf, gpu.Magma_CSRL, gpu.MagmaLower, gpu.Magma_UNITY)
    HU = _dm(gpu, Uf, gpu.Magma_CSRU, gpu.MagmaUpper, gpu.Magma_VALUE)
    rhs = gpu.HostVec(r)
    y = gpu.HostVec(np.zeros(n))
    z = gpu.HostVec(np.zeros(n))
    it = C.c_int(0)
    orc.set_threads(1)
    try:
        assert gpu.lib().data_forward_solve(HL.ref(), y.ref(), rhs.ref(), 1.0, C.byref(it)) == 0
        want_y, it_y = orc.parcsrtrsv(Lf, r, False, 1.0)
        assert np.array_equal(y.v, want_y) and it.value == it_y
        assert gpu.lib().data_backward_solve(HU.ref(), z.ref(), y.ref(), 1.0, C.byref(it)) == 0
        want_z, it_z = orc.parcsrtrsv(Uf, want_y, True, 1.0)
        assert np.array_equal(z.v, want_z) and it.value == it_z
        its = it_y + it_z
    finally:
        orc.set_threads(0)
    out = capfd.readouterr().out
    lines = [l for l in out.splitlines() if l.startswith("% iteration")]
    assert len(lines) == its
    assert lines[0].startswith("% iteration = 1 step = ")
    if it_y == 2:
        assert lines[1] == "% iteration = 2 step = 0.000000e+00"
    # the solution of the pair solves A ~ LU
    LU = orc.spmv(Lf, orc.spmv(Uf, z.v))
    assert np.allclose(LU, r, rtol=1e-10, atol=1e-10 * np.abs(r).max())
    # the wrong container is refused with info -1, as in the reference
    assert gpu.lib().data_forward_solve(HU.ref(), y.ref(), rhs.ref(), 1.0, C.byref(it)) == -1


def test_parcsrtrsv_config2_factor(gpu, orc):
    """The FP64 ParILU(3) factors of the 64^3 convection-diffusion operator
    (config 4's size): exact solves bit for bit with one-thread substitution."""
    A = orc.convdiff3d(64, 64, 64)
    Lf, Uf = _factors(orc, A)
    r = np.random.default_rng(13).standard_normal(A.n_rows)
    orc.set_threads(1)
    try:
        for T, upper in ((Lf, False), (Uf, True)):
            want, it_want = orc.parcsrtrsv(T, r, upper, 1.0)
            rc, got, it = _parcsrtrsv(gpu, T, r, np.zeros(A.n_rows), 1.0, upper)
            assert rc == 0 and it == it_want == 2
            assert np.array_equal(got, want)
    finally:
        orc.set_threads(0)
