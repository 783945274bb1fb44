"""The allocation-free lace_parcsrtrsv_noalloc (device arrays, scratch in a
fixed per-GPU array) against data_parcsrtrsv (host arrays) through the C
ABI: the same y bit for bit and the same sweep count, on two calls in a row
-- the second on a smaller system, which finds the first call's contents in
the shared scratch -- for a lower and an upper factor; sizes beyond the
scratch's maximum are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_parcsrtrsv(gpu, T, r, y0, tol, upper):
    y = np.array(y0, dtype=np.float64, copy=True)
    it = C.c_int(-7)
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    row = np.ascontiguousarray(T.row, dtype=np.int32)
    col = np.ascontiguousarray(T.col, dtype=np.int32)
    val = np.ascontiguousarray(T.val, dtype=np.float64)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    rc = gpu.lib().data_parcsrtrsv(gpu.MagmaUpper if upper else gpu.MagmaLower,
                                   gpu.Magma_CSRU if upper else gpu.Magma_CSRL,
                                   gpu.Magma_VALUE if upper else gpu.Magma_UNITY, T.n_rows,
                                   val.ctypes.data_as(dp), row.ctypes.data_as(ip),
                                   col.ctypes.data_as(ip), rr.ctypes.data_as(dp),
                                   y.ctypes.data_as(dp), tol, C.byref(it))
    assert rc == 0
    return y, it.value


def _noalloc(gpu, T, r, y0, tol, upper):
    n = T.n_rows
    dval = gpu.DevVec(T.nnz, np.float64, T.val)
    drow = gpu.DevVec(n + 1, np.int32, T.row)
    dcol = gpu.DevVec(T.nnz, np.int32, T.col)
    dr = gpu.DevVec(n, np.float64, r)
    dy = gpu.DevVec(n, np.float64, y0)
    dit = gpu.DevVec(1, np.int32, np.array([-7], np.int32))
    rc = gpu.lib().lace_parcsrtrsv_noalloc(gpu.MagmaUpper if upper else gpu.MagmaLower,
                                           gpu.Magma_CSRU if upper else gpu.Magma_CSRL,
                                           gpu.Magma_VALUE if upper else gpu.Magma_UNITY, n,
                                           dval.ptr, drow.ptr, dcol.ptr, dr.ptr, dy.ptr, tol, dit.ptr)
    assert rc == 0, gpu.lib().lace_last_error()
    gpu.check(gpu.lib().lace_device_synchronize())
    return dy.download(), int(dit.download()[0])


@pytest.mark.parametrize("upper", [False, True])
def test_noalloc_matches_dropin_two_calls(gpu, orc, upper):
    print(f"noalloc: upper={upper}", flush=True)
    rng = np.random.default_rng(31)
    for N, tol, y0kind in ((40, 1.0, "zero"), (24, 1e-3, "random")):
        A = orc.convdiff3d(N, N, N)
        Lf, Uf = orc.parilu(A, 3).factors()
        T = Uf if upper else Lf
        n = A.n_rows
        r = rng.standard_normal(n)
        y0 = np.zeros(n) if y0kind == "zero" else rng.standard_normal(n)
        want, it_want = _host_parcsrtrsv(gpu, T, r, y0, tol, upper)
        got, it = _noalloc(gpu, T, r, y0, tol, upper)
        print(f"noalloc: N={N} n={n} iter={it} (drop-in {it_want})", flush=True)
        assert np.array_equal(got, want), (N, upper)
        assert it == it_want == 2


def test_noalloc_tol_rules(gpu, orc):
    A = orc.convdiff3d(10, 10, 10)
    Lf, _ = orc.parilu(A, 3).factors()
    n = A.n_rows
    r = np.random.default_rng(32).standard_normal(n)
    exact, _ = _host_parcsrtrsv(gpu, Lf, r, np.zeros(n), 1.0, False)
    # y0 already the solution: step 0 <= tol, one sweep
    got, it = _noalloc(gpu, Lf, r, exact, 1.0, False)
    assert it == 1 and np.array_equal(got, exact)
    # tol >= 1e8 (the loop's initial step): no sweep, y untouched
    y0 = np.full(n, 3.0)
    got, it = _noalloc(gpu, Lf, r, y0, 1e9, False)
    assert it == 0 and np.array_equal(got, y0)


def test_noalloc_refuses_oversize_and_bad_storage(gpu):
    lib = gpu.lib()
    big = 4194304 + 1
    rc = lib.lace_parcsrtrsv_noalloc(gpu.MagmaLower, gpu.Magma_CSRL, gpu.Magma_UNITY, big, None, None,
                                     None, None, None, 1.0, None)
    assert rc == -104
    assert b"at most 4194304" in lib.lace_last_error()
    rc = lib.lace_parcsrtrsv_noalloc(gpu.MagmaUpper, gpu.Magma_CSRL, gpu.Magma_UNITY, 10, None, None,
                                     None, None, None, 1.0, None)
    assert rc == -1
