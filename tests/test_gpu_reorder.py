"""Reverse Cuthill-McKee as a pre-step of the solve (data_sparse_reorder,
src/reorder.cpp:167-253, with the textbook ordering lace_reorder_rcm,
reorder = 4): the permuted system solves to the same tolerance and maps back
to the solution of the original one; the ordering changes the factors' level
counts and the storage forms.  (The reference's own ordering, reorder = 2,
makes the synchronous ParILU sweeps overflow on this operator; the CPU test
test_reorder.py::test_reference_ordering_breaks_synchronous_sweeps shows
that on the oracle.)"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rcm(gpu, orc, A, reorder=4):
    H = gpu.HostCSR(A.row, A.col, A.val)
    M = gpu.DMatrix()
    gpu.check(gpu.lib().data_zmconvert(H.m, C.byref(M), gpu.Magma_CSR, gpu.Magma_CSR))
    n = A.n_rows
    P = np.zeros(n, np.int32)
    Pinv = np.zeros(n, np.int32)
    ip = C.POINTER(C.c_int)
    gpu.check(gpu.lib().data_sparse_reorder(C.byref(M), P.ctypes.data_as(ip), Pinv.ctypes.data_as(ip), reorder))
    row, col, val = gpu.owned_to_numpy(M)
    return orc.CSR(n, n, row, col, val), P


def test_rcm_prestep_solve(gpu, orc):
    A = orc.convdiff3d(16, 14, 12)
    n = A.n_rows
    Ar, P = _rcm(gpu, orc, A)
    assert np.array_equal(Ar.dense(), A.dense()[np.ix_(P, P)])
    b = gpu.gen_normal(0, n)
    o = gpu.default_opts()
    o.rtol, o.parilu_sweeps, o.search_max, o.inner_iters = 1e-8, 3, 10, 10
    o.trsv, o.jacobi_sweeps, o.prec_precond, o.restart_max = gpu.TRSV_JACOBI, 5, 1, 100
    out = {}
    for name, M, rhs in (("natural", A, b), ("rcm", Ar, b[P])):
        H = gpu.HostCSR.from_csr(M)
        S = gpu.Solver(H, o)
        S.setup()
        x, st = S.solve(rhs)
        assert st.converged == 1 and st.rel_true_residual <= 1e-8
        F = gpu.ParILU(H, gpu.LACE_F32)
        out[name] = (x, F.levels(), gpu.DevCSRMatrix(H).forms(), F.forms(False))
    # y = P x solves the permuted system, so x = y mapped back solves A x = b
    x = np.empty(n)
    x[P] = out["rcm"][0]
    assert np.linalg.norm(b - A.scipy() @ x) <= 1.01e-8 * np.linalg.norm(b)
    # natural order: levels nx + ny + nz - 2 per factor, the diagonal forms
    assert out["natural"][1] == (40, 40)
    assert out["natural"][2][0] == 7 and out["natural"][3][0] == 3
    # RCM numbers the grid by breadth-first fronts from a corner: the 40
    # fronts are the levels again, and the entries spread over many
    # diagonals (no diagonal form)
    assert out["rcm"][1] == (40, 40) and out["rcm"][2][0] == 0
