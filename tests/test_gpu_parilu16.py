"""Scalar ParILU with 16-bit factor storage (BF16, FP16) on the GPU against
the oracle's model of the same rounding: FP32 sweeps, each factor value
rounded to 16 bits relative to the power-of-two scale of its row after
every sweep (L's unit diagonal exact).  Bar: bit-exact factors after every
sweep, bit-exact triangular solves on those factors (the Jacobi sweeps read
16-bit words), and FP64 solves to 1e-8 with the 16-bit preconditioner."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

STORE = {"bf16": 2, "f16": 3}  # LACE_BF16, LACE_F16


def _nonsym(orc):
    """7-point pattern with every third strictly upper entry beyond the
    first super-diagonal dropped: the general CSR-L / CSC-U kernel."""
    A = orc.convdiff3d(10, 9, 8)
    ri = np.repeat(np.arange(A.n_rows), np.diff(A.row))
    keep = np.ones(A.nnz, bool)
    keep[np.nonzero(A.col > ri + 1)[0][::3]] = False
    return orc._coo_to_csr(A.n_rows, ri[keep], A.col[keep], A.val[keep])


def _problem(orc, name):
    if name == "convdiff16":
        return orc.convdiff3d(16, 16, 16)
    if name == "scaled_rows":
        # rows of very different magnitude: the scales must follow the rows
        A = orc.convdiff3d(12, 11, 10)
        ri = np.repeat(np.arange(A.n_rows), np.diff(A.row))
        f = 2.0 ** (np.arange(A.n_rows) % 17 - 8)
        return orc.CSR(A.n_rows, A.n_cols, A.row, A.col, A.val * f[ri])
    if name == "trefethen20":
        return orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx"))
    return _nonsym(orc)


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("name", ["convdiff16", "scaled_rows", "trefethen20", "nonsym"])
@pytest.mark.parametrize("variant", [0, 1])
def test_parilu16_bitexact(gpu, orc, storage, name, variant):
    A = _problem(orc, name)
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), STORE[storage])
    if variant:
        gpu.check(gpu.lib().lace_parilu_set_variant(P.h, variant))
    st = orc.ParILUState(A, storage=storage)
    (_, _, lv), (_, _, uv) = P.factors()
    assert np.array_equal(lv.astype(np.float32), st.Lval) and np.array_equal(uv.astype(np.float32), st.Uval)
    anorm = np.linalg.norm(A.val)
    for k in range(1, 6):
        s = P.sweeps(1, want_step=True)
        step = st.sweep("jacobi")
        (_, _, lv), (_, _, uv) = P.factors()
        assert np.array_equal(lv.astype(np.float32), st.Lval), (k, "L")
        assert np.array_equal(uv.astype(np.float32), st.Uval), (k, "U")
        assert abs(s - step / anorm) <= 1e-6 * (step / anorm) + 1e-300


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("name", ["convdiff16", "scaled_rows", "nonsym"])
def test_parilu16_solves_bitexact(gpu, orc, storage, name):
    """Jacobi-5 (the diagonal form reads the 16-bit words on the stencil)
    and exact solves with the 16-bit factors equal the oracle's FP32 solves
    on the oracle's factors, and the Jacobi result is the same with the FP32
    copy of the diagonal form."""
    A = _problem(orc, name)
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), STORE[storage])
    P.sweeps(3)
    st = orc.parilu(A, 3, storage=storage)
    Lf, Uf = st.factors()
    r = np.random.default_rng(7).standard_normal(A.n_rows).astype(np.float32)
    y = orc.trsv_jacobi(Lf, r, False, 5, dtype=np.float32)
    want = orc.trsv_jacobi(Uf, y, True, 5, dtype=np.float32)
    got = P.apply(r, kind=gpu.TRSV_JACOBI, sweeps=5)
    assert np.array_equal(got, want)
    gpu.check(gpu.lib().lace_set_option(b"jacobi_dia16", 0))
    try:
        got32 = P.apply(r, kind=gpu.TRSV_JACOBI, sweeps=5)
    finally:
        gpu.check(gpu.lib().lace_set_option(b"jacobi_dia16", 1))
    assert np.array_equal(got32, want)
    ex = orc.trsv(Uf, orc.trsv(Lf, r, False, dtype=np.float32), True, dtype=np.float32)
    assert np.array_equal(P.apply(r, kind=gpu.TRSV_LEVEL), ex)
    if name == "convdiff16":
        d, _ = P.forms(False)
        assert d == 3  # the stencil's factors take the diagonal form


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_solver_with_16bit_factors(gpu, orc, storage):
    """FP64 FGMRES(10) x inner FP32 GMRES(10) with 16-bit ILU factors
    (Jacobi-5) reaches an FP64 true relative residual <= 1e-8."""
    A = orc.convdiff3d(24, 24, 24)
    o = gpu.default_opts()
    o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = 10, 100, 1e-8, 3
    o.trsv, o.jacobi_sweeps, o.inner_iters = gpu.TRSV_JACOBI, 5, 10
    o.prec_precond = STORE[storage]
    S = gpu.Solver(gpu.HostCSR.from_csr(A), o)
    S.setup()
    b = gpu.gen_normal(0, A.n_rows)
    x, st = S.solve(b)
    assert st.converged == 1
    assert np.linalg.norm(b - A.scipy() @ x) <= 1e-8 * np.linalg.norm(b)
