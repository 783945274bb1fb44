"""Storage forms of the products and Jacobi sweeps.  The diagonal (DIA),
ELLPACK and CSR forms give bit-identical results (the oracle's, in FP64);
which forms a matrix gets follows its pattern.  The Jacobi sweeps on the
diagonal form run pipelined (all sweeps of a factor in one launch) or one
launch per sweep with 16 bytes of rows or one row per thread, also
bit-identical."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FORMS = {"dia": (1, 1), "ell": (0, 1), "csr": (0, 0)}


@pytest.fixture
def forms(gpu):
    def set_forms(kind, form):
        dia, ell = FORMS[form]
        gpu.check(gpu.lib().lace_set_option(f"{kind}_dia".encode(), dia))
        gpu.check(gpu.lib().lace_set_option(f"{kind}_ell".encode(), ell))
    yield set_forms
    for kind in ("spmv", "jacobi"):
        set_forms(kind, "dia")
    gpu.check(gpu.lib().lace_set_option(b"jacobi_fuse", 1))
    gpu.check(gpu.lib().lace_set_option(b"jacobi_dia_vec", 1))
    gpu.check(gpu.lib().lace_set_option(b"jacobi_pipe", 0))


def _holes(orc, A, frac, seed):
    """A with a random fraction of its off-diagonal entries removed (the
    remaining pattern still lies on A's diagonals, with gaps)."""
    rng = np.random.default_rng(seed)
    ri = A.rowidx()
    keep = (ri == A.col) | (rng.random(A.nnz) >= frac)
    row = np.zeros(A.n_rows + 1, dtype=np.int32)
    np.add.at(row, ri[keep] + 1, 1)
    return orc.CSR(A.n_rows, A.n_cols, np.cumsum(row).astype(np.int32), A.col[keep].copy(), A.val[keep].copy())


def problems(orc):
    c = orc.convdiff3d(16, 16, 16)
    return {
        "laplace64": orc.laplace2d(64, 64),
        "convdiff16": c,
        "convdiff16_holes": _holes(orc, c, 0.1, 5),
        "trefethen20": orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx")),
        # offsets 1, 15, 180: y windows at every shift within a 16-byte word
        "convdiff15x12x10": orc.convdiff3d(15, 12, 10),
        # n = 2730: not a multiple of 4 rows (FP32 keeps one row per thread)
        "convdiff15x14x13": orc.convdiff3d(15, 14, 13),
    }


NAMES = ["laplace64", "convdiff16", "convdiff16_holes", "trefethen20", "convdiff15x12x10",
         "convdiff15x14x13"]


def test_forms_follow_the_pattern(gpu, orc):
    P = problems(orc)
    got = {}
    for name in NAMES:
        H = gpu.HostCSR.from_csr(P[name])
        F = gpu.ParILU(H, gpu.LACE_F64)
        got[name] = (F.forms(False), F.forms(True), gpu.DevCSRMatrix(H).forms())
    # 2-D 5-point: L and U on 2 diagonals, A on 5; 3-D 7-point: 3 and 7
    assert got["laplace64"] == ((2, 2), (2, 2), (5, 5))
    assert got["convdiff16"] == ((3, 3), (3, 3), (7, 7))
    # gaps keep the diagonals (10 % padding < 25 %)
    assert got["convdiff16_holes"][0][0] == 3 and got["convdiff16_holes"][2][0] == 7
    # Trefethen: A on 11 diagonals (0, +-1, +-2, +-4, +-8, +-16); L's 5
    # diagonals would be 48 % padding
    assert got["trefethen20"][0][0] == 0 and got["trefethen20"][1][0] == 0
    assert got["trefethen20"][2][0] == 0


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_jacobi_forms_bitexact(gpu, orc, forms, name, prec):
    A = problems(orc)[name]
    lp = gpu.LACE_F64 if prec == "f64" else gpu.LACE_F32
    dt = np.float64 if prec == "f64" else np.float32
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), lp)
    P.sweeps(3)
    st = orc.parilu(A, 3, dtype=dt)
    Lf = orc.CSR(A.n_rows, A.n_cols, st.L.row, st.L.col, st.Lval)
    Ur = orc.transpose(orc.CSR(A.n_rows, A.n_cols, st.U.row, st.U.col, st.Uval))
    r = np.random.default_rng(11).standard_normal(A.n_rows).astype(dt)
    # odd sweep counts let U's first sweep ride on L's last (diagonal form)
    for sweeps in (5, 4, 3, 2):
        y = orc.trsv_jacobi(Lf, r, upper=False, sweeps=sweeps, dtype=dt)
        ref = orc.trsv_jacobi(Ur, y, upper=True, sweeps=sweeps, dtype=dt)
        for form in FORMS:
            forms("jacobi", form)
            for fuse in (1, 0):
                gpu.check(gpu.lib().lace_set_option(b"jacobi_fuse", fuse))
                # diagonal form: all sweeps of a factor in one pipelined
                # launch, or a launch per sweep with a 16-byte word of rows
                # per thread or one row
                for pipe, vec in (((1, 1), (0, 1), (0, 0)) if form == "dia" else ((0, 1),)):
                    gpu.check(gpu.lib().lace_set_option(b"jacobi_pipe", pipe))
                    gpu.check(gpu.lib().lace_set_option(b"jacobi_dia_vec", vec))
                    got = P.apply(r, gpu.TRSV_JACOBI, sweeps=sweeps)
                    assert np.array_equal(got, ref), (form, fuse, pipe, vec, sweeps)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_pipelined_jacobi_many_tiles(gpu, orc, forms, prec):
    """The pipelined sweeps with many tiles waiting on one another: 64^3
    (256 FP32 tiles of 1024 rows, 512 FP64 tiles; the z-offset 4096 reaches
    4 or 8 tiles back), against a launch per sweep, bit for bit, for
    several sweep counts and repeated applies (flags count up across
    launches); no wait gave up."""
    A = orc.convdiff3d(64, 64, 64)
    lp = gpu.LACE_F64 if prec == "f64" else gpu.LACE_F32
    dt = np.float64 if prec == "f64" else np.float32
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), lp)
    P.sweeps(3)
    rng = np.random.default_rng(17)
    for sweeps in (3, 5, 4, 7, 5):
        r = rng.standard_normal(A.n_rows).astype(dt)
        gpu.check(gpu.lib().lace_set_option(b"jacobi_pipe", 0))
        ref = P.apply(r, gpu.TRSV_JACOBI, sweeps=sweeps)
        gpu.check(gpu.lib().lace_set_option(b"jacobi_pipe", 1))
        for _ in range(2):
            assert np.array_equal(P.apply(r, gpu.TRSV_JACOBI, sweeps=sweeps), ref), sweeps
    err = C.c_int(-1)
    gpu.check(gpu.lib().lace_syncfree_error(P.h, C.byref(err)))
    assert err.value == 0


def _csr_order_sum(A, x, dt):
    """sum_k val[k] * x[col[k]] per row, in CSR order from zero, in dt."""
    n = A.n_rows
    lens = np.diff(A.row)
    s = np.zeros(n, dt)
    for q in range(int(lens.max())):
        rows = np.nonzero(lens > q)[0]
        k = A.row[rows] + q
        s[rows] = s[rows] + A.val[k].astype(dt) * x[A.col[k]].astype(dt)
    return s


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_spmv_forms_bitexact(gpu, orc, forms, name, prec):
    A = problems(orc)[name]
    lp = gpu.LACE_F64 if prec == "f64" else gpu.LACE_F32
    dt = np.float64 if prec == "f64" else np.float32
    M = gpu.DevCSRMatrix(gpu.HostCSR.from_csr(A), lp)
    rng = np.random.default_rng(13)
    x = rng.standard_normal(A.n_rows).astype(dt)
    y0 = rng.standard_normal(A.n_rows).astype(dt)
    ref = orc.spmv(A, x) if prec == "f64" else _csr_order_sum(A, x, np.float32)
    outs = {}
    for form in FORMS:
        forms("spmv", form)
        assert np.array_equal(M.spmv(x), ref), form
        outs[form] = M.spmv(x, alpha=2.0, beta=-0.5, y=y0)
    assert np.array_equal(outs["dia"], outs["ell"]) and np.array_equal(outs["dia"], outs["csr"])
    assert np.allclose(outs["dia"], 2.0 * ref - 0.5 * y0, rtol=1e-5 if prec == "f32" else 1e-14, atol=1e-5)


def test_solve_is_identical_across_forms(gpu, orc, forms):
    """Every product and sweep is bit-identical across forms, so a whole
    mixed-precision FGMRES solve is too."""
    A = orc.convdiff3d(24, 24, 24)
    b = np.random.default_rng(0).standard_normal(A.n_rows)
    o = gpu.default_opts()
    o.rtol = 1e-8
    o.parilu_sweeps = 3
    o.search_max = 10
    o.inner_iters = 5
    o.trsv = gpu.TRSV_JACOBI
    o.jacobi_sweeps = 5
    o.prec_precond = 1
    o.restart_max = 50
    xs, its = [], []
    for form in ("dia", "ell", "csr"):
        forms("spmv", form)
        forms("jacobi", form)
        S = gpu.Solver(gpu.HostCSR.from_csr(A), o)
        S.setup()
        x, st = S.solve(b)
        assert st.converged == 1 and st.rel_true_residual <= 1e-8
        xs.append(x)
        its.append(st.iterations)
    assert its[0] == its[1] == its[2]
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])
