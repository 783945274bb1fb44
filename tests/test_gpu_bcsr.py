"""Block (BCSR) path on the GPU against the oracle: block ParILU sweeps
(symmetric-pair kernels with and without LDS staging, per-block kernel, FP64/FP32 bit-exact, BF16/F16
storage bit-exact against the oracle's rounding), block triangular solves
(exact level-scheduled and Jacobi-k, FP32 bit-exact), block SpMV (FP64
bit-exact), the data_PariLU_v0_3_bcsr drop-in, and block-preconditioned
FGMRES to 1e-8 (config 3 at reduced and at full size)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _block7(orc, nx, ny, nz, bs, seed, shift=5.0):
    row, col = orc.gen_bcsr7_pattern(nx, ny, nz)
    rng = np.random.default_rng(seed)
    val = rng.normal(0.0, 0.1, size=(len(col), bs, bs))
    for I in range(len(row) - 1):
        for k in range(row[I], row[I + 1]):
            if col[k] == I:
                val[k] += shift * np.eye(bs)
    return orc.BCSR(len(row) - 1, bs, row, col, val)


def _bcsr_test(orc):
    return orc.bcsr_from_csr(orc.read_mtx(os.path.join(GOLDEN, "bcsr_test.mtx")), 5)


def _host(gpu, B):
    return gpu.HostBCSR(B.row, B.col, B.val, B.bs)


_PREC = {"f64": (0, np.float64, None), "f32": (1, np.float32, None),
         "bf16": (2, None, "bf16"), "f16": (3, None, "f16")}


def _compare(gpu, orc, B, prec, sweeps, variant=0):
    code, dt, storage = _PREC[prec]
    P = gpu.BILU(_host(gpu, B), code)
    if variant:
        P.set_variant(variant)
    P.sweeps(sweeps)
    (lr, lc, lv, lm), (ur, uc, uv, um) = P.factors()
    ref = orc.bilu(B, sweeps, "jacobi", dtype=dt or np.float32, storage=storage)
    Lf, Uf = ref.factors()
    assert np.array_equal(lr, Lf.row) and np.array_equal(lc, Lf.col)
    assert np.array_equal(ur, Uf.row) and np.array_equal(uc, Uf.col)
    assert lm["storage"] == gpu.Magma_BCSR and um["storage"] == gpu.Magma_BCSCU
    return lv, Lf.val, uv, Uf.val


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("bs", [5, 3, 1])
def test_sym_sweeps_bit_exact(gpu, orc, prec, bs):
    B = _block7(orc, 5, 4, 3, bs, seed=bs)
    for k in (1, 2, 4):
        lv, lw, uv, uw = _compare(gpu, orc, B, prec, k)
        assert np.array_equal(lv, lw), (prec, bs, k)
        assert np.array_equal(uv, uw), (prec, bs, k)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_general_kernel_bit_exact(gpu, orc, prec):
    """bcsr_test.mtx has an unpaired block (0,1): the per-block kernel runs.
    On a symmetric pattern the per-block kernel equals the paired one."""
    B = _bcsr_test(orc)
    P = gpu.BILU(_host(gpu, B), _PREC[prec][0])
    assert not P.symmetric()
    for k in (1, 2, 5):
        lv, lw, uv, uw = _compare(gpu, orc, B, prec, k)
        assert np.array_equal(lv, lw) and np.array_equal(uv, uw), k
    S = _block7(orc, 4, 3, 3, 5, seed=9)
    lv, lw, uv, uw = _compare(gpu, orc, S, prec, 3, variant=1)
    assert np.array_equal(lv, lw) and np.array_equal(uv, uw)


@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_16bit_storage_bit_exact(gpu, orc, prec):
    B = _block7(orc, 5, 4, 3, 5, seed=21)
    for k in (1, 3):
        lv, lw, uv, uw = _compare(gpu, orc, B, prec, k)
        assert np.array_equal(lv, lw), (prec, k)
        assert np.array_equal(uv, uw), (prec, k)
    lv, lw, uv, uw = _compare(gpu, orc, _bcsr_test(orc), prec, 2)
    assert np.array_equal(lv, lw) and np.array_equal(uv, uw)


def _block_sym_random(orc, nbr, bs, seed, most=6, shift=8.0):
    """Random structurally symmetric block pattern (each row its diagonal and
    up to `most` partners, each pair in both rows): unlike a grid, it holds
    triangles, so positions (I, J) have block columns K < J common to rows I
    and J (products in the sweep)."""
    rng = np.random.default_rng(seed)
    pairs = set()
    for i in range(nbr):
        for j in rng.choice(nbr, size=rng.integers(0, most + 1), replace=False):
            if j != i:
                pairs.add((i, int(j)))
                pairs.add((int(j), i))
    pairs |= {(i, i) for i in range(nbr)}
    ent = sorted(pairs)
    row = np.zeros(nbr + 1, np.int32)
    for i, _ in ent:
        row[i + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    col = np.array([j for _, j in ent], np.int32)
    val = rng.normal(0.0, 0.1, size=(len(col), bs, bs))
    for k, (i, j) in enumerate(ent):
        if i == j:
            val[k] += shift * np.eye(bs)
    return orc.BCSR(nbr, bs, row, col, val)


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16", "f16"])
@pytest.mark.parametrize("bs", [5, 4, 3, 2])
@pytest.mark.parametrize("pattern", ["grid", "random"])
def test_paired_sweep_forms_identical(gpu, orc, prec, bs, pattern):
    """The forms of the paired sweep: the device-side merge of the two block
    rows (variant 2) against the set-up tables without staging (3), with
    each workgroup's blocks staged through LDS wherever that applies (4),
    with only L, U and U_JJ staged (5), the wave-run form (6: whole-word
    runs per wave, columns of inv(U_JJ) split over a group's lanes), and
    the default choice (0): identical factors and step norms over many
    workgroups with a partial last one.  A 7-point grid has no products
    over K < J (no triangles); the random pattern has many."""
    if pattern == "grid":
        B = _block7(orc, 7, 6, 5, bs, seed=30 + bs)
    else:
        B = _block_sym_random(orc, 300, bs, seed=40 + bs)
    code = _PREC[prec][0]
    Ps = {v: gpu.BILU(_host(gpu, B), code) for v in (2, 3, 4, 5, 6, 0)}
    for v, P in Ps.items():
        assert P.symmetric()
        if v:
            P.set_variant(v)
    for k in range(3):
        steps = {v: P.sweeps(1, want_step=True) for v, P in Ps.items()}
        (_, _, l2, _), (_, _, u2, _) = Ps[2].factors()
        for v in (3, 4, 5, 6, 0):
            assert steps[v] == steps[2], (v, k)
            (_, _, lv, _), (_, _, uv, _) = Ps[v].factors()
            assert np.array_equal(lv, l2) and np.array_equal(uv, u2), (v, k)
    if pattern == "random" and prec in ("f64", "f32"):
        lv, lw, uv, uw = _compare(gpu, orc, B, prec, 3)
        assert np.array_equal(lv, lw) and np.array_equal(uv, uw)


def test_step_matches_oracle(gpu, orc):
    B = _block7(orc, 5, 4, 3, 5, seed=4)
    P = gpu.BILU(_host(gpu, B), gpu.LACE_F64)
    st = orc.BILUState(B)
    for k in range(4):
        s_gpu = P.sweeps(1, want_step=True)
        s_orc = st.sweep("jacobi")
        assert abs(s_gpu - s_orc) <= 1e-12 * max(s_orc, 1e-300), k


def _factors_for_solves(orc, B, sweeps):
    st = orc.bilu(B, sweeps, "jacobi", dtype=np.float32)
    Lf, Uf = st.factors()
    return Lf, orc.bcsr_transpose(Uf)


@pytest.mark.parametrize("case", ["sym", "general"])
def test_block_solves_bit_exact(gpu, orc, case):
    B = _block7(orc, 6, 5, 4, 5, seed=2) if case == "sym" else _bcsr_test(orc)
    P = gpu.BILU(_host(gpu, B), gpu.LACE_F32)
    P.sweeps(3)
    Lf, Ur = _factors_for_solves(orc, B, 3)
    r = np.random.default_rng(8).normal(size=B.nbr * B.bs).astype(np.float32)
    z = P.apply(r, gpu.TRSV_LEVEL)
    want = orc.btrsv(Ur, orc.btrsv(Lf, r, upper=False), upper=True)
    assert np.array_equal(z, want)
    for k in (1, 2, 5):
        z = P.apply(r, gpu.TRSV_JACOBI, sweeps=k)
        want = orc.btrsv(Ur, orc.btrsv(Lf, r, upper=False, jac=k), upper=True, jac=k)
        assert np.array_equal(z, want), k
    if case == "sym":
        nl, nu = P.levels()
        assert nl == nu == 6 + 5 + 4 - 2


def test_block_solves_f64_vs_dense(gpu, orc):
    B = _block7(orc, 4, 4, 3, 5, seed=6)
    P = gpu.BILU(_host(gpu, B), gpu.LACE_F64)
    P.sweeps(3)
    st = orc.bilu(B, 3, "jacobi")
    Lf, Uf = st.factors()
    M = Lf.dense() @ orc.bcsr_transpose(Uf).dense()
    r = np.random.default_rng(3).normal(size=B.nbr * B.bs)
    z = P.apply(r, gpu.TRSV_LEVEL)
    assert np.allclose(z, np.linalg.solve(M, r), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("bs", [5, 2])
def test_bcsr_spmv_bit_exact(gpu, orc, bs):
    B = _block7(orc, 7, 5, 4, bs, seed=bs + 30)
    x = np.random.default_rng(1).normal(size=B.nbr * bs)
    A = gpu.DevBCSR(_host(gpu, B), gpu.LACE_F64)
    assert np.array_equal(A.spmv(x), orc.bcsr_spmv(B, x))
    A32 = gpu.DevBCSR(_host(gpu, B), gpu.LACE_F32)
    assert np.allclose(A32.spmv(x), B.dense() @ x, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("case", ["sym", "general"])
def test_parilu_bcsr_dropin(gpu, orc, case):
    """data_PariLU_v0_3_bcsr: sweeps until sum((new-old)^2) <= reduction
    (absolute); factors equal the oracle's after the same number of
    synchronous sweeps; L comes back as BCSRL, U as BCSCU."""
    B = _block7(orc, 4, 4, 3, 5, seed=5) if case == "sym" else _bcsr_test(orc)
    H = _host(gpu, B)
    L = gpu.DMatrix(); U = gpu.DMatrix(); log = gpu.PrecondLog()
    rc = gpu.lib().data_PariLU_v0_3_bcsr(H.ref(), C.byref(L), C.byref(U), 1e-20, C.byref(log))
    assert rc == 0
    (lr, lc, lv, lm), (ur, uc, uv, um) = gpu.owned_bcsr_to_numpy(L), gpu.owned_bcsr_to_numpy(U)
    st = orc.BILUState(B)
    steps = [st.sweep("jacobi") for _ in range(log.sweeps)]
    assert steps[-1] <= 1e-20 and all(s > 1e-20 for s in steps[:-1])
    Lf, Uf = st.factors()
    assert np.array_equal(lv, Lf.val) and np.array_equal(uv, Uf.val)
    assert lm["storage"] == gpu.Magma_BCSR and um["storage"] == gpu.Magma_BCSCU


def _solve(gpu, B, prec, trsv, m=40, sweeps=3, rtol=1e-8, b=None):
    o = gpu.default_opts()
    o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = m, 50, rtol, sweeps
    o.prec_precond, o.trsv, o.jacobi_sweeps = prec, trsv, 5
    S = gpu.Solver(_host(gpu, B), o)
    S.setup()
    if b is None:
        b = gpu.gen_normal(2, S.n)
    x, st = S.solve(b)
    return x, st, b


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16", "f16"])
@pytest.mark.parametrize("trsv", [0, 1])
def test_block_fgmres_converges(gpu, orc, prec, trsv):
    """FGMRES(40) in FP64 with block ILU factors in the given storage to
    1e-8 on a 10^3 block grid (bs 5); the true FP64 residual is checked on
    the host against the oracle's block SpMV."""
    B = _block7(orc, 10, 10, 10, 5, seed=1)
    x, st, b = _solve(gpu, B, _PREC[prec][0], trsv)
    assert st.converged == 1
    r = b - orc.bcsr_spmv(B, x)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(b)
    assert st.iterations <= 40


def test_config3_full_size(gpu, orc):
    """Config 3: 64^3 block grid, 5x5 blocks (n = 1,310,720), off-diagonal
    entries 0.1 N(0,1) (seed 1), diagonal blocks + 5 I, b ~ N(0,1) seed 2;
    FP32 and BF16 factors under FP64 FGMRES(40) to 1e-8, the true residual
    recomputed on the host with the oracle's block product.  The error of
    each solution is stated against an FP64 solve to 1e-13."""
    import lace_amd as L
    row, col, val = L.gen_bcsr7(64, 64, 64, 5, seed=1, sigma=0.1, shift=5.0)
    B = orc.BCSR(len(row) - 1, 5, row, col, val)
    b = L.gen_normal(2, B.nbr * 5)
    xr, st, _ = _solve(gpu, B, gpu.LACE_F64, gpu.TRSV_JACOBI, rtol=1e-13, b=b)
    assert st.converged == 1
    assert np.linalg.norm(b - orc.bcsr_spmv(B, xr)) <= 1.01e-13 * np.linalg.norm(b)
    for prec in ("f32", "bf16"):
        x, st, _ = _solve(gpu, B, _PREC[prec][0], gpu.TRSV_JACOBI, b=b)
        assert st.converged == 1 and st.rel_true_residual <= 1.01e-8, prec
        # the FP64 true residual recomputed on the host with the oracle's
        # block product
        rel = np.linalg.norm(b - orc.bcsr_spmv(B, x)) / np.linalg.norm(b)
        assert rel <= 1e-8, (prec, rel)
        assert abs(rel - st.rel_true_residual) <= 1e-3 * rel
        err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        print(f"config3 {prec}: its={st.iterations} rel_res={st.rel_true_residual:.3e} "
              f"err_vs_fp64={err:.3e}")
        assert err < 1e-6
