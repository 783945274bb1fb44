"""Pins of the oracle's block (BCSR) routines: block ParILU
(src/parilu_v0_3_bcsr.cpp:118-250), the block inverse, block SpMV and block
triangular solves, and the 16-bit rounding used for factor storage.

Checks chosen so a plausible slip fails one of them: a transposed block or
a swapped product order fails the dense sweep-1 closed form and the block
ILU(0) brute force; a wrong merge index fails the block-size-1 reduction to
the scalar sweep; a wrong inverse fails numpy.linalg.inv."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _rand_block7(orc, nx, ny, nz, bs, seed, shift=None):
    row, col = orc.gen_bcsr7_pattern(nx, ny, nz)
    rng = np.random.default_rng(seed)
    val = rng.normal(0.0, 0.1, size=(len(col), bs, bs))
    shift = 5.0 if shift is None else shift
    for I in range(len(row) - 1):
        for k in range(row[I], row[I + 1]):
            if col[k] == I:
                val[k] += shift * np.eye(bs)
    return orc.BCSR(len(row) - 1, bs, row, col, val)


def _block_ilu0_dense(B):
    """Block ILU(0) by the IKJ elimination restricted to the block pattern
    (dense numpy, independent of the sweep formulation)."""
    bs, nbr = B.bs, B.nbr
    blocks = {}
    for I in range(nbr):
        for k in range(B.row[I], B.row[I + 1]):
            blocks[(I, int(B.col[k]))] = B.val[k].copy()
    for I in range(nbr):
        cols = sorted(J for (i, J) in blocks if i == I)
        for K in [c for c in cols if c < I]:
            blocks[(I, K)] = blocks[(I, K)] @ np.linalg.inv(blocks[(K, K)])
            for J in cols:
                if J > K and (K, J) in blocks:
                    blocks[(I, J)] = blocks[(I, J)] - blocks[(I, K)] @ blocks[(K, J)]
    L = np.eye(nbr * bs); U = np.zeros((nbr * bs, nbr * bs))
    for (I, J), v in blocks.items():
        if I > J:
            L[I * bs:(I + 1) * bs, J * bs:(J + 1) * bs] = v
        else:
            U[I * bs:(I + 1) * bs, J * bs:(J + 1) * bs] = v
    return L, U


def _dense_factors(orc, st):
    Lf, Uf = st.factors()
    return Lf.dense(), orc.bcsr_transpose(Uf).dense()


def test_block_inverse_vs_numpy(orc):
    rng = np.random.default_rng(1)
    for bs in range(1, 9):
        for _ in range(5):
            M = rng.normal(size=(bs, bs)) + bs * np.eye(bs)
            assert np.allclose(orc.binv(M), np.linalg.inv(M), rtol=1e-13, atol=1e-14)


def test_bcsr_test_mtx_conversion(orc):
    """testing/matrices/bcsr_test.mtx: 15 x 15 = 3 x 3 blocks of 5 x 5."""
    A = orc.read_mtx(os.path.join(GOLDEN, "bcsr_test.mtx"))
    B = orc.bcsr_from_csr(A, 5)
    assert B.nbr == 3 and B.nblocks * 25 == 150
    assert np.array_equal(B.dense(), A.dense())
    # BCSRU = A's blocks with J >= I; BCSRL(unit) = blocks J < I plus I
    D = A.dense()
    blk = np.kron(np.arange(3)[:, None] <= np.arange(3)[None, :], np.ones((5, 5)))
    Ur = orc.bcsr_split(B, True, "value")
    assert np.array_equal(Ur.dense(), D * blk)
    Lr = orc.bcsr_split(B, False, "unit")
    assert np.array_equal(Lr.dense(), D * (1 - blk) + np.eye(15))
    # BCSCU: block columns, rows ascending with the diagonal last, block
    # interiors as stored (not transposed)
    U = orc.bcsr_transpose(Ur)
    for J in range(3):
        rows = U.col[U.row[J]:U.row[J + 1]]
        assert list(rows) == sorted(rows) and rows[-1] == J
        for k in range(U.row[J], U.row[J + 1]):
            I = U.col[k]
            assert np.array_equal(U.val[k], D[I * 5:(I + 1) * 5, J * 5:(J + 1) * 5])


@pytest.mark.parametrize("mode", ["jacobi", "gs"])
def test_bs1_reduces_to_scalar_parilu(orc, mode):
    """Block size 1: blocks are scalars, inv(u) = 1/u; the sweep equals the
    scalar sweep up to s*(1/u) vs s/u rounding."""
    A = orc.convdiff3d(6, 5, 4)
    B = orc.bcsr_from_csr(A, 1)
    sc = orc.ParILUState(A)
    bl = orc.BILUState(B)
    for _ in range(3):
        sc.sweep(mode)
        bl.sweep(mode)
    assert np.allclose(bl.Lval, sc.Lval, rtol=1e-14, atol=0)
    assert np.allclose(bl.Uval, sc.Uval, rtol=1e-14, atol=0)


def test_sweep1_closed_form(orc):
    """One Jacobi sweep from L0 = strict-lower(A) + I, U0 = upper(A):
    L_IJ = (A_IJ - sum_{K<J} A_IK A_KJ) inv(A_JJ), U_IJ = A_IJ - sum_{K<I}
    A_IK A_KJ, sums over blocks present in both."""
    B = _rand_block7(orc, 3, 3, 2, 3, seed=5)
    st = orc.BILUState(B)
    st.sweep("jacobi")
    L1, U1 = _dense_factors(orc, st)
    bs = B.bs
    blk = {(I, int(B.col[k])): B.val[k] for I in range(B.nbr) for k in range(B.row[I], B.row[I + 1])}
    for (I, J), a in blk.items():
        s = a.copy()
        for K in range(min(I, J)):
            if (I, K) in blk and (K, J) in blk:
                s = s - blk[(I, K)] @ blk[(K, J)]
        got = (L1 if I > J else U1)[I * bs:(I + 1) * bs, J * bs:(J + 1) * bs]
        want = s @ np.linalg.inv(blk[(J, J)]) if I > J else s
        assert np.allclose(got, want, rtol=1e-12, atol=1e-14), (I, J)


@pytest.mark.parametrize("case", ["bcsr_test", "block7"])
def test_fixed_point_is_block_ilu0(orc, case):
    """Synchronous sweeps reach the block ILU(0) factors (finite
    termination on the DAG of the pattern); one in-place (serial, the
    reference's order) sweep already gives them."""
    if case == "bcsr_test":
        B = orc.bcsr_from_csr(orc.read_mtx(os.path.join(GOLDEN, "bcsr_test.mtx")), 5)
    else:
        B = _rand_block7(orc, 4, 3, 3, 3, seed=7)
    Lw, Uw = _block_ilu0_dense(B)
    scale = np.abs(Uw).max()
    st = orc.bilu(B, 40, "jacobi")
    L, U = _dense_factors(orc, st)
    assert np.allclose(L, Lw, rtol=1e-10, atol=1e-12 * np.abs(Lw).max())
    assert np.allclose(U, Uw, rtol=1e-10, atol=1e-12 * scale)
    gs = orc.bilu(B, 1, "gs")
    L, U = _dense_factors(orc, gs)
    assert np.allclose(L, Lw, rtol=1e-10, atol=1e-12 * np.abs(Lw).max())
    assert np.allclose(U, Uw, rtol=1e-10, atol=1e-12 * scale)
    # invariant of ILU(0): (LU)_IJ = A_IJ on A's block pattern
    D = B.dense(); LU = L @ U
    mask = D != 0
    assert np.allclose(LU[mask], D[mask], rtol=1e-9, atol=1e-11 * np.abs(D).max())


def test_bcsr_spmv_vs_dense(orc):
    B = _rand_block7(orc, 4, 3, 2, 5, seed=3)
    x = np.random.default_rng(4).normal(size=B.nbr * B.bs)
    assert np.allclose(orc.bcsr_spmv(B, x), B.dense() @ x, rtol=1e-13, atol=1e-13)


def test_block_triangular_solves(orc):
    B = _rand_block7(orc, 4, 3, 3, 5, seed=11)
    st = orc.bilu(B, 3, "jacobi", dtype=np.float32)
    Lf, Uf = st.factors()
    Ur = orc.bcsr_transpose(Uf)  # U by block rows, diagonal block first
    r = np.random.default_rng(2).normal(size=B.nbr * B.bs).astype(np.float32)
    Ld, Ud = Lf.dense(), Ur.dense()
    y = orc.btrsv(Lf, r, upper=False)
    assert np.allclose(y, np.linalg.solve(Ld, r.astype(np.float64)), rtol=1e-4, atol=1e-5)
    z = orc.btrsv(Ur, y, upper=True)
    assert np.allclose(z, np.linalg.solve(Ud, y.astype(np.float64)), rtol=1e-4, atol=1e-5)
    # Jacobi-k closed form: y_k = sum_{j<k} (-D^{-1} N)^j D^{-1} r
    for T, D, upper in ((Lf, Ld, False), (Ur, Ud, True)):
        bs = B.bs
        Dg = np.zeros_like(D)
        for I in range(B.nbr):
            Dg[I * bs:(I + 1) * bs, I * bs:(I + 1) * bs] = D[I * bs:(I + 1) * bs, I * bs:(I + 1) * bs]
        Di = np.linalg.inv(Dg)
        G = -Di @ (D - Dg)
        for k in (1, 2, 5):
            want = sum(np.linalg.matrix_power(G, j) @ Di @ r.astype(np.float64) for j in range(k))
            got = orc.btrsv(T, r, upper=upper, jac=k)
            assert np.allclose(got, want, rtol=1e-4, atol=1e-5), (upper, k)


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    import oracle as orc
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        # exact ties: 1 + 2^-8 (to even 1.0), 1 + 3*2^-8 (up)
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0]])
    x = x.astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = orc.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got[-5] == 1.0 and got[-4] == np.float32(1.0 + 2.0 ** -6)


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_16bit_storage_tracks_fp32(orc, storage):
    """16-bit factor storage with row scales: after the sweeps the stored
    factors sit within the format's rounding of the FP32 ones."""
    B = _rand_block7(orc, 4, 4, 3, 5, seed=13)
    ref = orc.bilu(B, 5, "jacobi", dtype=np.float32)
    q = orc.bilu(B, 5, "jacobi", storage=storage)
    eps = 2.0 ** -8 if storage == "bf16" else 2.0 ** -11
    Lr, Ur = ref.factors(); Lq, Uq = q.factors()
    for a, b in ((Lr.val, Lq.val), (Ur.val, Uq.val)):
        err = np.abs(a - b).max() / np.abs(a).max()
        assert err < 20 * eps, err
    # every stored entry is a 16-bit number times its row scale
    rows = q._urows()
    m = (q.Uval / q.rs[rows]).astype(np.float32)
    if storage == "bf16":
        assert np.array_equal(orc.bf16_round(m), m)
    else:
        assert np.array_equal(m.astype(np.float16).astype(np.float32), m)
