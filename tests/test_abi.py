"""The C-ABI library on the CPU: it loads, exports every symbol that
include/lace.h declares, the header is plain C with the reference's
container layout, and the host-side routines (reader, format conversions,
merge, residual check, exact ILU(0), level sets, generators) agree with the
oracle.  No device calls."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "lace.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\(", src)
    skip = {"if", "sizeof", "__attribute__", "aligned", "visibility", "push", "pop", "defined"}
    # function declarations: identifier followed by '(' at statement level
    decls = set()
    for m in re.finditer(r"[\w\*\s]+?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src):
        decls.add(m.group(1))
    return sorted(d for d in decls if d not in skip)


def test_library_exports_every_declared_symbol(lace):
    names = declared_functions()
    assert len(names) > 60
    out = subprocess.run(["nm", "-D", "--defined-only", lace.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and every binding in the Python module resolves
    for n in lace.SIGNATURES:
        assert hasattr(lace.lib(), n)


def test_header_is_plain_c_with_reference_layout(tmp_path):
    prog = tmp_path / "probe.c"
    prog.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "lace.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(data_d_matrix),
         offsetof(data_d_matrix, val), offsetof(data_d_matrix, row),
         offsetof(data_d_matrix, col), offsetof(data_d_matrix, major),
         sizeof(data_d_preconditioner_log), sizeof(data_d_gmres_log),
         offsetof(data_d_gmres_param, parilu_reduction));
  return 0;
}
''')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER),
                    str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    # reference include/sparse_types.h: pointer members aligned to 64 bytes
    assert [int(v) for v in out] == [384, 64, 192, 320, 360, 128, 64, 40]


def _read(lace, path):
    A = lace.DMatrix()
    lace.check(lace.lib().data_z_csr_mtx(C.byref(A), path.encode()))
    return lace.owned_to_numpy(A)


@pytest.mark.parametrize("name", ["io_test.mtx", "Trefethen_20.mtx", "bcsr_test.mtx"])
def test_reader_matches_oracle(lace, orc, name):
    row, col, val = _read(lace, os.path.join(GOLDEN, name))
    ref = orc.read_mtx(os.path.join(GOLDEN, name))
    assert np.array_equal(row, ref.row) and np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.val)


def test_writer_round_trip(lace, orc, tmp_path):
    A = orc.convdiff3d(4, 3, 2)
    H = lace.HostCSR.from_csr(A)
    p = str(tmp_path / "a.mtx")
    lace.check(lace.lib().data_zwrite_csr_mtx(H.m, lace.MagmaRowMajor, p.encode()))
    row, col, val = _read(lace, p)
    assert np.array_equal(row, A.row) and np.array_equal(col, A.col) and np.array_equal(val, A.val)
    first = open(p).read().splitlines()[:2]
    assert first[0] == "%%MatrixMarket matrix coordinate real general"
    assert first[1] == f"{A.n_rows} {A.n_cols} {A.nnz}"


def _convert(lace, A, old, new, diag=None):
    H = lace.HostCSR.from_csr(A)
    H.m.storage_type = old
    B = lace.DMatrix()
    if diag is not None:
        B.diagorder_type = diag
    lace.check(lace.lib().data_zmconvert(H.m, C.byref(B), old, new), "zmconvert")
    return lace.owned_to_numpy(B)


@pytest.mark.parametrize("name", ["io_test.mtx", "Trefethen_20.mtx"])
def test_conversions_match_oracle(lace, orc, name):
    A = orc.read_mtx(os.path.join(GOLDEN, name))
    L = _convert(lace, A, lace.Magma_CSR, lace.Magma_CSRL, lace.Magma_UNITY)
    Lr = orc.csr_split(A, False, orc.DIAG_UNIT)
    assert all(np.array_equal(a, b) for a, b in zip(L, (Lr.row, Lr.col, Lr.val)))
    U = _convert(lace, A, lace.Magma_CSR, lace.Magma_CSRU, lace.Magma_VALUE)
    Ur = orc.csr_split(A, True, orc.DIAG_VALUE)
    assert all(np.array_equal(a, b) for a, b in zip(U, (Ur.row, Ur.col, Ur.val)))
    Uc = _convert(lace, A, lace.Magma_CSR, lace.Magma_CSCU, lace.Magma_VALUE)
    Ucr = orc.transpose(Ur)
    assert all(np.array_equal(a, b) for a, b in zip(Uc, (Ucr.row, Ucr.col, Ucr.val)))
    T = _convert(lace, A, lace.Magma_CSR, lace.Magma_CSC)
    Tr = orc.transpose(A)
    assert all(np.array_equal(a, b) for a, b in zip(T, (Tr.row, Tr.col, Tr.val)))
    # CSC -> CSR undoes the transpose (testing/test_linearSolver.cpp:670)
    back = _convert(lace, orc.CSR(A.n_rows, A.n_cols, *T), lace.Magma_CSC, lace.Magma_CSR)
    assert all(np.array_equal(a, b) for a, b in zip(back, (A.row, A.col, A.val)))


def test_lumerge_matches_split(lace, orc):
    A = orc.convdiff3d(5, 4, 3)
    Lr = orc.csr_split(A, False, orc.DIAG_UNIT)
    Ur = orc.csr_split(A, True, orc.DIAG_VALUE)
    Lh = lace.HostCSR.from_csr(Lr); Lh.m.storage_type = lace.Magma_CSR
    Uh = lace.HostCSR.from_csr(Ur); Uh.m.storage_type = lace.Magma_CSR
    M = lace.DMatrix()
    lace.check(lace.lib().data_zmlumerge(Lh.m, Uh.m, C.byref(M)))
    row, col, val = lace.owned_to_numpy(M)
    # strict-L then U in A's own pattern and order
    assert np.array_equal(row, A.row) and np.array_equal(col, A.col)
    assert np.array_equal(val, A.val)


def test_ilures_and_spgemm_match_oracle(lace, orc):
    A = orc.convdiff3d(6, 5, 4)
    st = orc.parilu(A, 2)
    Lf, Uf = st.factors()
    Lh = lace.HostCSR.from_csr(Lf); Lh.m.storage_type = lace.Magma_CSRL
    Uh = lace.HostCSR.from_csr(Uf); Uh.m.storage_type = lace.Magma_CSRU
    LU = lace.DMatrix()
    r = C.c_double(); nl = C.c_double()
    Ah = lace.HostCSR.from_csr(A)  # keep the arrays alive across the call
    lace.check(lace.lib().data_zilures(Ah.m, Lh.m, Uh.m, C.byref(LU), C.byref(r), C.byref(nl)))
    lace.owned_to_numpy(LU)
    rr, nlr = orc.ilures(A, Lf, Uf)
    assert abs(r.value - rr) <= 1e-13 * rr and abs(nl.value - nlr) <= 1e-13 * nlr
    Cm = lace.DMatrix()
    lace.check(lace.lib().data_z_spmm(1.0, Lh.m, Uh.m, C.byref(Cm)))
    cr, cc, cv = lace.owned_to_numpy(Cm)
    ref = orc.spgemm(Lf, Uf)
    assert np.array_equal(cr, ref.row) and np.array_equal(cc, ref.col)
    assert np.allclose(cv, ref.val, rtol=1e-15, atol=1e-15)


def test_exact_ilu0_matches_oracle(lace, orc):
    A = orc.convdiff3d(6, 6, 5)
    H = lace.HostCSR.from_csr(A)
    lace.check(lace.lib().data_dcsrilu0_mkl(H.ref()))
    ref, bad = orc.ilu0(A)
    assert bad < 0
    assert np.array_equal(H.val, ref.val)  # same IKJ operation order: bitwise


def test_level_sets_match_oracle(lace, orc):
    A = orc.laplace2d(11, 9)
    for upper in (False, True):
        T = orc.csr_split(A, upper, orc.DIAG_VALUE if upper else orc.DIAG_UNIT)
        n = T.n_rows
        perm = np.zeros(n, np.int32); lptr = np.zeros(n + 1, np.int32)
        nl = lace.lib().lace_level_sets(n, T.row.ctypes.data_as(C.POINTER(C.c_int)),
                                        T.col.ctypes.data_as(C.POINTER(C.c_int)), int(upper),
                                        perm.ctypes.data_as(C.POINTER(C.c_int)),
                                        lptr.ctypes.data_as(C.POINTER(C.c_int)))
        lev, nlr = orc.levels(T, upper)
        assert nl == nlr == 11 + 9 - 1
        # rows sorted by level, ties by index
        assert np.array_equal(perm, np.lexsort((np.arange(n), lev)))
        assert np.array_equal(lptr[: nl + 1], np.concatenate([[0], np.cumsum(np.bincount(lev))]))


def test_generators_match_oracle(lace, orc):
    for dims in ((5, 4, 3), (16, 16, 16)):
        r, c, v = lace.gen_convdiff3d(*dims)
        ref = orc.convdiff3d(*dims)
        assert np.array_equal(r, ref.row) and np.array_equal(c, ref.col)
        assert np.array_equal(v, ref.val)
    r, c, v = lace.gen_laplace2d(64, 64)
    ref = orc.laplace2d(64, 64)
    assert np.array_equal(r, ref.row) and np.array_equal(c, ref.col) and np.array_equal(v, ref.val)
    assert r[-1] == 20224  # config 1 nnz


def test_config2_sizes(lace):
    r, c, v = lace.gen_convdiff3d(128, 128, 128)
    assert len(r) - 1 == 2097152 and r[-1] == 14581760
    d = v[np.arange(len(r) - 1)[:, None] == 0] if False else None  # noqa
    # row sums: interior rows sum to 0 (h-scaled upwind stencil is conservative)
    i = 64 + 128 * 64 + 16384 * 64
    assert abs(v[r[i]:r[i + 1]].sum()) < 1e-14


def test_gen_normal_philox_box_muller(lace, orc):
    """Sample 2t / 2t+1 = Box-Muller pair from Philox4x32-10(counter t,
    key = seed): two 53-bit uniforms in (0, 1]."""
    seed = 12345
    out = lace.gen_normal(seed, 11)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for t in range(6):
        w = [int(x) for x in orc.philox4x32_10([t, 0, 0, 0], key)]
        a = ((w[0] << 21) ^ (w[1] >> 11)) & ((1 << 53) - 1)
        b = ((w[2] << 21) ^ (w[3] >> 11)) & ((1 << 53) - 1)
        u1 = (a + 1) / 2.0 ** 53
        u2 = (b + 1) / 2.0 ** 53
        rad = np.sqrt(-2 * np.log(u1))
        assert abs(out[2 * t] - rad * np.cos(2 * np.pi * u2)) <= 1e-14 * max(1, abs(out[2 * t]))
        if 2 * t + 1 < 11:
            assert abs(out[2 * t + 1] - rad * np.sin(2 * np.pi * u2)) <= 1e-14 * max(1, abs(out[2 * t + 1]))
    big = lace.gen_normal(0, 200000)
    assert abs(big.mean()) < 0.01 and abs(big.std() - 1) < 0.01


def test_givens_dropin(lace, orc):
    c = C.c_double(); s = C.c_double()
    for a, b in [(3.0, 4.0), (-1.0, 2.5), (0.0, 1.0), (2.0, 0.0)]:
        lace.lib().givens_rotation(a, b, 4.9e-324, C.byref(c), C.byref(s))
        assert (c.value, s.value) == orc.givens(a, b, 4.9e-324)


def test_dnrm2_and_frobenius(lace, orc):
    x = np.random.default_rng(0).standard_normal(100)
    assert abs(lace.lib().data_dnrm2(100, x.ctypes.data_as(C.POINTER(C.c_double)), 1)
               - np.linalg.norm(x)) < 1e-13
    A = orc.convdiff3d(4, 4, 4)
    f = C.c_double()
    Ah = lace.HostCSR.from_csr(A)
    lace.check(lace.lib().data_zfrobenius(Ah.m, C.byref(f)))
    assert abs(f.value - np.linalg.norm(A.val)) < 1e-12


def test_device_calls_fail_cleanly_without_gpu(lace):
    if lace.gpu_available():
        pytest.skip("a GPU is present")
    p = C.c_void_p()
    rc = lace.lib().lace_dmalloc(C.byref(p), 16)
    assert rc != 0 and b"device" in lace.lib().lace_last_error()


# ------------------------------------------------------------ block formats
def _bcsr_equal(got, ref):
    row, col, val = got[:3]
    return (np.array_equal(row, ref.row) and np.array_equal(col, ref.col)
            and np.array_equal(val.reshape(ref.val.shape), ref.val))


def test_gen_bcsr7_matches_oracle_pattern(lace, orc):
    """7-point block pattern (Dirichlet, x fastest, columns ascending); the
    values are sigma * gen_normal(seed) in storage order plus the shift on
    the diagonal entries of the diagonal blocks."""
    nx, ny, nz, bs = 5, 4, 3, 5
    row, col, val = lace.gen_bcsr7(nx, ny, nz, bs, seed=1, sigma=0.1, shift=5.0)
    prow, pcol = orc.gen_bcsr7_pattern(nx, ny, nz)
    assert np.array_equal(row, prow) and np.array_equal(col, pcol)
    want = 0.1 * lace.gen_normal(1, len(col) * bs * bs).reshape(-1, bs, bs)
    for I in range(len(row) - 1):
        for k in range(row[I], row[I + 1]):
            if col[k] == I:
                want[k] += 5.0 * np.eye(bs)
    assert np.array_equal(val, want)


def test_config3_block_counts(lace, orc):
    """64^3 7-point block grid: 1,810,432 blocks (Dirichlet), 5x5 blocks."""
    prow, _ = orc.gen_bcsr7_pattern(64, 64, 64)
    assert prow[-1] == 1810432 and len(prow) - 1 == 262144


def _to_bcsr(lace, A, bs):
    H = lace.HostCSR(A.row, A.col, A.val, A.n_cols)
    B = lace.DMatrix()
    B.blocksize = bs
    lace.check(lace.lib().data_zmconvert(H.m, C.byref(B), lace.Magma_CSR, lace.Magma_BCSR))
    return lace.owned_bcsr_to_numpy(B)


@pytest.mark.parametrize("name,bs", [("bcsr_test.mtx", 5), ("io_test.mtx", 2), ("Trefethen_20.mtx", 3)])
def test_csr_to_bcsr_matches_oracle(lace, orc, name, bs):
    A = orc.read_mtx(os.path.join(GOLDEN, name))
    got = _to_bcsr(lace, A, bs)
    assert got[3]["storage"] == lace.Magma_BCSR and got[3]["bs"] == bs
    assert _bcsr_equal(got, orc.bcsr_from_csr(A, bs))


def test_csr_to_bcsr_needs_blocksize(lace, orc):
    A = orc.read_mtx(os.path.join(GOLDEN, "io_test.mtx"))
    H = lace.HostCSR(A.row, A.col, A.val, A.n_cols)
    B = lace.DMatrix()
    rc = lace.lib().data_zmconvert(H.m, C.byref(B), lace.Magma_CSR, lace.Magma_BCSR)
    assert rc != 0 and b"blocksize" in lace.lib().lace_last_error()


def test_bcsr_splits_and_block_transpose(lace, orc):
    """BCSR -> BCSRL (unit) / BCSRU / BCSCU and data_zmtranspose on a BCSR
    matrix against the oracle (blocks move, interiors stay)."""
    A = orc.read_mtx(os.path.join(GOLDEN, "bcsr_test.mtx"))
    Bo = orc.bcsr_from_csr(A, 5)
    H = lace.HostBCSR(Bo.row, Bo.col, Bo.val, 5)

    def conv(new, diag):
        M = lace.DMatrix()
        M.diagorder_type = diag
        lace.check(lace.lib().data_zmconvert(H.m, C.byref(M), lace.Magma_BCSR, new))
        return lace.owned_bcsr_to_numpy(M)

    L = conv(lace.Magma_BCSRL, lace.Magma_UNITY)
    assert L[3]["storage"] == lace.Magma_BCSR  # the reference's tag for BCSRL
    assert _bcsr_equal(L, orc.bcsr_split(Bo, False, "unit"))
    U = conv(lace.Magma_BCSRU, lace.Magma_VALUE)
    assert _bcsr_equal(U, orc.bcsr_split(Bo, True, "value"))
    Uc = conv(lace.Magma_BCSCU, lace.Magma_VALUE)
    assert Uc[3]["storage"] == lace.Magma_BCSCU and Uc[3]["major"] == lace.MagmaColMajor
    assert _bcsr_equal(Uc, orc.bcsr_transpose(orc.bcsr_split(Bo, True, "value")))
    T = lace.DMatrix()
    lace.check(lace.lib().data_zmtranspose(H.m, C.byref(T)))
    assert _bcsr_equal(lace.owned_bcsr_to_numpy(T), orc.bcsr_transpose(Bo))


def _dense_dm(lace, A):
    """Column-major dense data_d_matrix over a numpy array (kept alive by
    the caller through the returned tuple)."""
    a = np.asfortranarray(np.array(A, dtype=np.float64))
    m = lace.DMatrix()
    m.storage_type = lace.Magma_DENSE
    m.num_rows, m.num_cols = a.shape
    m.nnz = a.size
    m.ld = a.shape[0]
    m.major = lace.MagmaColMajor
    flat = a.reshape(-1, order="F")
    m.val = flat.ctypes.data_as(C.POINTER(C.c_double))
    return m, flat


def test_hqrd_and_housegen_dropins_match_goldens(lace, orc):
    """testing/matrices/magic7_Householder_{U,R}.mtx: the reference's
    Householder factors of magic(7)."""
    X = orc.read_mtx_dense(os.path.join(GOLDEN, "magic7.mtx"))
    Uref = orc.read_mtx_dense(os.path.join(GOLDEN, "magic7_Householder_U.mtx"))
    Rref = orc.read_mtx_dense(os.path.join(GOLDEN, "magic7_Householder_R.mtx"))
    Xm, xf = _dense_dm(lace, X)
    Um, uf = _dense_dm(lace, np.zeros((7, 7)))
    Rm, rf = _dense_dm(lace, np.zeros((7, 7)))
    lace.lib().data_hqrd(C.byref(Xm), C.byref(Um), C.byref(Rm))
    U = uf.reshape(7, 7, order="F"); R = rf.reshape(7, 7, order="F")
    assert np.allclose(U, Uref, rtol=1e-12, atol=1e-12)
    assert np.allclose(R, Rref, rtol=1e-12, atol=1e-11)
    # housegen on its own against the oracle and the reflector property
    x = np.array([3.0, -1.0, 2.0, 0.5])
    u = np.zeros(4); nu = C.c_double()
    lace.lib().data_housegen(4, x.ctypes.data_as(C.POINTER(C.c_double)),
                             u.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nu))
    uo, nuo = orc.housegen(x)
    assert np.allclose(u, uo, rtol=1e-15) and nu.value == nuo
    Hx = x - u * np.dot(u, x)
    assert np.allclose(Hx, [nu.value, 0, 0, 0], atol=1e-14)


def test_dense_mtx_read_write_round_trip(lace, orc, tmp_path):
    """control/magma_zmio.cpp:1221-1282: array files in storage order; the
    reader agrees with the oracle's column-major reader on the golden file
    and the writer's output reads back bit for bit."""
    path = os.path.join(GOLDEN, "magic7.mtx")
    D = lace.DMatrix()
    D.storage_type = lace.Magma_DENSE
    lace.check(lace.lib().data_z_dense_mtx(C.byref(D), lace.MagmaColMajor, path.encode()))
    assert (D.num_rows, D.num_cols, D.ld) == (7, 7, 7)
    got = lace.owned_to_numpy(D, kind="dense")
    assert np.array_equal(got, orc.read_mtx_dense(path).reshape(-1, order="F"))

    x = np.random.default_rng(5).standard_normal(13) * 10.0 ** np.arange(-6, 7)
    M, keep = _dense_dm(lace, x.reshape(13, 1))
    p = str(tmp_path / "x.mtx")
    lace.check(lace.lib().data_zwrite_dense(M, p.encode()))
    assert open(p).readline().split()[:4] == ["%%MatrixMarket", "matrix", "array", "real"]
    R = lace.DMatrix()
    R.storage_type = lace.Magma_DENSE
    lace.check(lace.lib().data_z_dense_mtx(C.byref(R), lace.MagmaRowMajor, p.encode()))
    assert (R.num_rows, R.num_cols, R.ld) == (13, 1, 1)
    assert np.array_equal(lace.owned_to_numpy(R, kind="dense"), x)


def test_zaxpy_strided(lace):
    x = np.arange(1.0, 13.0)
    y = np.full(12, 0.5)
    P = C.POINTER(C.c_double)
    lace.lib().data_zaxpy(4, -2.0, x.ctypes.data_as(P), 3, y.ctypes.data_as(P), 2)
    ref = np.full(12, 0.5)
    ref[0:8:2] += -2.0 * x[0:12:3]
    assert np.array_equal(y, ref)


def test_zdot_dropins_match_reference_goldens(lace, orc):
    """data_zdot / data_zdot_mkl (blas/zdot.cpp:50-78) on the reference's own
    dense fixture and expected values (4-ulp tolerance, as DoubleEq)."""
    import json
    gold = json.load(open(os.path.join(GOLDEN, "reference_goldens.json")))["dense_dot"]
    D = lace.DMatrix()
    D.storage_type = lace.Magma_DENSE
    path = os.path.join(GOLDEN, "io_dense_test.mtx")
    lace.check(lace.lib().data_z_dense_mtx(C.byref(D), lace.MagmaRowMajor, path.encode()))
    a = np.ascontiguousarray(lace.owned_to_numpy(D, kind="dense"), dtype=np.float64)
    assert a.size == 35
    P = C.POINTER(C.c_double)
    for case in gold["cases"]:
        sub = a[case["offset"]:]
        ptr = sub.ctypes.data_as(P)
        exp = case["expected"]
        for fn in (lace.lib().data_zdot, lace.lib().data_zdot_mkl):
            got = fn(case["n"], ptr, case["inc"], ptr, case["inc"])
            assert abs(got - exp) <= 4 * np.spacing(abs(exp)), (case, got)
