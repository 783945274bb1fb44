"""CPU oracle for the LACE preconditioned-solve path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker (or the timed CPU baseline); the
shipped library never calls it.  The numeric work lives in lace_oracle.c
(built by oracle/Makefile); this module wraps it with numpy and adds the
pure-Python pieces: the MatrixMarket reader (control/magma_zmio.cpp:92-350)
and independent restatements of the synthetic generators.

Parity status per function is recorded in lace_oracle.c's header; every
function there is pinned by at least one check in tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liblace_oracle.so")
_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc, a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: {_LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_i = C.c_int
_d = C.c_double
_p = C.c_void_p


def _declare(L):
    sig = {
        "orc_csr_split": (_i, [_i, _p, _p, _p, _i, _i, _p, _p, _p]),
        "orc_transpose": (None, [_i, _i, _p, _p, _p, _p, _p, _p]),
        "orc_spgemm": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
        "orc_ilures": (None, [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
        "orc_parilu_sweep_d": (_d, [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
        "orc_parilu_sweep_f": (_d, [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
        "orc_parilu_v03_omp": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _d, _p]),
        "orc_ilu0": (_i, [_i, _p, _p, _p]),
        "orc_trsv_lower_d": (None, [_i, _p, _p, _p, _p, _p]),
        "orc_trsv_upper_d": (None, [_i, _p, _p, _p, _p, _p]),
        "orc_trsv_lower_f": (None, [_i, _p, _p, _p, _p, _p]),
        "orc_trsv_upper_f": (None, [_i, _p, _p, _p, _p, _p]),
        "orc_trsv_jacobi_d": (None, [_i, _p, _p, _p, _i, _i, _p, _p]),
        "orc_trsv_jacobi_f": (None, [_i, _p, _p, _p, _i, _i, _p, _p]),
        "orc_parcsrtrsv": (_i, [_i, _p, _p, _p, _i, _p, _p, _d]),
        "orc_levels": (_i, [_i, _p, _p, _i, _p]),
        "orc_spmv": (None, [_i, _p, _p, _p, _d, _p, _d, _p]),
        "orc_givens": (None, [_d, _d, _d, _p, _p]),
        "orc_fgmres_householder": (_i, [_i, _p, _p, _p, _p, _p, _i, _i, _d, _i, _i, _i,
                                        _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p, _i]),
        "orc_fgmres_mgs": (_i, [_i, _p, _p, _p, _p, _p, _i, _i, _d, _i, _i, _i, _i,
                                _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p]),
        "orc_housegen": (None, [_i, _p, _p, _p]),
        "orc_hqrd": (None, [_i, _i, _p, _p, _p]),
        "orc_philox4x32_10": (None, [_p, _p, _p]),
        "orc_set_threads": (None, [_i]),
        "orc_bilu_sweep_d": (_d, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
        "orc_bilu_sweep_f": (_d, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
        "orc_binv_double": (None, [_i, _p, _p]),
        "orc_bcsr_spmv": (None, [_i, _i, _p, _p, _p, _p, _p]),
        "orc_btrsv": (None, [_i, _i, _p, _p, _p, _i, _i, _p, _p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _ci(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cd(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


@dataclass
class CSR:
    """Plain CSR (int32 indices).  For CSC views, `row` is the column
    pointer and `col` the row indices."""
    n_rows: int
    n_cols: int
    row: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row[-1])

    def dense(self) -> np.ndarray:
        D = np.zeros((self.n_rows, self.n_cols))
        for i in range(self.n_rows):
            for k in range(self.row[i], self.row[i + 1]):
                D[i, self.col[k]] += self.val[k]
        return D

    def rowidx(self) -> np.ndarray:
        return np.repeat(np.arange(self.n_rows, dtype=np.int32), np.diff(self.row))

    def scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.row), shape=(self.n_rows, self.n_cols))


# ---------------------------------------------------------------- I/O
def read_mtx(path: str) -> CSR:
    """MatrixMarket coordinate reader following control/magma_zmio.cpp:92-350:
    1-based COO; symmetric/hermitian files duplicate every off-diagonal entry
    (:221-273); counting sort into CSR; each row sorted by column (:323-336).
    Pattern files get value 1.0; complex files keep the real part."""
    with open(path) as f:
        banner = f.readline().lower().split()
        if len(banner) < 5 or banner[0] != "%%matrixmarket" or banner[2] != "coordinate":
            raise ValueError("only coordinate MatrixMarket files are supported")
        field, symm = banner[3], banner[4]
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        m, n, nz = (int(t) for t in line.split()[:3])
        rows, cols, vals = [], [], []
        for _ in range(nz):
            t = f.readline().split()
            while not t:
                t = f.readline().split()
            rows.append(int(t[0]) - 1)
            cols.append(int(t[1]) - 1)
            vals.append(1.0 if field == "pattern" else float(t[2]))
    r = np.array(rows, dtype=np.int64)
    c = np.array(cols, dtype=np.int64)
    v = np.array(vals, dtype=np.float64)
    if symm in ("symmetric", "hermitian"):
        rr, cc, vv = [], [], []
        for a, b, x in zip(r, c, v):
            rr.append(a); cc.append(b); vv.append(x)
            if a != b:
                rr.append(b); cc.append(a); vv.append(x)
        r, c, v = np.array(rr), np.array(cc), np.array(vv)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    row = np.zeros(m + 1, dtype=np.int32)
    np.add.at(row, r + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    return CSR(m, n, row, c.astype(np.int32), v)


def read_mtx_dense(path: str) -> np.ndarray:
    """MatrixMarket `array` (column-major) reader."""
    with open(path) as f:
        f.readline()
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        m, n = (int(t) for t in line.split()[:2])
        vals = [float(t) for t in f.read().split()]
    return np.array(vals[: m * n]).reshape((n, m)).T.copy()


# ---------------------------------------------------------------- generators
def laplace2d(nx: int, ny: int) -> CSR:
    """5-point [4,-1,-1,-1,-1] with Dirichlet boundaries, x fastest."""
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx):
            p = i + nx * j
            ent = []
            if j > 0: ent.append((p - nx, -1.0))
            if i > 0: ent.append((p - 1, -1.0))
            ent.append((p, 4.0))
            if i < nx - 1: ent.append((p + 1, -1.0))
            if j < ny - 1: ent.append((p + nx, -1.0))
            for q, a in ent:
                rows.append(p); cols.append(q); vals.append(a)
    return _coo_to_csr(nx * ny, rows, cols, vals)


def convdiff3d(nx: int, ny: int, nz: int, pe=(2.0, 1.0, 0.5)) -> CSR:
    """7-point first-order upwind convection-diffusion, h-scaled:
    a_ii = 6 + sum(pe), a_{i,i-e_d} = -1 - pe_d, a_{i,i+e_d} = -1,
    Dirichlet (neighbours outside the box dropped), x fastest."""
    import scipy.sparse as sp

    def axis(m, pe_d):
        lo = np.full(m - 1, -1.0 - pe_d)
        hi = np.full(m - 1, -1.0)
        return sp.diags([lo, hi], [-1, 1], shape=(m, m))

    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    A = (sp.kron(Iz, sp.kron(Iy, axis(nx, pe[0])))
         + sp.kron(Iz, sp.kron(axis(ny, pe[1]), Ix))
         + sp.kron(axis(nz, pe[2]), sp.kron(Iy, Ix)))
    A = A + sp.identity(nx * ny * nz) * (6.0 + sum(pe))
    A = sp.csr_matrix(A)
    A.sort_indices()
    return CSR(A.shape[0], A.shape[1], A.indptr.astype(np.int32),
               A.indices.astype(np.int32), A.data.astype(np.float64))


def _coo_to_csr(n, rows, cols, vals) -> CSR:
    r = np.array(rows); c = np.array(cols); v = np.array(vals, dtype=np.float64)
    o = np.lexsort((c, r))
    r, c, v = r[o], c[o], v[o]
    row = np.zeros(n + 1, dtype=np.int32)
    np.add.at(row, r + 1, 1)
    return CSR(n, n, np.cumsum(row).astype(np.int32), c.astype(np.int32), v)


def set_threads(n: int) -> None:
    """OpenMP thread count of the oracle's parallel loops (0: default)."""
    lib().orc_set_threads(n)


def philox4x32_10(ctr, key):
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(np.asarray(ctr, dtype=np.uint32)),
                            _ptr(np.asarray(key, dtype=np.uint32)), _ptr(out))
    return out


# ---------------------------------------------------------------- formats
DIAG_OMIT, DIAG_UNIT, DIAG_VALUE, DIAG_ZERO = 0, 1, 2, 3


def csr_split(A: CSR, upper: bool, diag_mode: int) -> CSR:
    """CSRL (upper=False) / CSRU (upper=True): magma_zmconverter.cpp:303-460."""
    L = lib()
    row, col, val = _ci(A.row), _ci(A.col), _cd(A.val)
    nz = L.orc_csr_split(A.n_rows, _ptr(row), _ptr(col), _ptr(val), int(upper), diag_mode,
                         None, None, None)
    tr = np.zeros(A.n_rows + 1, np.int32); tc = np.zeros(nz, np.int32); tv = np.zeros(nz)
    L.orc_csr_split(A.n_rows, _ptr(row), _ptr(col), _ptr(val), int(upper), diag_mode,
                    _ptr(tr), _ptr(tc), _ptr(tv))
    return CSR(A.n_rows, A.n_cols, tr, tc, tv)


def transpose(A: CSR) -> CSR:
    tr = np.zeros(A.n_cols + 1, np.int32); tc = np.zeros(A.nnz, np.int32); tv = np.zeros(A.nnz)
    row, col, val = _ci(A.row), _ci(A.col), _cd(A.val)
    lib().orc_transpose(A.n_rows, A.n_cols, _ptr(row), _ptr(col), _ptr(val),
                        _ptr(tr), _ptr(tc), _ptr(tv))
    return CSR(A.n_cols, A.n_rows, tr, tc, tv)


def spgemm(A: CSR, B: CSR) -> CSR:
    L = lib()
    a = (_ci(A.row), _ci(A.col), _cd(A.val)); b = (_ci(B.row), _ci(B.col), _cd(B.val))
    nz = L.orc_spgemm(A.n_rows, B.n_cols, *map(_ptr, a), *map(_ptr, b), None, None, None)
    cr = np.zeros(A.n_rows + 1, np.int32); cc = np.zeros(nz, np.int32); cv = np.zeros(nz)
    L.orc_spgemm(A.n_rows, B.n_cols, *map(_ptr, a), *map(_ptr, b), _ptr(cr), _ptr(cc), _ptr(cv))
    return CSR(A.n_rows, B.n_cols, cr, cc, cv)


def ilures(A: CSR, Lf: CSR, Ucsr: CSR):
    """(res, nonlinres) of control/norms.cpp:301-405; L CSRL, U CSR upper."""
    res = C.c_double(); nl = C.c_double()
    arrs = [_ci(A.row), _ci(A.col), _cd(A.val), _ci(Lf.row), _ci(Lf.col), _cd(Lf.val),
            _ci(Ucsr.row), _ci(Ucsr.col), _cd(Ucsr.val)]
    lib().orc_ilures(A.n_rows, *map(_ptr, arrs), C.byref(res), C.byref(nl))
    return res.value, nl.value


# ---------------------------------------------------------------- ParILU
def scalar_row_scales(A: CSR) -> np.ndarray:
    """Power-of-two scale per row: 2^e with 2^e <= max_j |a_ij| < 2^(e+1)
    (1 for a zero row), e clamped to the FP32 normal range."""
    rs = np.ones(A.n_rows, np.float32)
    for i in range(A.n_rows):
        v = A.val[A.row[i]:A.row[i + 1]]
        if len(v) == 0:
            continue
        m = np.abs(v).max()
        if m > 0 and np.isfinite(m):
            _, e = np.frexp(m)
            rs[i] = np.float32(2.0) ** int(np.clip(e - 1, -126, 127))
    return rs


def round16(v, s, storage: str) -> np.ndarray:
    """The 16-bit storage model: v / s rounded to BF16 or FP16 (nearest
    even), times s, all in FP32 (s a power of two, so the scalings are
    exact)."""
    t = (np.asarray(v, np.float32) * (np.float32(1.0) / s)).astype(np.float32)
    q = bf16_round(t) if storage == "bf16" else t.astype(np.float16).astype(np.float32)
    return (q * s).astype(np.float32)


class ParILUState:
    """L (CSRL, unit diagonal last) and U (CSCU: column pointer in U.row,
    row indices in U.col, diagonal last per column) initialised from A as
    src/parilu_v0_3.cpp:23-27 does.  storage 'bf16' / 'f16' iterates in
    FP32 and, at the start and after every sweep, rounds each factor value
    to 16 bits relative to the power-of-two scale of its row (L's unit
    diagonal is kept exact)."""

    def __init__(self, A: CSR, dtype=np.float64, storage=None):
        self.A = A
        self.storage = storage
        if storage:
            dtype = np.float32
        self.dtype = dtype
        self.L = csr_split(A, False, DIAG_UNIT)
        Ucsr = csr_split(A, True, DIAG_VALUE)
        self.U = transpose(Ucsr)
        self.Lval = self.L.val.astype(dtype)
        self.Uval = self.U.val.astype(dtype)
        self.Aval = A.val.astype(dtype)
        if storage:
            self.rs = scalar_row_scales(A)
            self._lrow = np.repeat(np.arange(A.n_rows), np.diff(self.L.row))
            self._ldiag = self.L.col == self._lrow
            self._urow = np.asarray(self.U.col)  # CSC: row index of each U entry
            self.Lval, self.Uval = self._round(self.Lval, self.Uval)

    def _round(self, Lv, Uv):
        Lq = round16(Lv, self.rs[self._lrow], self.storage)
        Lq[self._ldiag] = Lv[self._ldiag]
        return Lq, round16(Uv, self.rs[self._urow], self.storage)

    def sweep(self, mode: str = "jacobi") -> float:
        A, L, U = self.A, self.L, self.U
        fn = lib().orc_parilu_sweep_d if self.dtype == np.float64 else lib().orc_parilu_sweep_f
        a = (_ci(A.row), _ci(A.col), self.Aval)
        lrc = (_ci(L.row), _ci(L.col)); urc = (_ci(U.row), _ci(U.col))
        if mode == "jacobi":
            Lout, Uout = self.Lval.copy(), self.Uval.copy()
        elif mode == "gs":
            Lout, Uout = self.Lval, self.Uval
        else:
            raise ValueError(mode)
        step = fn(A.n_rows, *map(_ptr, a), *map(_ptr, lrc), _ptr(self.Lval), _ptr(Lout),
                  *map(_ptr, urc), _ptr(self.Uval), _ptr(Uout))
        if self.storage:
            if mode != "jacobi":
                raise ValueError("16-bit storage runs Jacobi sweeps only")
            Lout, Uout = self._round(Lout, Uout)
            step = float(np.sum((self.Lval.astype(np.float64) - Lout) ** 2) +
                         np.sum((self.Uval.astype(np.float64) - Uout) ** 2))
        self.Lval, self.Uval = Lout, Uout
        return step

    def factors(self):
        """(L as CSRL, U as CSRU) with the current values (float64)."""
        Lf = CSR(self.L.n_rows, self.L.n_cols, self.L.row, self.L.col, self.Lval.astype(np.float64))
        Uc = CSR(self.U.n_rows, self.U.n_cols, self.U.row, self.U.col, self.Uval.astype(np.float64))
        return Lf, transpose(Uc)


def parilu(A: CSR, sweeps: int, mode: str = "jacobi", dtype=np.float64, storage=None):
    st = ParILUState(A, dtype, storage)
    for _ in range(sweeps):
        st.sweep(mode)
    return st


def parilu_v03_omp(A: CSR, max_sweeps: int):
    """The reference's OpenMP in-place loop (CPU baseline); returns
    (state, sweeps_run, final_step, seconds)."""
    import time
    st = ParILUState(A)
    anorm = float(np.sqrt(np.sum(A.val ** 2)))
    rowidx = _ci(A.rowidx())
    fs = C.c_double()
    t0 = time.perf_counter()
    it = lib().orc_parilu_v03_omp(A.n_rows, A.nnz, _ptr(_ci(A.row)), _ptr(rowidx), _ptr(_ci(A.col)),
                                  _ptr(_cd(A.val)), _ptr(_ci(st.L.row)), _ptr(_ci(st.L.col)),
                                  _ptr(st.Lval), _ptr(_ci(st.U.row)), _ptr(_ci(st.U.col)),
                                  _ptr(st.Uval), max_sweeps, 1.0 / anorm, C.byref(fs))
    return st, it, fs.value, time.perf_counter() - t0


def ilu0(A: CSR):
    """Exact ILU(0) in A's pattern (merged strict-L / U); returns (CSR, bad_row)."""
    v = _cd(A.val).copy()
    bad = lib().orc_ilu0(A.n_rows, _ptr(_ci(A.row)), _ptr(_ci(A.col)), _ptr(v))
    return CSR(A.n_rows, A.n_cols, A.row.copy(), A.col.copy(), v), bad


def parilu_tiled(A: CSR, tile: int, max_iter: int = 1000):
    """The tiled ParILU iteration of src/parilu_v3_0.cpp:15-229 as the
    reference runs it on one thread, restated with dense tiles in pure
    Python (small matrices only): for each tile (ti, tj) of the padded grid
    in row-major order, Lblock (tile x span, strict L, columns < span) and
    Ublock (span x tile, strict U), span = min(ti, tj) + tile, are copied
    out (control/sparse_sub.cpp:77-96), C = Lblock Ublock is summed over the
    inner index in ascending order from zero, and then (ti <= tj) the tile's
    strict U entries become Uc - C, on the diagonal tile D_i = 1/(Dc_i -
    C_ii) and the tile's L entries (Lc - C) D_j; (ti > tj) the L entries.
    The step is sum (old - new)^2 / ||A||_F; sweeps run while it exceeds
    1e-15, at most max_iter.  U's diagonal ends as 1 / D.  Returns (L values
    in CSRL order with the unit diagonal last, U values in CSRU order with
    the diagonal first, sweeps)."""
    import math
    n = A.n_rows
    lp = [[] for _ in range(n)]
    up = [[] for _ in range(n)]
    Lc = {}; Uc = {}; Dc = [0.0] * n
    for i in range(n):
        for k in range(A.row[i], A.row[i + 1]):
            j = int(A.col[k]); v = float(A.val[k])
            if j < i:
                lp[i].append(j); Lc[(i, j)] = v
            elif j > i:
                up[i].append(j); Uc[(i, j)] = v
            else:
                Dc[i] = v
    Lv = dict(Lc); Uv = dict(Uc)
    D = [1.0 / d for d in Dc]
    pad = -(-n // tile) * tile
    anorm = math.sqrt(sum(float(v) * float(v) for v in A.val[:A.row[n]]))
    step, it = 1.0, 0
    while step > 1e-15 and it < max_iter:
        step = 0.0
        for ti in range(0, pad, tile):
            for tj in range(0, pad, tile):
                span = min(ti + tile, tj + tile)
                Lb = [[Lv.get((i, k), 0.0) for k in range(span)] for i in range(ti, ti + tile)]
                Ub = [[Uv.get((k, j), 0.0) for j in range(tj, tj + tile)] for k in range(span)]
                Ct = [[0.0] * tile for _ in range(tile)]
                for a in range(tile):
                    for b in range(tile):
                        acc = 0.0
                        for k in range(span):
                            acc = acc + Lb[a][k] * Ub[k][b]
                        Ct[a][b] = acc
                rows = range(ti, min(ti + tile, n))
                if ti <= tj:
                    for i in rows:
                        for j in up[i]:
                            if tj <= j < tj + tile:
                                t = Uc[(i, j)] - Ct[i - ti][j - tj]
                                step += math.pow(Uv[(i, j)] - t, 2)
                                Uv[(i, j)] = t
                    if ti == tj:
                        for i in rows:
                            t = 1.0 / (Dc[i] - Ct[i - ti][i - tj])
                            step += math.pow(D[i] - t, 2)
                            D[i] = t
                if ti >= tj:
                    for i in rows:
                        for j in lp[i]:
                            if tj <= j < tj + tile:
                                t = (Lc[(i, j)] - Ct[i - ti][j - tj]) * D[j]
                                step += math.pow(Lv[(i, j)] - t, 2)
                                Lv[(i, j)] = t
        step /= anorm
        it += 1
    lval = []
    uval = []
    for i in range(n):
        lval += [Lv[(i, j)] for j in lp[i]] + [1.0]
        uval += [1.0 / D[i]] + [Uv[(i, j)] for j in up[i]]
    return np.array(lval), np.array(uval), it


def split_merged(LU: CSR):
    """Merged ILU (data_zmlumerge layout) -> (CSRL unit diag, CSRU)."""
    Lf = csr_split(LU, False, DIAG_UNIT)
    Uf = csr_split(LU, True, DIAG_VALUE)
    return Lf, Uf


# ---------------------------------------------------------------- trsv
def trsv(T: CSR, r, upper: bool, dtype=np.float64):
    L = lib()
    y = np.zeros(T.n_rows, dtype=dtype)
    rr = _cd(r, dtype); tv = _cd(T.val, dtype)
    if dtype == np.float64:
        fn = L.orc_trsv_upper_d if upper else L.orc_trsv_lower_d
    else:
        fn = L.orc_trsv_upper_f if upper else L.orc_trsv_lower_f
    fn(T.n_rows, _ptr(_ci(T.row)), _ptr(_ci(T.col)), _ptr(tv), _ptr(rr), _ptr(y))
    return y


def trsv_jacobi(T: CSR, r, upper: bool, sweeps: int, dtype=np.float64):
    L = lib()
    y = np.zeros(T.n_rows, dtype=dtype)
    fn = L.orc_trsv_jacobi_d if dtype == np.float64 else L.orc_trsv_jacobi_f
    fn(T.n_rows, _ptr(_ci(T.row)), _ptr(_ci(T.col)), _ptr(_cd(T.val, dtype)), int(upper), sweeps,
       _ptr(_cd(r, dtype)), _ptr(y))
    return y


def parcsrtrsv(T: CSR, r, upper: bool, tol: float, y0=None):
    y = np.zeros(T.n_rows) if y0 is None else _cd(y0).copy()
    it = lib().orc_parcsrtrsv(T.n_rows, _ptr(_ci(T.row)), _ptr(_ci(T.col)), _ptr(_cd(T.val)),
                              int(upper), _ptr(_cd(r)), _ptr(y), tol)
    return y, it


def levels(T: CSR, upper: bool):
    lev = np.zeros(T.n_rows, np.int32)
    nl = lib().orc_levels(T.n_rows, _ptr(_ci(T.row)), _ptr(_ci(T.col)), int(upper), _ptr(lev))
    return lev, nl


def spmv(A: CSR, x, alpha=1.0, beta=0.0, y=None):
    yy = np.zeros(A.n_rows) if y is None else _cd(y).copy()
    lib().orc_spmv(A.n_rows, _ptr(_ci(A.row)), _ptr(_ci(A.col)), _ptr(_cd(A.val)), alpha,
                   _ptr(_cd(x)), beta, _ptr(yy))
    return yy


def givens(a, b, tol):
    c = C.c_double(); s = C.c_double()
    lib().orc_givens(a, b, tol, C.byref(c), C.byref(s))
    return c.value, s.value


# ---------------------------------------------------------------- FGMRES
@dataclass
class FgmresResult:
    x: np.ndarray
    info: int
    searches: int
    restarts: int
    hist: np.ndarray
    final_res: float


def _fgmres_common(kind, A, b, x0, m, restart_max, rtol, tol_type, choice, jac, Lf, Uf,
                   reorth=0, hist_len=4096, nthreads=1):
    x = _cd(x0).copy() if x0 is not None else np.zeros(A.n_rows)
    hist = np.zeros(hist_len)
    so = C.c_int(); ro = C.c_int(); fr = C.c_double()
    if Lf is None:
        Lf = CSR(A.n_rows, A.n_cols, np.zeros(A.n_rows + 1, np.int32), np.zeros(1, np.int32),
                 np.zeros(1))
        Uf = Lf
    fac = [_ci(Lf.row), _ci(Lf.col), _cd(Lf.val), _ci(Uf.row), _ci(Uf.col), _cd(Uf.val)]
    a = [_ci(A.row), _ci(A.col), _cd(A.val)]
    if kind == "householder":
        info = lib().orc_fgmres_householder(
            A.n_rows, *map(_ptr, a), _ptr(_cd(b)), _ptr(x), m, restart_max, rtol, tol_type,
            choice, jac, *map(_ptr, fac), C.byref(so), C.byref(ro), _ptr(hist), hist_len,
            C.byref(fr), nthreads)
    else:
        info = lib().orc_fgmres_mgs(
            A.n_rows, *map(_ptr, a), _ptr(_cd(b)), _ptr(x), m, restart_max, rtol, tol_type,
            reorth, choice, jac, *map(_ptr, fac), C.byref(so), C.byref(ro), _ptr(hist),
            hist_len, C.byref(fr))
    return FgmresResult(x, info, so.value, ro.value, hist[: min(so.value, hist_len)], fr.value)


def fgmres_householder(A, b, x0=None, m=30, restart_max=1, rtol=1e-10, tol_type=1, choice=0,
                       jac=5, Lf=None, Uf=None, nthreads=1):
    """src/fgmres_householder_restart.cpp:63-788.  nthreads > 1 runs the
    n-length loops under OpenMP (CPU baseline); 1 is deterministic."""
    return _fgmres_common("householder", A, b, x0, m, restart_max, rtol, tol_type, choice, jac,
                          Lf, Uf, nthreads=nthreads)


def fgmres_mgs(A, b, x0=None, m=30, restart_max=1, rtol=1e-10, tol_type=1, choice=0, jac=5,
               Lf=None, Uf=None, reorth=0):
    """src/fgmres.cpp:64-517 with restarts (src/fgmres_restart.cpp)."""
    return _fgmres_common("mgs", A, b, x0, m, restart_max, rtol, tol_type, choice, jac, Lf, Uf,
                          reorth=reorth)


# ---------------------------------------------------------------- Householder QR
def housegen(x):
    x = _cd(x)
    u = np.zeros_like(x); nu = C.c_double()
    lib().orc_housegen(len(x), _ptr(x), _ptr(u), C.byref(nu))
    return u, nu.value


def hqrd(X):
    X = np.asfortranarray(np.array(X, dtype=np.float64))
    m, n = X.shape
    U = np.zeros((m, n), order="F"); R = np.zeros((n, n), order="F")
    lib().orc_hqrd(m, n, _ptr(X), _ptr(U), _ptr(R))
    return U, R


# ---------------------------------------------------------------- BCSR
@dataclass
class BCSR:
    """Block CSR: nbr block rows, bs x bs blocks stored row-major in
    val[numblocks, bs, bs].  For block-column (BCSC) views, `row` is the
    block column pointer and `col` the block row indices; block interiors
    are never transposed (control/magma_zmtranspose.cpp:237-380)."""
    nbr: int
    bs: int
    row: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nblocks(self) -> int:
        return int(self.row[-1])

    def dense(self) -> np.ndarray:
        bs = self.bs
        D = np.zeros((self.nbr * bs, self.nbr * bs))
        for I in range(self.nbr):
            for k in range(self.row[I], self.row[I + 1]):
                J = self.col[k]
                D[I * bs:(I + 1) * bs, J * bs:(J + 1) * bs] += self.val[k]
        return D


def bcsr_from_csr(A: CSR, bs: int) -> BCSR:
    """CSR -> BCSR (control/magma_zmconverter.cpp:639-677 via
    mkl_dcsrbsr, zero-based so blocks are row-major): a block is kept when
    any stored entry falls in it; block columns ascending."""
    nbr = (A.n_rows + bs - 1) // bs
    row = [0]; col = []; blocks = []
    for I in range(nbr):
        cols = sorted({int(A.col[k]) // bs for i in range(I * bs, min(A.n_rows, (I + 1) * bs))
                       for k in range(A.row[i], A.row[i + 1])})
        for J in cols:
            blk = np.zeros((bs, bs))
            for i in range(I * bs, min(A.n_rows, (I + 1) * bs)):
                for k in range(A.row[i], A.row[i + 1]):
                    if A.col[k] // bs == J:
                        blk[i - I * bs, A.col[k] - J * bs] += A.val[k]
            col.append(J); blocks.append(blk)
        row.append(len(col))
    val = np.array(blocks) if blocks else np.zeros((0, bs, bs))
    return BCSR(nbr, bs, np.array(row, np.int32), np.array(col, np.int32), val)


def bcsr_split(B: BCSR, upper: bool, diag: str) -> BCSR:
    """BCSRL / BCSRU (control/magma_zmconverter.cpp:1248-1436): strict
    lower (upper) blocks and the diagonal block -- 'unit' identity,
    'value' A's block, 'omit'.  Rows sorted, so BCSRL keeps the diagonal
    block last and BCSRU first."""
    row = [0]; col = []; blocks = []
    for I in range(B.nbr):
        for k in range(B.row[I], B.row[I + 1]):
            J = int(B.col[k])
            if (upper and J > I) or (not upper and J < I):
                col.append(J); blocks.append(B.val[k].copy())
            elif J == I and diag != "omit":
                col.append(J)
                blocks.append(np.eye(B.bs) if diag == "unit" else B.val[k].copy())
        row.append(len(col))
    return BCSR(B.nbr, B.bs, np.array(row, np.int32), np.array(col, np.int32), np.array(blocks))


def bcsr_transpose(B: BCSR) -> BCSR:
    """Block transpose by counting sort; blocks move, interiors stay."""
    nb = B.nblocks
    cnt = np.zeros(B.nbr + 1, np.int64)
    for k in range(nb):
        cnt[B.col[k] + 1] += 1
    ptr = np.cumsum(cnt).astype(np.int32)
    fill = ptr[:-1].copy()
    col = np.zeros(nb, np.int32); val = np.zeros_like(B.val)
    for I in range(B.nbr):
        for k in range(B.row[I], B.row[I + 1]):
            q = fill[B.col[k]]; fill[B.col[k]] += 1
            col[q] = I; val[q] = B.val[k]
    return BCSR(B.nbr, B.bs, ptr, col, val)


def gen_bcsr7_pattern(nx: int, ny: int, nz: int):
    """(row, col) of the 7-point block stencil, x fastest, Dirichlet,
    block columns ascending."""
    row = [0]; col = []
    sy, sz = nx, nx * ny
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                p = i + sy * j + sz * k
                nb = []
                if k > 0: nb.append(p - sz)
                if j > 0: nb.append(p - sy)
                if i > 0: nb.append(p - 1)
                nb.append(p)
                if i < nx - 1: nb.append(p + 1)
                if j < ny - 1: nb.append(p + sy)
                if k < nz - 1: nb.append(p + sz)
                col.extend(nb); row.append(len(col))
    return np.array(row, np.int32), np.array(col, np.int32)


def bf16_round(x) -> np.ndarray:
    """float32 -> bfloat16 (round to nearest even), returned as float32."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7fffffff) > 0x7f800000
    r = ((u + 0x7fff + ((u >> 16) & 1)) >> 16) & 0xffff
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32)


def row_scales(B: BCSR) -> np.ndarray:
    """Power-of-two scale per scalar row: 2^e with 2^e <= max|row| < 2^(e+1)."""
    bs = B.bs
    rs = np.ones(B.nbr * bs, np.float32)
    for I in range(B.nbr):
        blk = B.val[B.row[I]:B.row[I + 1]]
        if len(blk) == 0:
            continue
        m = np.abs(blk).max(axis=(0, 2))
        for a in range(bs):
            if m[a] > 0 and np.isfinite(m[a]):
                _, e = np.frexp(m[a])
                rs[I * bs + a] = np.float32(2.0) ** int(np.clip(e - 1, -126, 127))
    return rs


class BILUState:
    """Block ParILU(0) of src/parilu_v0_3_bcsr.cpp:16-276: L as BCSRL
    (identity diagonal block last), U as BCSCU (block columns, diagonal
    block last, interiors as stored).  dtype float64 or float32 iterates
    in that precision; storage 'bf16' / 'f16' keeps the factors rounded to
    16 bits with a power-of-two scale per scalar row (FP32 iteration) and
    runs Jacobi sweeps only."""

    def __init__(self, B: BCSR, dtype=np.float64, storage=None):
        self.B = B
        self.bs = B.bs
        self.dtype = np.float32 if storage else dtype
        self.storage = storage
        self.L = bcsr_split(B, False, "unit")
        self.U = bcsr_transpose(bcsr_split(B, True, "value"))
        self.Aval = np.ascontiguousarray(B.val.reshape(-1), dtype=np.float64)
        self.Lval = self.L.val.reshape(-1).astype(self.dtype)
        self.Uval = self.U.val.reshape(-1).astype(self.dtype)
        if storage:
            self.rs = row_scales(B)
            self.Lval = self._q(self.Lval, self._lrows(), diag_unit=True)
            self.Uval = self._q(self.Uval, self._urows())

    # scalar row of each stored entry
    def _lrows(self):
        bs = self.bs
        I = np.repeat(np.arange(self.B.nbr), np.diff(self.L.row))
        return (I[:, None] * bs + np.arange(bs)[None, :]).repeat(bs, axis=1).reshape(-1)

    def _urows(self):
        bs = self.bs
        K = self.U.col
        return (K[:, None] * bs + np.arange(bs)[None, :]).repeat(bs, axis=1).reshape(-1)

    def _q(self, v, rows, diag_unit=False):
        s = self.rs[rows]
        t = (v.astype(np.float32) * (np.float32(1.0) / s)).astype(np.float32)
        q = bf16_round(t) if self.storage == "bf16" else t.astype(np.float16).astype(np.float32)
        out = (q * s).astype(np.float32)
        if diag_unit:  # the identity blocks are never stored or rounded
            bs = self.bs
            I = np.repeat(np.arange(self.B.nbr), np.diff(self.L.row))
            d = (self.L.col == I).repeat(bs * bs)
            out[d] = v[d]
        return out

    def sweep(self, mode: str = "jacobi") -> float:
        L, U, B = self.L, self.U, self.B
        fn = lib().orc_bilu_sweep_d if self.dtype == np.float64 else lib().orc_bilu_sweep_f
        if mode == "jacobi":
            Lout, Uout = self.Lval.copy(), self.Uval.copy()
        elif mode == "gs" and not self.storage:
            Lout, Uout = self.Lval, self.Uval
        else:
            raise ValueError(mode)
        step = fn(B.nbr, B.bs, _ptr(_ci(B.row)), _ptr(_ci(B.col)), _ptr(self.Aval),
                  _ptr(_ci(L.row)), _ptr(_ci(L.col)), _ptr(self.Lval), _ptr(Lout),
                  _ptr(_ci(U.row)), _ptr(_ci(U.col)), _ptr(self.Uval), _ptr(Uout))
        if self.storage:
            Lout = self._q(Lout, self._lrows(), diag_unit=True)
            Uout = self._q(Uout, self._urows())
            step = float(np.sum((self.Lval.astype(np.float64) - Lout) ** 2) +
                         np.sum((self.Uval.astype(np.float64) - Uout) ** 2))
        self.Lval, self.Uval = Lout, Uout
        return step

    def factors(self):
        """(L as BCSRL, U as BCSCU) with the current values (float64)."""
        bs = self.bs
        Lf = BCSR(self.B.nbr, bs, self.L.row, self.L.col, self.Lval.astype(np.float64).reshape(-1, bs, bs))
        Uf = BCSR(self.B.nbr, bs, self.U.row, self.U.col, self.Uval.astype(np.float64).reshape(-1, bs, bs))
        return Lf, Uf


def bilu(B: BCSR, sweeps: int, mode: str = "jacobi", dtype=np.float64, storage=None):
    st = BILUState(B, dtype, storage)
    for _ in range(sweeps):
        st.sweep(mode)
    return st


def binv(M) -> np.ndarray:
    M = np.ascontiguousarray(M, dtype=np.float64)
    out = np.zeros_like(M)
    lib().orc_binv_double(M.shape[0], _ptr(M), _ptr(out))
    return out


def bcsr_spmv(B: BCSR, x) -> np.ndarray:
    x = _cd(x); y = np.zeros(B.nbr * B.bs)
    lib().orc_bcsr_spmv(B.nbr, B.bs, _ptr(_ci(B.row)), _ptr(_ci(B.col)),
                        _ptr(_cd(B.val.reshape(-1))), _ptr(x), _ptr(y))
    return y


def btrsv(T: BCSR, r, upper: bool, jac: int = 0) -> np.ndarray:
    """FP32 block triangular solve: T lower with the diagonal block last
    (BCSRL) or upper with it first (BCSRU); jac > 0 runs that many Jacobi
    sweeps from zero instead of exact substitution."""
    r = np.ascontiguousarray(r, dtype=np.float32)
    y = np.zeros(T.nbr * T.bs, np.float32)
    lib().orc_btrsv(T.nbr, T.bs, _ptr(_ci(T.row)), _ptr(_ci(T.col)),
                    _ptr(np.ascontiguousarray(T.val.reshape(-1), dtype=np.float32)),
                    1 if upper else 0, jac, _ptr(r), _ptr(y))
    return y


# ---------------------------------------------------------------- orderings
def cuthill_mckee(A: CSR, reverse: bool = False):
    """Restatement of src/cuthill_mckee.cpp:53-265 written from the reference
    text with its indexing as-is: 1-offset neighbour store (:100-115), the
    window [nnsn[seed-1], nnsn[seed]) (:155-157, nnsn[-1] taken as 0), seed
    scan starting from row 1's degree (:130-142), queue sorted stably by the
    degree of node v+1 descending (:183-189), position-swap reverse
    (:241-252).  Returns nn_map, or None where the reference would read an
    uninitialised queue entry."""
    nn = A.n_rows
    nnsn = [0] * (nn + 1)
    nsn = [0] * (max(A.nnz, 1) + 1)
    for i in range(nn):
        offset = nnsn[i - 1] if i > 0 else 0
        for j in range(A.row[i], A.row[i + 1]):
            if A.col[j] != i:
                nnsn[i] += 1
                nsn[offset + nnsn[i]] = int(A.col[j])
        if i > 0:
            nnsn[i] += nnsn[i - 1]
    seed, maxdegree, degree = 0, nnsn[1] - nnsn[0], nnsn[0]
    for i in range(nn):
        if i > 0:
            degree = nnsn[i] - nnsn[i - 1]
        if degree > maxdegree:
            maxdegree, seed = degree, i
    nn_map = [-1] * nn
    queue = [0] * nn
    inode = 0
    nn_map[seed] = inode; inode += 1; queue[seed] = 1
    S = []
    start_s = 0
    while inode < nn:
        lo = nnsn[seed - 1] if seed > 0 else 0
        R = []
        for i in range(lo, nnsn[seed]):
            if queue[nsn[i]] == 0:
                R.append(nsn[i]); queue[nsn[i]] = 1
        R.sort(key=lambda v: -(nnsn[v + 1] - nnsn[v]))  # stable
        for v in R:
            nn_map[v] = inode; inode += 1; S.append(v)
        if inode >= nn:
            break
        if start_s >= len(S):
            return None
        seed = S[start_s]; start_s += 1
    if reverse:
        counter = nn - 1
        for i in range(nn // 2 + 1):
            nn_map[counter], nn_map[i] = nn_map[i], nn_map[counter]
            counter -= 1
    return np.array(nn_map, dtype=np.int32)


def rcm_textbook(A: CSR):
    """Reverse Cuthill-McKee as George & Liu (1981, ch. 4) state it, the
    ordering lace_reorder_rcm computes (an addition: not the reference's):
    symmetrised graph without self loops; per component a pseudo-peripheral
    start (from the least (degree, number) node, move to the last BFS
    level's least node while the level count grows); breadth-first
    numbering with unnumbered neighbours by increasing (degree, number);
    the order reversed.  Returns P with P[k] = old number of new row k."""
    n = A.n_rows
    nb = [set() for _ in range(n)]
    for i in range(n):
        for k in range(A.row[i], A.row[i + 1]):
            j = int(A.col[k])
            if j != i:
                nb[i].add(j)
                nb[j].add(i)
    adj = [sorted(s) for s in nb]
    deg = [len(a) for a in adj]
    key = lambda v: (deg[v], v)  # noqa: E731
    numbered = [False] * n

    def levels_from(r):
        seen = {r}
        level = [r]
        depth = 0
        last = level
        while level:
            last = level
            depth += 1
            nxt = []
            for v in level:
                for u in adj[v]:
                    if not numbered[u] and u not in seen:
                        seen.add(u)
                        nxt.append(u)
            level = nxt
        return depth, min(last, key=key)

    order = []
    for start in sorted(range(n), key=key):
        if numbered[start]:
            continue
        depth, far = levels_from(start)
        while True:
            d2, far2 = levels_from(far)
            root = far
            if d2 <= depth:
                break
            depth, far = d2, far2
        numbered[root] = True
        comp = [root]
        t = 0
        while t < len(comp):
            cand = sorted((u for u in adj[comp[t]] if not numbered[u]), key=key)
            for u in cand:
                numbered[u] = True
                comp.append(u)
            t += 1
        order.extend(comp)
    return np.array(order[::-1], dtype=np.int32)
