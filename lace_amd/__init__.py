"""ctypes binding of liblace_mi355x.so, the MI355X-native LACE hot path.

The engine is C++/HIP behind the C ABI in include/lace.h; this module only
loads the shared library, mirrors the container structs and moves numpy
arrays across the boundary.  It is used by the tests, smoke() and bench.py.
Loading fails loudly when the library has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblace_mi355x.so")
_lib = None

# enum values (include/lace.h, MAGMA numbering)
MagmaRowMajor, MagmaColMajor = 101, 102
MagmaUpper, MagmaLower = 121, 122
Magma_CSR, Magma_DENSE, Magma_CSC = 611, 614, 616
Magma_CSRL, Magma_CSRU, Magma_CSCU, Magma_CSCL = 627, 628, 635, 634
Magma_UNITY, Magma_VALUE, Magma_NODIAG = 593, 594, 596
Magma_BCSR, Magma_BCSRL, Magma_BCSRU, Magma_BCSCU = 615, 641, 642, 646
LACE_F64, LACE_F32, LACE_BF16, LACE_F16 = 0, 1, 2, 3
TRSV_LEVEL, TRSV_JACOBI, TRSV_NONE = 0, 1, 2
DEV_SUCCESS = 0


class DMatrix(C.Structure):
    """data_d_matrix with the reference layout (pointer members 64-byte
    aligned, struct size 384)."""
    _fields_ = [
        ("storage_type", C.c_int), ("sym", C.c_int), ("diagorder_type", C.c_int),
        ("fill_mode", C.c_int), ("num_rows", C.c_int), ("num_cols", C.c_int),
        ("nnz", C.c_int), ("max_nnz_row", C.c_int), ("diameter", C.c_int),
        ("true_nnz", C.c_int), ("_p0", C.c_char * 24),
        ("val", C.POINTER(C.c_double)), ("_p1", C.c_char * 56),
        ("diag", C.POINTER(C.c_double)), ("_p2", C.c_char * 56),
        ("row", C.POINTER(C.c_int)), ("_p3", C.c_char * 56),
        ("rowidx", C.POINTER(C.c_int)), ("_p4", C.c_char * 56),
        ("col", C.POINTER(C.c_int)),
        ("list", C.POINTER(C.c_int)), ("blockinfo", C.POINTER(C.c_int)),
        ("blocksize", C.c_int), ("ldblock", C.c_int), ("numblocks", C.c_int),
        ("alignment", C.c_int), ("major", C.c_int), ("ld", C.c_int),
        ("pad_rows", C.c_int), ("pad_cols", C.c_int), ("_p5", C.c_char * 8),
    ]


assert C.sizeof(DMatrix) == 384 and DMatrix.val.offset == 64 and DMatrix.col.offset == 320


class PrecondLog(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("maxSweeps", C.c_int), ("tol", C.c_double),
                ("finalStep", C.c_double), ("A_Frobenius", C.c_double),
                ("precond_generation_time", C.c_double), ("initial_residual", C.c_double),
                ("initial_nonlinear_residual", C.c_double), ("residual", C.c_double),
                ("nonlinear_residual", C.c_double), ("omp_num_threads", C.c_int),
                ("_pad", C.c_char * 52)]


class GmresLog(C.Structure):
    _fields_ = [("restarts", C.c_int), ("search_directions", C.c_int),
                ("solve_time", C.c_double), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("scaled_residual", C.c_double),
                ("original_residual", C.c_double), ("_pad", C.c_char * 16)]


class GmresParam(C.Structure):
    _fields_ = [("search_max", C.c_int), ("tol_type", C.c_int), ("rtol", C.c_double),
                ("reorth", C.c_int), ("user_csrtrsv_choice", C.c_int),
                ("monitorOrthog", C.c_int), ("restart_max", C.c_int),
                ("precondition", C.c_int), ("parilu_reduction", C.c_double),
                ("reorder", C.c_int), ("_pad", C.c_char * 12)]


class SolverOpts(C.Structure):
    _fields_ = [("krylov", C.c_int), ("search_max", C.c_int), ("restart_max", C.c_int),
                ("rtol", C.c_double), ("prec_precond", C.c_int), ("trsv", C.c_int),
                ("jacobi_sweeps", C.c_int), ("parilu_sweeps", C.c_int),
                ("inner_iters", C.c_int), ("transpose", C.c_int), ("inner_fp64", C.c_int)]


class SolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int), ("inner_iterations", C.c_int),
                ("initial_residual", C.c_double), ("final_residual_est", C.c_double),
                ("true_residual", C.c_double), ("rel_true_residual", C.c_double),
                ("setup_ms", C.c_double), ("solve_ms", C.c_double), ("converged", C.c_int)]


NRHS = 32


class BlockStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int), ("panel_passes", C.c_int),
                ("shifted_panels", C.c_int), ("converged", C.c_int),
                ("max_rel_true_residual", C.c_double),
                ("rel_true_residual", C.c_double * NRHS),
                ("initial_residual", C.c_double * NRHS), ("solve_ms", C.c_double)]


assert C.sizeof(PrecondLog) == 128 and C.sizeof(GmresLog) == 64 and C.sizeof(GmresParam) == 64
assert C.sizeof(SolverOpts) == 56 and C.sizeof(SolverStats) == 72
assert C.sizeof(BlockStats) == 24 + 8 + 2 * 8 * NRHS + 8

_P = C.POINTER
_DM = _P(DMatrix)
_vp = C.c_void_p
_i, _d = C.c_int, C.c_double

# name -> (restype, argtypes); every symbol declared in include/lace.h
SIGNATURES = {
    "data_zvinit": (_i, [_DM, _i, _i, _d, _i]),
    "data_rowindex": (_i, [_DM, _P(_P(_i))]),
    "data_zmconvert": (_i, [DMatrix, _DM, _i, _i]),
    "data_zmtranspose": (_i, [DMatrix, _DM]),
    "data_zmfree": (_i, [_DM]),
    "data_zmlumerge": (_i, [DMatrix, DMatrix, _DM]),
    "data_z_csr_mtx": (_i, [_DM, C.c_char_p]),
    "data_zwrite_csr_mtx": (_i, [DMatrix, _i, C.c_char_p]),
    "data_dnrm2": (_d, [_i, _P(_d), _i]),
    "data_zdot": (_d, [_i, _P(_d), _i, _P(_d), _i]),
    "data_zdot_mkl": (_d, [_i, _P(_d), _i, _P(_d), _i]),
    "data_z_dense_mtx": (_i, [_DM, _i, C.c_char_p]),
    "data_zwrite_dense_mtx": (_i, [_i, _i, _i, _i, _P(_P(_d)), C.c_char_p]),
    "data_zwrite_dense": (_i, [DMatrix, C.c_char_p]),
    "data_zaxpy": (None, [_i, _d, _P(_d), _i, _P(_d), _i]),
    "data_zfrobenius": (_i, [DMatrix, _P(_d)]),
    "data_zilures": (_i, [DMatrix, DMatrix, DMatrix, _DM, _P(_d), _P(_d)]),
    "data_z_spmv": (_i, [_d, _DM, _DM, _d, _DM]),
    "data_z_spmm": (_i, [_d, DMatrix, DMatrix, _DM]),
    "data_dcsrilu0_mkl": (_i, [_DM]),
    "data_PariLU_v0_3": (None, [_DM, _DM, _DM, _d, _P(PrecondLog)]),
    "data_PariLU_v0_3_prescribedSweeps": (None, [_DM, _DM, _DM, _d, _P(PrecondLog)]),
    "data_PariLU_v0_3_gpu": (None, [_DM, _DM, _DM, _d, _P(PrecondLog), _i]),
    "data_PariLU_v0_3_gpu_prescribedSweeps": (None, [_DM, _DM, _DM, _d, _P(PrecondLog)]),
    "data_PariLU_v0_3_bcsr": (_i, [_DM, _DM, _DM, _d, _P(PrecondLog)]),
    "data_forward_solve": (_i, [_DM, _DM, _DM, _d, _P(_i)]),
    "data_backward_solve": (_i, [_DM, _DM, _DM, _d, _P(_i)]),
    "data_parcsrtrsv": (_i, [_i, _i, _i, _i, _P(_d), _P(_i), _P(_i), _P(_d), _P(_d), _d, _P(_i)]),
    "lace_parcsrtrsv_noalloc": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _d, _vp]),
    "data_fgmres": (_i, [_DM, _DM, _DM, _DM, _DM, _P(GmresParam), _P(GmresLog)]),
    "data_fgmres_restart": (_i, [_DM, _DM, _DM, _DM, _DM, _P(GmresParam), _P(GmresLog)]),
    "data_fgmres_householder": (_i, [_DM, _DM, _DM, _DM, _DM, _P(GmresParam), _P(GmresLog)]),
    "data_fgmres_householder_restart": (_i, [_DM, _DM, _DM, _DM, _DM, _P(GmresParam),
                                             _P(GmresLog)]),
    "givens_rotation": (None, [_d, _d, _d, _P(_d), _P(_d)]),
    "data_housegen": (None, [_i, _P(_d), _P(_d), _P(_d)]),
    "data_hqrd": (None, [_DM, _DM, _DM]),
    "data_PariLU_v0_2": (None, [_DM, _DM, _DM, _d]),
    **{f"data_PariLU_v3_{v}": (None, [_DM, _DM, _DM, _i]) for v in (0, 1, 2, 5, 6, 7, 8, 9)},
    "data_MKL_FGMRES": (_i, [_DM, _DM, _DM, _P(GmresParam)]),
    "data_reorder_cuthill_mckee": (_i, [_DM, _P(_i), C.c_bool]),
    "data_sparse_reorder": (_i, [_DM, _P(_i), _P(_i), _i]),
    "lace_reorder_rcm": (_i, [_DM, _P(_i)]),
    "lace_init": (_i, [_i]),
    "lace_last_error": (C.c_char_p, []),
    "lace_device_synchronize": (_i, []),
    "lace_dmalloc": (_i, [_P(_vp), C.c_size_t]),
    "lace_dfree": (_i, [_vp]),
    "lace_memcpy_h2d": (_i, [_vp, _vp, C.c_size_t]),
    "lace_memcpy_d2h": (_i, [_vp, _vp, C.c_size_t]),
    "lace_stream_copy": (_i, [_vp, _vp, C.c_size_t]),
    "lace_csr_create": (_i, [_DM, _i, _P(_vp)]),
    "lace_csr_destroy": (_i, [_vp]),
    "lace_csr_n": (_i, [_vp]),
    "lace_csr_spmv": (_i, [_vp, _d, _vp, _d, _vp]),
    "lace_csr_forms": (_i, [_vp, _P(_i), _P(_i)]),
    "lace_parilu_create": (_i, [_DM, _i, _P(_vp)]),
    "lace_parilu_destroy": (_i, [_vp]),
    "lace_parilu_reset": (_i, [_vp]),
    "lace_parilu_set_values": (_i, [_vp, _P(_d)]),
    "lace_parilu_sweeps": (_i, [_vp, _i, _P(_d)]),
    "lace_parilu_get": (_i, [_vp, _DM, _DM]),
    "lace_parilu_nonlinres": (_i, [_vp, _P(_d)]),
    "lace_parilu_sweep_bytes": (_d, [_vp]),
    "lace_parilu_apply": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "lace_parilu_levels": (_i, [_vp, _P(_i), _P(_i)]),
    "lace_parilu_forms": (_i, [_vp, _i, _P(_i), _P(_i)]),
    "lace_parilu_tiles": (_i, [_vp, _i, _P(_i)]),
    "lace_parilu_set_variant": (_i, [_vp, _i]),
    "lace_syncfree_error": (_i, [_vp, _P(_i)]),
    "lace_bcsr_create": (_i, [_DM, _i, _P(_vp)]),
    "lace_bcsr_destroy": (_i, [_vp]),
    "lace_bcsr_spmv": (_i, [_vp, _vp, _vp]),
    "lace_bcsr_spmv_bytes": (_d, [_vp]),
    "lace_bilu_create": (_i, [_DM, _i, _P(_vp)]),
    "lace_bilu_destroy": (_i, [_vp]),
    "lace_bilu_reset": (_i, [_vp]),
    "lace_bilu_sweeps": (_i, [_vp, _i, _P(_d)]),
    "lace_bilu_get": (_i, [_vp, _DM, _DM]),
    "lace_bilu_sweep_bytes": (_d, [_vp]),
    "lace_bilu_apply": (_i, [_vp, _i, _i, _vp, _vp]),
    "lace_bilu_levels": (_i, [_vp, _P(_i), _P(_i)]),
    "lace_bilu_set_variant": (_i, [_vp, _i]),
    "lace_bilu_is_symmetric": (_i, [_vp]),
    "lace_solver_default_opts": (None, [_P(SolverOpts)]),
    "lace_solver_create": (_i, [_DM, _P(SolverOpts), _P(_vp)]),
    "lace_solver_destroy": (_i, [_vp]),
    "lace_solver_setup": (_i, [_vp]),
    "lace_solver_solve": (_i, [_vp, _vp, _vp, _P(SolverStats)]),
    "lace_solver_grad_values": (_i, [_vp, _vp, _vp, _vp]),
    "lace_solver_solve_trans": (_i, [_vp, _vp, _vp, _i, _P(SolverStats)]),
    "lace_solver_set_values": (_i, [_vp, _P(_d)]),
    "lace_set_option": (_i, [C.c_char_p, _i]),
    "lace_solver_solve_block": (_i, [_vp, _vp, _vp, _i, _P(BlockStats)]),
    "lace_solver_grad_values_block": (_i, [_vp, _vp, _vp, _vp]),
    "lace_csr_spmm32": (_i, [_vp, _vp, _i, _vp, _vp]),
    "lace_parilu_apply_block": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "lace_block_panel_qr": (_i, [_i, _i, _vp, _vp, _vp, _vp, _P(_i)]),
    "lace_daxpby": (_i, [C.c_size_t, _d, _vp, _d, _vp, _vp]),
    "lace_tsgemm_tn": (_i, [_i, _i, _P(_vp), _vp, _i, _i, _vp]),
    "lace_tsgemm_nn": (_i, [_i, _i, _P(_vp), _vp, _vp, _i, _i, _i, _vp]),
    "lace_timing_enable": (_i, [_i]),
    "lace_timing_reset": (_i, []),
    "lace_timing_only": (_i, [C.c_char_p]),
    "lace_timing_get": (_i, [C.c_char_p, _P(_d), _P(C.c_long)]),
    "lace_timing_get_work": (_i, [C.c_char_p, _P(_d), _P(_d)]),
    "lace_timing_get_calls": (_i, [C.c_char_p, _P(C.c_long)]),
    "lace_level_sets": (_i, [_i, _P(_i), _P(_i), _i, _P(_i), _P(_i)]),
    "lace_gen_laplace2d": (_i, [_i, _i, _DM]),
    "lace_gen_convdiff3d": (_i, [_i, _i, _i, _d, _d, _d, _DM]),
    "lace_gen_bcsr7": (_i, [_i, _i, _i, _i, C.c_uint64, _d, _d, _DM]),
    "lace_gen_normal": (_i, [C.c_uint64, _i, _P(_d)]),
    "lace_philox4x32_10": (_P(C.c_uint32), [_P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint32)]),
}


class LaceError(RuntimeError):
    pass


def lib():
    """Load liblace_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LaceError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != DEV_SUCCESS:
        msg = lib().lace_last_error().decode(errors="replace")
        raise LaceError(f"{what} failed with {rc}: {msg}")


# ------------------------------------------------------------ host matrices
def _iptr(a):
    return a.ctypes.data_as(_P(_i))


def _dptr(a):
    return a.ctypes.data_as(_P(_d))


class HostCSR:
    """data_d_matrix viewing numpy arrays owned by this object (inputs)."""

    def __init__(self, row, col, val, n_cols=None, storage=Magma_CSR):
        # private copies: in-place routines (data_dcsrilu0_mkl) must not
        # write through to the caller's arrays
        self.row = np.array(row, dtype=np.int32, copy=True, order="C")
        self.col = np.array(col, dtype=np.int32, copy=True, order="C")
        self.val = np.array(val, dtype=np.float64, copy=True, order="C")
        n = len(self.row) - 1
        m = DMatrix()
        m.storage_type = storage
        m.sym = 581
        m.diagorder_type = Magma_VALUE
        m.fill_mode = 123
        m.num_rows = n
        m.num_cols = n if n_cols is None else n_cols
        m.nnz = int(self.row[-1])
        m.true_nnz = m.nnz
        m.major = MagmaRowMajor
        m.row = _iptr(self.row)
        m.col = _iptr(self.col)
        m.val = _dptr(self.val)
        self.m = m

    @classmethod
    def from_csr(cls, A):
        return cls(A.row, A.col, A.val, A.n_cols)

    def ref(self):
        return C.byref(self.m)

    def release_rowidx(self):
        """Free a rowidx array the library attached (ParILU does, as the
        reference does, src/parilu_v0_3.cpp:61)."""
        if self.m.rowidx:
            C.CDLL(None).free(C.cast(self.m.rowidx, C.c_void_p))
            self.m.rowidx = C.POINTER(C.c_int)()


class HostBCSR:
    """BCSR data_d_matrix over numpy arrays owned by this object: num_rows
    block rows, blocks of bs x bs stored row-major (val has shape
    (numblocks, bs, bs) or is flat)."""

    def __init__(self, row, col, val, bs: int):
        self.row = np.array(row, dtype=np.int32, copy=True, order="C")
        self.col = np.array(col, dtype=np.int32, copy=True, order="C")
        self.val = np.array(val, dtype=np.float64, copy=True, order="C").reshape(-1)
        nbr = len(self.row) - 1
        nb = int(self.row[-1])
        assert self.val.size == nb * bs * bs
        m = DMatrix()
        m.storage_type = Magma_BCSR
        m.sym = 581
        m.diagorder_type = Magma_VALUE
        m.fill_mode = 123
        m.num_rows = m.num_cols = nbr
        m.blocksize = bs
        m.ldblock = bs * bs
        m.numblocks = nb
        m.nnz = m.true_nnz = nb * bs * bs
        m.major = MagmaRowMajor
        m.row = _iptr(self.row)
        m.col = _iptr(self.col)
        m.val = _dptr(self.val)
        self.m = m
        self.bs = bs
        self.n = nbr * bs

    def ref(self):
        return C.byref(self.m)


def owned_bcsr_to_numpy(m: DMatrix):
    """(row, col, val[numblocks, bs, bs]) of a library-allocated block
    matrix (BCSR / BCSRL / BCSCU), freed afterwards."""
    n, bs = m.num_rows, m.blocksize
    nb = m.row[n]
    row = np.ctypeslib.as_array(m.row, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(m.col, shape=(max(nb, 1),))[:nb].copy()
    val = np.ctypeslib.as_array(m.val, shape=(max(nb, 1) * bs * bs,))[:nb * bs * bs].copy()
    meta = dict(storage=m.storage_type, bs=bs, numblocks=m.numblocks, major=m.major)
    lib().data_zmfree(C.byref(m))
    return row, col, val.reshape(nb, bs, bs), meta


class HostVec:
    """Dense n x 1 data_d_matrix over a numpy float64 array."""

    def __init__(self, v):
        self.v = np.ascontiguousarray(v, dtype=np.float64).copy()
        m = DMatrix()
        m.storage_type = Magma_DENSE
        m.num_rows = len(self.v)
        m.num_cols = 1
        m.nnz = len(self.v)
        m.ld = len(self.v)
        m.major = MagmaColMajor
        m.val = _dptr(self.v)
        self.m = m

    def ref(self):
        return C.byref(self.m)


def owned_to_numpy(m: DMatrix, kind="csr"):
    """Copy a library-allocated matrix out and free it with data_zmfree."""
    n = m.num_rows
    if kind == "csr":
        nnz = m.row[n]
        row = np.ctypeslib.as_array(m.row, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(m.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
        val = np.ctypeslib.as_array(m.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
        out = (row, col, val)
    else:
        cnt = m.num_rows * m.num_cols
        out = np.ctypeslib.as_array(m.val, shape=(cnt,)).copy()
    lib().data_zmfree(C.byref(m))
    return out


# ------------------------------------------------------------ device vectors
class DevVec:
    """Device buffer allocated by the library (lace_dmalloc)."""

    def __init__(self, n: int, dtype=np.float64, data=None):
        self.n = n
        self.dtype = np.dtype(dtype)
        self.ptr = C.c_void_p()
        check(lib().lace_dmalloc(C.byref(self.ptr), max(1, n) * self.dtype.itemsize), "dmalloc")
        if data is not None:
            self.upload(data)

    def upload(self, data):
        a = np.ascontiguousarray(data, dtype=self.dtype)
        assert a.size == self.n
        check(lib().lace_memcpy_h2d(self.ptr, a.ctypes.data_as(_vp), a.nbytes), "h2d")

    def download(self):
        a = np.empty(self.n, dtype=self.dtype)
        check(lib().lace_memcpy_d2h(a.ctypes.data_as(_vp), self.ptr, a.nbytes), "d2h")
        return a

    def free(self):
        if self.ptr:
            lib().lace_dfree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gpu_available() -> bool:
    """True if a gfx950 device can be initialised (counts devices without
    creating a GPU context beforehand)."""
    try:
        return lib().lace_init(0) == DEV_SUCCESS
    except LaceError:
        return False


# ------------------------------------------------------------ handles
class ParILU:
    def __init__(self, A: HostCSR, prec=LACE_F64):
        self.h = C.c_void_p()
        check(lib().lace_parilu_create(A.ref(), prec, C.byref(self.h)), "parilu_create")
        self.prec = prec
        self.n = A.m.num_rows

    def reset(self):
        check(lib().lace_parilu_reset(self.h), "parilu_reset")

    def set_values(self, vals):
        v = np.ascontiguousarray(vals, dtype=np.float64)
        check(lib().lace_parilu_set_values(self.h, _dptr(v)), "parilu_set_values")

    def sweeps(self, k: int, want_step=False):
        st = C.c_double()
        check(lib().lace_parilu_sweeps(self.h, k, C.byref(st) if want_step else None), "sweeps")
        return st.value if want_step else None

    def factors(self):
        L = DMatrix(); U = DMatrix()
        check(lib().lace_parilu_get(self.h, C.byref(L), C.byref(U)), "parilu_get")
        return owned_to_numpy(L), owned_to_numpy(U)

    def nonlinres(self):
        v = C.c_double()
        check(lib().lace_parilu_nonlinres(self.h, C.byref(v)), "nonlinres")
        return v.value

    def sweep_bytes(self):
        return lib().lace_parilu_sweep_bytes(self.h)

    def levels(self):
        a = C.c_int(); b = C.c_int()
        check(lib().lace_parilu_levels(self.h, C.byref(a), C.byref(b)), "levels")
        return a.value, b.value

    def forms(self, upper: bool):
        """(diagonals, ELLPACK width) of L's or U's off-diagonal entries
        (0 where that form does not apply)."""
        d = C.c_int(); w = C.c_int()
        check(lib().lace_parilu_forms(self.h, int(upper), C.byref(d), C.byref(w)), "forms")
        return d.value, w.value

    def tiles(self, which: int):
        """Box sides of the exact solve's block order for L (0), U (1),
        U^T (2), L^T (3); (0, 0, 0) for blocks of consecutive rows."""
        d = (C.c_int * 3)()
        check(lib().lace_parilu_tiles(self.h, int(which), d), "tiles")
        return tuple(d)

    def apply(self, r: np.ndarray, kind=TRSV_LEVEL, sweeps=5, trans=0):
        dt = np.float64 if self.prec == LACE_F64 else np.float32
        dr = DevVec(self.n, dt, r)
        dz = DevVec(self.n, dt)
        check(lib().lace_parilu_apply(self.h, kind, sweeps, trans, dr.ptr, dz.ptr), "apply")
        err = C.c_int()
        check(lib().lace_syncfree_error(self.h, C.byref(err)), "syncfree_error")
        if err.value:
            raise LaceError("sync-free solve exceeded its poll budget")
        return dz.download()

    def apply_block(self, R: np.ndarray, kind=TRSV_JACOBI, sweeps=5, trans=0):
        """Z = M^{-1} R (or M^{-T} R) for an (n, 32) block in the factor
        precision."""
        dt = np.float64 if self.prec == LACE_F64 else np.float32
        R = np.ascontiguousarray(R, dtype=dt)
        dr = DevVec(R.size, dt, R.reshape(-1))
        dz = DevVec(R.size, dt)
        check(lib().lace_parilu_apply_block(self.h, kind, sweeps, trans, dr.ptr, dz.ptr),
              "apply_block")
        return dz.download().reshape(R.shape)

    def __del__(self):
        try:
            if self.h:
                lib().lace_parilu_destroy(self.h)
        except Exception:
            pass


class DevCSRMatrix:
    """Device CSR matrix (lace_csr) for the product kernels."""

    def __init__(self, A: HostCSR, prec=LACE_F64):
        self.h = C.c_void_p()
        check(lib().lace_csr_create(A.ref(), prec, C.byref(self.h)), "csr_create")
        self.n = A.m.num_rows
        self.prec = prec

    def forms(self):
        """(diagonals, ELLPACK width) held besides CSR (0: not applicable)."""
        d = C.c_int(); w = C.c_int()
        check(lib().lace_csr_forms(self.h, C.byref(d), C.byref(w)), "csr_forms")
        return d.value, w.value

    def spmv(self, x: np.ndarray, alpha=1.0, beta=0.0, y=None) -> np.ndarray:
        """alpha A x + beta y on device vectors of the matrix precision."""
        dt = np.float64 if self.prec == LACE_F64 else np.float32
        dx = DevVec(self.n, dt, np.asarray(x, dt))
        dy = DevVec(self.n, dt, np.zeros(self.n, dt) if y is None else np.asarray(y, dt))
        check(lib().lace_csr_spmv(self.h, alpha, dx.ptr, beta, dy.ptr), "csr_spmv")
        return dy.download()

    def spmm32(self, X: np.ndarray, B=None) -> np.ndarray:
        """A X, or B - A X, for (n, 32) blocks; X FP64 or FP32."""
        xp = LACE_F64 if X.dtype == np.float64 else LACE_F32
        X = np.ascontiguousarray(X)
        dx = DevVec(X.size, X.dtype, X.reshape(-1))
        dy = DevVec(self.n * NRHS, np.float64)
        db = None
        if B is not None:
            db = DevVec(self.n * NRHS, np.float64, np.ascontiguousarray(B, np.float64).reshape(-1))
        check(lib().lace_csr_spmm32(self.h, dx.ptr, xp, db.ptr if db is not None else None,
                                    dy.ptr), "spmm32")
        return dy.download().reshape(self.n, NRHS)

    def __del__(self):
        try:
            if self.h:
                lib().lace_csr_destroy(self.h)
        except Exception:
            pass


def block_panel_qr(W: np.ndarray, r0: int):
    """Householder QR of rows r0.. of an (n, 32) FP64 block on the device:
    returns (Y, Tb, Rt, passes)."""
    n = W.shape[0]
    W = np.ascontiguousarray(W, np.float64)
    dW = DevVec(W.size, np.float64, W.reshape(-1))
    dY = DevVec(W.size, np.float64)
    dT = DevVec(1024, np.float64)
    dR = DevVec(1024, np.float64)
    p = C.c_int()
    check(lib().lace_block_panel_qr(n, r0, dW.ptr, dY.ptr, dT.ptr, dR.ptr, C.byref(p)), "panel_qr")
    return (dY.download().reshape(n, NRHS), dT.download().reshape(32, 32),
            dR.download().reshape(32, 32), p.value)


class DevBCSR:
    """Device BCSR matrix (lace_bcsr)."""

    def __init__(self, A: HostBCSR, prec=LACE_F64):
        self.h = C.c_void_p()
        check(lib().lace_bcsr_create(A.ref(), prec, C.byref(self.h)), "bcsr_create")
        self.n = A.n
        self.prec = prec

    def spmv(self, x: np.ndarray) -> np.ndarray:
        dt = np.float64 if self.prec == LACE_F64 else np.float32
        dx = DevVec(self.n, dt, x)
        dy = DevVec(self.n, dt)
        check(lib().lace_bcsr_spmv(self.h, dx.ptr, dy.ptr), "bcsr_spmv")
        return dy.download()

    def bytes(self):
        return lib().lace_bcsr_spmv_bytes(self.h)

    def __del__(self):
        try:
            if self.h:
                lib().lace_bcsr_destroy(self.h)
        except Exception:
            pass


class BILU:
    """Device block ParILU(0) (lace_bilu)."""

    def __init__(self, A: HostBCSR, prec=LACE_F32):
        self.h = C.c_void_p()
        check(lib().lace_bilu_create(A.ref(), prec, C.byref(self.h)), "bilu_create")
        self.prec = prec
        self.n = A.n

    def reset(self):
        check(lib().lace_bilu_reset(self.h), "bilu_reset")

    def sweeps(self, k: int, want_step=False):
        st = C.c_double()
        check(lib().lace_bilu_sweeps(self.h, k, C.byref(st) if want_step else None), "bilu_sweeps")
        return st.value if want_step else None

    def factors(self):
        """(L as BCSRL, U as BCSCU), each (row, col, val[nb, bs, bs], meta)."""
        L = DMatrix(); U = DMatrix()
        check(lib().lace_bilu_get(self.h, C.byref(L), C.byref(U)), "bilu_get")
        return owned_bcsr_to_numpy(L), owned_bcsr_to_numpy(U)

    def sweep_bytes(self):
        return lib().lace_bilu_sweep_bytes(self.h)

    def set_variant(self, v: int):
        check(lib().lace_bilu_set_variant(self.h, v), "bilu_set_variant")

    def symmetric(self) -> bool:
        return bool(lib().lace_bilu_is_symmetric(self.h))

    def levels(self):
        a = C.c_int(); b = C.c_int()
        check(lib().lace_bilu_levels(self.h, C.byref(a), C.byref(b)), "bilu_levels")
        return a.value, b.value

    def apply(self, r: np.ndarray, kind=TRSV_LEVEL, sweeps=5):
        dt = np.float64 if self.prec == LACE_F64 else np.float32
        dr = DevVec(self.n, dt, r)
        dz = DevVec(self.n, dt)
        check(lib().lace_bilu_apply(self.h, kind, sweeps, dr.ptr, dz.ptr), "bilu_apply")
        return dz.download()

    def __del__(self):
        try:
            if self.h:
                lib().lace_bilu_destroy(self.h)
        except Exception:
            pass


def default_opts() -> SolverOpts:
    o = SolverOpts()
    lib().lace_solver_default_opts(C.byref(o))
    return o


class Solver:
    """FGMRES solver handle over a HostCSR or a HostBCSR operator."""

    def __init__(self, A, opts: SolverOpts):
        self.h = C.c_void_p()
        check(lib().lace_solver_create(A.ref(), C.byref(opts), C.byref(self.h)), "solver_create")
        self.n = A.n if isinstance(A, HostBCSR) else A.m.num_rows
        self.nnz = A.m.nnz
        self.opts = opts

    def setup(self):
        check(lib().lace_solver_setup(self.h), "solver_setup")

    def solve_dev(self, db: DevVec, dx: DevVec) -> SolverStats:
        st = SolverStats()
        check(lib().lace_solver_solve(self.h, db.ptr, dx.ptr, C.byref(st)), "solver_solve")
        return st

    def solve_dev_trans(self, db: DevVec, dx: DevVec, trans: int) -> SolverStats:
        st = SolverStats()
        check(lib().lace_solver_solve_trans(self.h, db.ptr, dx.ptr, trans, C.byref(st)),
              "solver_solve_trans")
        return st

    def solve(self, b, x0=None, trans=0):
        db = DevVec(self.n, np.float64, b)
        dx = DevVec(self.n, np.float64, np.zeros(self.n) if x0 is None else x0)
        st = self.solve_dev_trans(db, dx, trans)
        return dx.download(), st

    def set_values(self, vals):
        v = np.ascontiguousarray(vals, dtype=np.float64)
        check(lib().lace_solver_set_values(self.h, _dptr(v)), "solver_set_values")

    def grad_values(self, lam: DevVec, x: DevVec, nnz: int) -> np.ndarray:
        g = DevVec(nnz, np.float64)
        check(lib().lace_solver_grad_values(self.h, lam.ptr, x.ptr, g.ptr), "grad_values")
        return g.download()

    # 32 right-hand sides at once: blocks are (n, 32) arrays, row-major
    def solve_block_dev(self, dB: DevVec, dX: DevVec, trans=0) -> BlockStats:
        st = BlockStats()
        check(lib().lace_solver_solve_block(self.h, dB.ptr, dX.ptr, trans, C.byref(st)),
              "solver_solve_block")
        return st

    def solve_block(self, B, X0=None, trans=0):
        B = np.ascontiguousarray(B, dtype=np.float64)
        assert B.shape == (self.n, NRHS)
        dB = DevVec(B.size, np.float64, B.reshape(-1))
        X0 = np.zeros_like(B) if X0 is None else np.ascontiguousarray(X0, dtype=np.float64)
        dX = DevVec(B.size, np.float64, X0.reshape(-1))
        st = self.solve_block_dev(dB, dX, trans)
        return dX.download().reshape(self.n, NRHS), st

    def grad_values_block(self, dLam: DevVec, dX: DevVec, nnz: int) -> np.ndarray:
        g = DevVec(nnz, np.float64)
        check(lib().lace_solver_grad_values_block(self.h, dLam.ptr, dX.ptr, g.ptr),
              "grad_values_block")
        return g.download()

    def __del__(self):
        try:
            if self.h:
                lib().lace_solver_destroy(self.h)
        except Exception:
            pass


# ------------------------------------------------------------ generators
def gen_convdiff3d(nx, ny, nz, pe=(2.0, 1.0, 0.5)):
    m = DMatrix()
    check(lib().lace_gen_convdiff3d(nx, ny, nz, pe[0], pe[1], pe[2], C.byref(m)), "gen")
    return owned_to_numpy(m)


def gen_laplace2d(nx, ny):
    m = DMatrix()
    check(lib().lace_gen_laplace2d(nx, ny, C.byref(m)), "gen")
    return owned_to_numpy(m)


def gen_bcsr7(nx, ny, nz, bs=5, seed=1, sigma=0.1, shift=5.0):
    """(row, col, val[nb, bs, bs]) of the 7-point block stencil matrix."""
    m = DMatrix()
    check(lib().lace_gen_bcsr7(nx, ny, nz, bs, seed, sigma, shift, C.byref(m)), "gen_bcsr7")
    row, col, val, _ = owned_bcsr_to_numpy(m)
    return row, col, val


def gen_normal(seed: int, n: int) -> np.ndarray:
    out = np.empty(n)
    check(lib().lace_gen_normal(seed, n, _dptr(out)), "gen_normal")
    return out


def timing(name: str):
    t = C.c_double(); c = C.c_long()
    lib().lace_timing_get(name.encode(), C.byref(t), C.byref(c))
    return t.value, c.value


def timing_calls(name: str) -> int:
    """Entries of a region while timing was on (timed or skipped by the
    "timing_sample" option)."""
    c = C.c_long()
    lib().lace_timing_get_calls(name.encode(), C.byref(c))
    return c.value


def timing_work(name: str):
    """(flops, bytes) summed over the timed launches of a region."""
    f = C.c_double(); b = C.c_double()
    lib().lace_timing_get_work(name.encode(), C.byref(f), C.byref(b))
    return f.value, b.value
