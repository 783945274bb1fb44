"""The kernel-timing regions that bench.py's live roofline figures come
from: with "timing_sample" k a region times its first four entries after a
reset and then every k-th, counts every entry, and sums the algorithmic
bytes of the timed launches only."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_sampled_region_counts(gpu, orc):
    lib = gpu.lib()
    A = orc.convdiff3d(16, 16, 16)
    M = gpu.DevCSRMatrix(gpu.HostCSR.from_csr(A), gpu.LACE_F64)
    x = np.random.default_rng(5).standard_normal(A.n_rows)
    M.spmv(x)  # first use outside the timings
    try:
        gpu.check(lib.lace_timing_reset())
        gpu.check(lib.lace_set_option(b"timing_sample", 16))
        gpu.check(lib.lace_timing_only(b"spmv"))
        gpu.check(lib.lace_timing_enable(1))
        for _ in range(40):
            M.spmv(x)
        gpu.check(lib.lace_timing_enable(0))
        total_ms, timed = gpu.timing("spmv")
        assert gpu.timing_calls("spmv") == 40
        assert timed == 6  # entries 0-3, 16 and 32
        assert total_ms > 0.0
        _, nbytes = gpu.timing_work("spmv")
        per = nbytes / timed
        # the FP64 diagonal form: 7 diagonals of values, a mask byte, x and y
        n = A.n_rows
        assert per == 7 * 8 * n + n + 8 * n + 8 * n
    finally:
        gpu.check(lib.lace_timing_enable(0))
        gpu.check(lib.lace_timing_only(None))
        gpu.check(lib.lace_set_option(b"timing_sample", 1))
        gpu.check(lib.lace_timing_reset())
