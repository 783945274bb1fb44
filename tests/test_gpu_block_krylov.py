"""Multi-RHS (block) Krylov path on the GPU: the tall-skinny FP64 MFMA
products against numpy (this also pins the MFMA operand layout)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _blocks(gpu, arrs):
    devs = [gpu.DevVec(a.size, np.float64, a.reshape(-1)) for a in arrs]
    ptrs = (C.c_void_p * len(devs))(*[d.ptr.value for d in devs])
    return devs, ptrs


@pytest.mark.parametrize("n,nb", [(1000, 1), (4099, 5), (20000, 9)])
def test_tsgemm_tn(gpu, n, nb):
    rng = np.random.default_rng(n + nb)
    X = [rng.standard_normal((n, 32)) for _ in range(nb)]
    W = rng.standard_normal((n, 32))
    devs, ptrs = _blocks(gpu, X)
    dW = gpu.DevVec(n * 32, np.float64, W.reshape(-1))
    dC = gpu.DevVec(nb * 1024, np.float64)
    rs_main, rs_last = 37, 5
    gpu.check(gpu.lib().lace_tsgemm_tn(n, nb, ptrs, dW.ptr, rs_main, rs_last, dC.ptr))
    Cg = dC.download().reshape(nb, 32, 32)
    for b in range(nb):
        rs = rs_last if b == nb - 1 else rs_main
        want = X[b][rs:].T @ W[rs:]
        assert np.allclose(Cg[b], want, rtol=1e-12, atol=1e-10), b


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_tsgemm_nn(gpu, mode):
    n, nb, rlo, e0 = 5003, 4, 70, 96
    rng = np.random.default_rng(mode)
    X = [rng.standard_normal((n, 32)) for _ in range(nb)]
    D = rng.standard_normal((nb, 32, 32))
    base = rng.standard_normal((n, 32))
    devs, ptrs = _blocks(gpu, X)
    dD = gpu.DevVec(nb * 1024, np.float64, D.reshape(-1))
    dB = gpu.DevVec(n * 32, np.float64, base.reshape(-1))
    dO = gpu.DevVec(n * 32, np.float64)
    gpu.check(gpu.lib().lace_tsgemm_nn(n, nb, ptrs, dD.ptr, dB.ptr, mode, e0, rlo, dO.ptr))
    out = dO.download().reshape(n, 32)
    if mode == 0:
        b0 = np.zeros((n, 32))
    elif mode == 1:
        b0 = base.copy()
    else:
        b0 = np.zeros((n, 32)); b0[e0:e0 + 32] = np.eye(32)
    want = b0.copy()
    want[rlo:] -= sum(X[b][rlo:] @ D[b] for b in range(nb))
    assert np.allclose(out, want, rtol=1e-12, atol=1e-10)
    # in place on the base block
    gpu.check(gpu.lib().lace_tsgemm_nn(n, nb, ptrs, dD.ptr, dB.ptr, 1, 0, rlo, dB.ptr))
    if mode == 1:
        assert np.allclose(dB.download().reshape(n, 32), want, rtol=1e-12, atol=1e-10)
