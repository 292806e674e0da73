"""GPU: the complex MFMA GEMM update C -= A B (FP64 and FP32) against numpy,
over shapes that exercise every lane group and tile edge."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (32, 3, 1), (1, 3, 32), (64, 64, 16), (70, 45, 33),
                                   (130, 129, 64), (5, 200, 7)])
def test_gemm_sub(gpu, prec, M, N, K):
    import qpms_amd
    L = qpms_amd.lib()
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    dt = np.complex128 if prec == "f64" else np.complex64
    A = (rng.normal(size=(M, K)) + 1j * rng.normal(size=(M, K))).astype(dt)
    B = (rng.normal(size=(K, N)) + 1j * rng.normal(size=(K, N))).astype(dt)
    C = (rng.normal(size=(M, N)) + 1j * rng.normal(size=(M, N))).astype(dt)
    ref = C.astype(np.complex128) - A.astype(np.complex128) @ B.astype(np.complex128)
    dA, dB, dC = (qpms_amd.DeviceBuffer.from_array(x) for x in (A, B, C))
    f = L.qh_gemm_sub_f64 if prec == "f64" else L.qh_gemm_sub_f32
    lg, vp, i32 = ctypes.c_long, ctypes.c_void_p, ctypes.c_int
    f.argtypes = [i32, i32, i32, i32, vp, lg, lg, vp, lg, lg, vp, lg, lg, vp]
    f(M, N, K, 1, dA.ptr, K, 0, dB.ptr, N, 0, dC.ptr, N, 0, None)
    L.qpms_hip_sync()
    out = dC.to_array(dt, (M, N))
    tol = 1e-13 if prec == "f64" else 2e-5
    assert np.max(np.abs(out - ref)) / max(1.0, np.max(np.abs(ref))) < tol * max(1, K)
