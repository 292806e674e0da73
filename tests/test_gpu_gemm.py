"""GPU: the complex MFMA GEMM update C -= A B (FP64 and FP32) against numpy,
over shapes that exercise every lane group and tile edge."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (32, 3, 1), (1, 3, 32), (64, 64, 16), (70, 45, 33),
                                   (130, 129, 64), (5, 200, 7), (9000, 32, 700)])   # last: the library path (FP64)
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


@pytest.mark.parametrize("N", [1, 2, 4])
def test_gemm_sub_few_rhs_batched(gpu, N):
    """FP64 C -= A B with N <= 4 right-hand sides (the refinement residuals:
    one wave per row of A), batched with per-matrix strides, rows longer and
    shorter than a wave."""
    import qpms_amd
    L = qpms_amd.lib()
    rng = np.random.default_rng(40 + N)
    batch, M, K = 3, 777, 1029
    A = rng.normal(size=(batch, M, K)) + 1j * rng.normal(size=(batch, M, K))
    B = rng.normal(size=(batch, K, N)) + 1j * rng.normal(size=(batch, K, N))
    C = rng.normal(size=(batch, M, N)) + 1j * rng.normal(size=(batch, M, N))
    ref = C - A @ B
    dA, dB, dC = (qpms_amd.DeviceBuffer.from_array(x) for x in (A, B, C))
    f = L.qh_gemm_sub_f64
    lg, vp, i32 = ctypes.c_long, ctypes.c_void_p, ctypes.c_int
    f.argtypes = [i32, i32, i32, i32, vp, lg, lg, vp, lg, lg, vp, lg, lg, vp]
    f(M, N, K, batch, dA.ptr, K, M * K, dB.ptr, N, K * N, dC.ptr, N, M * N, None)
    L.qpms_hip_sync()
    out = dC.to_array(np.complex128, (batch, M, N))
    assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) < 1e-13
    # a row shorter than a wave
    A2 = A[0, :5, :7].copy()
    B2 = B[0, :7].copy()
    C2 = C[0, :5].copy()
    d2 = [qpms_amd.DeviceBuffer.from_array(x) for x in (A2, B2, C2)]
    f(5, N, 7, 1, d2[0].ptr, 7, 0, d2[1].ptr, N, 0, d2[2].ptr, N, 0, None)
    L.qpms_hip_sync()
    out2 = d2[2].to_array(np.complex128, (5, N))
    ref2 = C2 - A2 @ B2
    assert np.max(np.abs(out2 - ref2)) / np.max(np.abs(ref2)) < 1e-13


@pytest.mark.parametrize("tile", [0, 1, 2])
@pytest.mark.parametrize("M,N,K,batch", [(1, 1, 1, 1), (32, 32, 2, 1), (64, 64, 16, 2), (70, 45, 33, 1),
                                         (130, 129, 64, 3), (257, 300, 256, 1), (5, 200, 7, 2), (400, 17, 300, 1),
                                         (100, 90, 48, 2), (64, 64, 80, 1)])
def test_cgemm3m_sub(gpu, tile, M, N, K, batch):
    """FP32 complex C -= A B by the three-product (3M) MFMA kernel, every tile
    shape, edges in M, N and K, batched with per-matrix strides and padded
    leading dimensions; K a multiple of 16 (the two-stage prefetch path) with
    an even and an odd number of stages, and K that is not."""
    import qpms_amd
    L = qpms_amd.lib()
    rng = np.random.default_rng(7 * M + 11 * N + 13 * K + tile)
    lda, ldb, ldc = K + 3, N + 5, N + 1
    A = (rng.normal(size=(batch, M, lda)) + 1j * rng.normal(size=(batch, M, lda))).astype(np.complex64)
    B = (rng.normal(size=(batch, K, ldb)) + 1j * rng.normal(size=(batch, K, ldb))).astype(np.complex64)
    C = (rng.normal(size=(batch, M, ldc)) + 1j * rng.normal(size=(batch, M, ldc))).astype(np.complex64)
    ref = C.astype(np.complex128)
    ref[:, :, :N] -= A[:, :, :K].astype(np.complex128) @ B[:, :, :N].astype(np.complex128)
    dA, dB, dC = (qpms_amd.DeviceBuffer.from_array(x) for x in (A, B, C))
    f = L.qh_cgemm3m_sub
    lg, vp, i32 = ctypes.c_long, ctypes.c_void_p, ctypes.c_int
    f.argtypes = [i32, i32, i32, i32, vp, lg, lg, vp, lg, lg, vp, lg, lg, i32, vp]
    f(M, N, K, batch, dA.ptr, lda, M * lda, dB.ptr, ldb, K * ldb, dC.ptr, ldc, M * ldc, tile, None)
    L.qpms_hip_sync()
    out = dC.to_array(np.complex64, (batch, M, ldc))
    assert np.array_equal(out[:, :, N:], C[:, :, N:])   # columns beyond N untouched
    err = np.max(np.abs(out - ref)) / max(1.0, np.max(np.abs(ref)))
    assert err < 4e-6 * max(1, K), err
