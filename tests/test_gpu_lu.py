"""GPU: the blocked LU (tall-strip panels of 32 columns, rank-256 trailing
updates) in FP64 and FP32, single and batched, across the outer-block and
strip boundaries, against P A = L U, numpy solves and zgetrf's info.

qpms_hip_zgetrf_dev (the factorisation the ABI returns, scatsystem.c:2214)
uses partial pivoting: its interchanges equal host LAPACK's zgetrf (Intel MKL
where present) and the oracle's unblocked zgetrf on generic matrices, and its
factors agree to rounding. The batched LU (serving, Beyn nodes) uses
tournament pivoting (strips of more than 8 x 256 rows go through one- and
two-level merges); its factors satisfy P A = L U with their own pivots."""
import ctypes

import numpy as np
import pytest

import oracle_ffi

pytestmark = pytest.mark.gpu


def rand(n, seed, shift=0.0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + shift * np.eye(n)


def check_plu(A, LU, piv, tol, nprobe=4):
    """||L U x - P A x|| / (||A|| ||x||) on random probe vectors (O(n^2))."""
    n = A.shape[0]
    rng = np.random.default_rng(7)
    X = rng.normal(size=(n, nprobe)) + 1j * rng.normal(size=(n, nprobe))
    PA = A.copy()
    for i in range(n):
        p = piv[i] - 1
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    lhs = np.tril(LU, -1) @ (np.triu(LU) @ X) + np.triu(LU) @ X
    rhs = PA @ X
    err = np.linalg.norm(lhs - rhs) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert err < tol, err


@pytest.mark.parametrize("n", [63, 64, 65, 255, 256, 257, 511, 512, 513, 1500, 3000, 8700])
def test_zgetrf_blocks(gpu, n):
    import qpms_amd
    A = rand(n, n)
    LU, piv, info = qpms_amd.zgetrf(A)
    assert info == 0
    assert np.all((piv >= np.arange(1, n + 1)) & (piv <= n))
    check_plu(A, LU, piv, 1e-14 * np.sqrt(n))
    if n <= 1500:
        b = np.random.default_rng(1).normal(size=(n, 2)) + 0j
        x = qpms_amd.zgetrs(LU, piv, b)
        xo = np.linalg.solve(A, b)
        assert np.max(np.abs(x - xo)) / np.max(np.abs(xo)) < 1e-9


def test_zgetrf_pivots_match_lapack_when_unambiguous(gpu):
    """Diagonally weighted matrix whose pivot choices have clear margins:
    the interchanges equal the oracle's zgetrf."""
    import qpms_amd
    n = 300
    rng = np.random.default_rng(3)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    perm = rng.permutation(n)
    A[perm, np.arange(n)] += 50 * n   # one dominant entry per column, rows permuted
    LU, piv, info = qpms_amd.zgetrf(A)
    LUo, pivo, infoo = oracle_ffi.zgetrf(A)
    assert info == infoo == 0
    assert np.array_equal(piv, pivo)
    assert np.max(np.abs(LU - LUo)) / np.max(np.abs(LUo)) < 1e-12


@pytest.mark.parametrize("n", [31, 257, 700, 2100])
def test_zgetrf_pivots_equal_lapack(gpu, n):
    """Generic random matrices (no designed margins): the interchanges are
    zgetrf's, element for element, against host LAPACK and the oracle; the
    factors agree to 1e-11 of their largest entry."""
    import qpms_amd
    A = rand(n, 900 + n)
    LU, piv, info = qpms_amd.zgetrf(A)
    LUh, pivh, infoh = oracle_ffi.host_lapack().zgetrf(A.copy())
    LUo, pivo, infoo = oracle_ffi.zgetrf(A, nthreads=8)
    assert info == infoh == infoo == 0
    assert np.array_equal(pivh, pivo)
    first = int(np.argmax(piv != pivh)) if not np.array_equal(piv, pivh) else -1
    assert first < 0, (first, piv[first], pivh[first])
    assert np.max(np.abs(LU - LUh)) / np.max(np.abs(LUh)) < 1e-11


@pytest.mark.parametrize("n,zc", [(600, 300), (200, 0), (130, 129)])
def test_zgetrf_zero_column_info(gpu, n, zc):
    import qpms_amd
    A = rand(n, 11)
    A[:, zc] = 0
    _, _, info = qpms_amd.zgetrf(A)
    _, _, infoo = oracle_ffi.zgetrf(A)
    assert info == infoo == zc + 1


# batch 1 above order 512: the lookahead path, whose strips run as whole-CU
# launches (k_gemm_cu) beside the trailing update
@pytest.mark.parametrize("prec,n,batch", [("f32", 700, 3), ("f64", 700, 3), ("f32", 1500, 1), ("f64", 1300, 1),
                                          ("f32", 2081, 1)])
def test_getrf_batched(gpu, prec, n, batch):
    import qpms_amd
    L = qpms_amd.lib()
    dt = np.complex64 if prec == "f32" else np.complex128
    As = np.stack([rand(n, 100 + z, shift=3 * np.sqrt(n)) for z in range(batch)]).astype(dt)
    dA = qpms_amd.DeviceBuffer.from_array(As)
    dpiv = qpms_amd.DeviceBuffer(4 * n * batch)
    dperm = qpms_amd.DeviceBuffer(4 * n * batch)
    dinfo = qpms_amd.DeviceBuffer.from_array(np.zeros(batch, np.int32))
    f = L.qh_getrf_batched_f32 if prec == "f32" else L.qh_getrf_batched_f64
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p]
    f(n, batch, dA.ptr, n * n, dpiv.ptr, dperm.ptr, dinfo.ptr, None)
    L.qpms_hip_sync()
    LU = dA.to_array(dt, (batch, n, n)).astype(np.complex128)
    piv = dpiv.to_array(np.int32, (batch, n))
    perm = dperm.to_array(np.int32, (batch, n))
    assert np.all(dinfo.to_array(np.int32, (batch,)) == 0)
    tol = 1e-5 if prec == "f32" else 1e-13
    for z in range(batch):
        A = As[z].astype(np.complex128)
        check_plu(A, LU[z], piv[z], tol)
        # perm is the row order of P A
        p = np.arange(n)
        for i in range(n):
            j = piv[z][i] - 1
            p[[i, j]] = p[[j, i]]
        assert np.array_equal(perm[z], p)


# many right-hand sides: 256-row diagonal blocks (k_trsm_blk) and K = 256
# updates; orders that are not multiples of 256 or 32, slabs of 32 columns
# with a partial last slab
@pytest.mark.parametrize("prec,n,nrhs,batch", [("f32", 300, 5, 2), ("f32", 777, 32, 2), ("f64", 777, 40, 2),
                                               ("f64", 2100, 32, 1), ("f32", 2100, 33, 1), ("f64", 255, 64, 3)])
def test_getrs_batched_many_rhs(gpu, prec, n, nrhs, batch):
    import qpms_amd
    L = qpms_amd.lib()
    dt = np.complex64 if prec == "f32" else np.complex128
    As = np.stack([rand(n, 300 + z, shift=3 * np.sqrt(n)) for z in range(batch)]).astype(dt)
    rng = np.random.default_rng(n + nrhs)
    Bs = (rng.normal(size=(batch, n, nrhs)) + 1j * rng.normal(size=(batch, n, nrhs))).astype(dt)
    dA = qpms_amd.DeviceBuffer.from_array(As)
    dB = qpms_amd.DeviceBuffer.from_array(Bs)
    dW = qpms_amd.DeviceBuffer(Bs.nbytes)
    dpiv = qpms_amd.DeviceBuffer(4 * n * batch)
    dperm = qpms_amd.DeviceBuffer(4 * n * batch)
    dinfo = qpms_amd.DeviceBuffer.from_array(np.zeros(batch, np.int32))
    i32, vp, lg = ctypes.c_int, ctypes.c_void_p, ctypes.c_long
    f = L.qh_getrf_batched_f32 if prec == "f32" else L.qh_getrf_batched_f64
    f.argtypes = [i32, i32, vp, lg, vp, vp, vp, vp]
    f(n, batch, dA.ptr, n * n, dpiv.ptr, dperm.ptr, dinfo.ptr, None)
    g = L.qh_getrs_batched_f32 if prec == "f32" else L.qh_getrs_batched_f64
    g.argtypes = [i32, i32, i32, vp, lg, vp, vp, lg, vp, vp]
    g(n, nrhs, batch, dA.ptr, n * n, dperm.ptr, dB.ptr, n * nrhs, dW.ptr, None)
    L.qpms_hip_sync()
    assert np.all(dinfo.to_array(np.int32, (batch,)) == 0)
    X = dB.to_array(dt, (batch, n, nrhs)).astype(np.complex128)
    tol = 2e-5 if prec == "f32" else 1e-12
    for z in range(batch):
        ref = np.linalg.solve(As[z].astype(np.complex128), Bs[z].astype(np.complex128))
        assert np.max(np.abs(X[z] - ref)) / np.max(np.abs(ref)) < tol
