"""GPU: the allocation-free LU entry point qpms_hip_zgetrf_dev_noalloc
against qpms_hip_zgetrf_dev through the C ABI, on two calls in a row (the
second smaller, so stale scratch from the first would show), plus the
rejection of an order above the scratch array's limit.

Bar: identical pivots; factors within 1e-12 of the largest entry (the two
paths differ only in the rounding of the trailing GEMM: hand-written MFMA
kernel against the library GEMM); info equal."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind(L):
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.qpms_hip_zgetrf_dev_noalloc.restype = ctypes.c_int
    L.qpms_hip_zgetrf_dev_noalloc.argtypes = [sz, vp, vp, vp]
    L.qpms_hip_zgetrf_noalloc_max_n.restype = sz
    L.qpms_hip_zgetrf_dev.restype = ctypes.c_int
    return L


def _lu(L, fn, A):
    import qpms_amd
    n = A.shape[0]
    dA = qpms_amd.DeviceBuffer.from_array(A)
    dp = qpms_amd.DeviceBuffer(4 * n)
    info = fn(n, dA.ptr, dp.ptr, None)
    LU, piv = dA.to_array(np.complex128, (n, n)), dp.to_array(np.int32, (n,))
    dA.free()
    dp.free()
    return LU, piv, info


def _matrix(n, seed, singular_at=None):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    if singular_at is not None:   # an exactly zero column: zgetrf's info is its index
        A[:, singular_at] = 0
    return A


@pytest.mark.parametrize("sizes", [(1500, 700), (2300, 257)])
def test_noalloc_matches_allocating_entry(gpu, sizes):
    L = _bind(gpu)
    for i, n in enumerate(sizes):
        print(f"no-alloc LU: call {i}, order {n}", flush=True)
        A = _matrix(n, 10 * n + i)
        LU0, p0, i0 = _lu(L, L.qpms_hip_zgetrf_dev, A)
        LU1, p1, i1 = _lu(L, L.qpms_hip_zgetrf_dev_noalloc, A)
        assert i0 == i1 == 0
        assert np.array_equal(p0, p1)
        assert np.max(np.abs(LU1 - LU0)) <= 1e-12 * np.max(np.abs(LU0))
        # and the factors are a factorisation of A: P A = L U
        lo = np.tril(LU1, -1) + np.eye(n)
        up = np.triu(LU1)
        PA = A.copy()
        for r in range(n):
            q = p1[r] - 1
            if q != r:
                PA[[r, q]] = PA[[q, r]]
        assert np.max(np.abs(lo @ up - PA)) <= 1e-11 * np.max(np.abs(A)) * n ** 0.5


def test_noalloc_info_and_reuse(gpu):
    """A singular matrix sets info; the next call (smaller) starts from a
    cleared info word."""
    L = _bind(gpu)
    A = _matrix(900, 5, singular_at=17)
    _, _, i0 = _lu(L, L.qpms_hip_zgetrf_dev, A)
    _, _, i1 = _lu(L, L.qpms_hip_zgetrf_dev_noalloc, A)
    assert i0 == i1 == 18
    B = _matrix(300, 6)
    LU0, p0, j0 = _lu(L, L.qpms_hip_zgetrf_dev, B)
    LU1, p1, j1 = _lu(L, L.qpms_hip_zgetrf_dev_noalloc, B)
    assert j0 == j1 == 0 and np.array_equal(p0, p1)
    assert np.max(np.abs(LU1 - LU0)) <= 1e-12 * np.max(np.abs(LU0))


def test_noalloc_rejects_orders_above_scratch(gpu):
    L = _bind(gpu)
    nmax = L.qpms_hip_zgetrf_noalloc_max_n()
    assert 100000 <= nmax < 200000
    # rejected before any device access: no buffers needed
    assert L.qpms_hip_zgetrf_dev_noalloc(nmax + 256, None, None, None) == -1
    assert L.qpms_hip_zgetrf_dev_noalloc(1 << 23, None, None, None) == -1
