"""GPU parity: translation coefficients, W / mode-problem builds, Mie, LU.

Every test calls the HIP kernels through libqpms_hip.so's C ABI and compares
with the CPU oracle on identical seeded inputs. Tolerance for FP64 builds:
max |gpu - oracle| <= 1e-10 * max |oracle| (per array / per block).
"""
import numpy as np
import pytest

import oracle_ffi

pytestmark = pytest.mark.gpu

EV = 1.602176634e-19 / 1.054571817e-34  # rad/s per eV (CODATA 2018)
C0 = 2.99792458e8
TOL = 1e-10


def relerr(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


@pytest.mark.parametrize("lMax", [1, 2, 3, 4])
@pytest.mark.parametrize("J", [1, 2, 3, 4])
def test_ab_arrays_single(gpu, lMax, J):
    import qpms_amd
    rng = np.random.default_rng(100 + 10 * lMax + J)
    c = qpms_amd.TransCalculator(lMax)
    o = oracle_ffi.OracleCalc(lMax)
    pts = [(2.0, 0.5, 0.5), (0.3, 1.2, -2.0), (15.0, 0.0, 0.0), (7.0, np.pi, 1.0), (40.0, np.pi / 2, 3.0),
           (1.0 + 0.2j, 0.7, 0.1)]
    for _ in range(4):
        pts.append((rng.uniform(0.5, 60), rng.uniform(0, np.pi), rng.uniform(-np.pi, np.pi)))
    for kr, th, ph in pts:
        A, B = c.get_AB_arrays(kr, th, ph, J=J)
        Ao, Bo = o.get_AB_arrays(kr, th, ph, J=J)
        assert relerr(A, Ao) < TOL, (kr, th, ph)
        assert relerr(B, Bo) < TOL, (kr, th, ph)


@pytest.mark.parametrize("J", [3, 4])
def test_ab_arrays_hankel_large_imaginary_argument(gpu, J):
    """Strongly absorbing media: kr with |Im| up to 12, where j +- i y cancels;
    the GPU forms h directly and must agree with the oracle's direct form."""
    import qpms_amd
    c = qpms_amd.TransCalculator(4)
    o = oracle_ffi.OracleCalc(4)
    for kr in (20 + 8j, 5 + 12j, 0.5 + 3j, 30 + 1j, 3 + 0.5j):
        for th, ph in ((0.4, 1.0), (np.pi / 2, -2.0)):
            A, B = c.get_AB_arrays(kr, th, ph, J=J)
            Ao, Bo = o.get_AB_arrays(kr, th, ph, J=J)
            assert relerr(A, Ao) < TOL, (kr, th, ph)
            assert relerr(B, Bo) < TOL, (kr, th, ph)


def test_ab_arrays_r0_nan(gpu):
    import qpms_amd
    c = qpms_amd.TransCalculator(2)
    A, B = c.get_AB_arrays(0.0, 0.0, 0.0, J=3)
    assert np.all(np.isnan(A)) and np.all(np.isnan(B))
    # regular kind at r = 0: finite, A is the identity-like regular operator
    A, B = c.get_AB_arrays(0.0, 0.0, 0.0, J=1)
    assert np.all(np.isfinite(A)) and np.all(np.isfinite(B))
    o = oracle_ffi.OracleCalc(2)
    Ao, Bo = o.get_AB_arrays(0.0, 0.0, 0.0, J=1)
    assert relerr(A, Ao) < TOL and np.max(np.abs(B - Bo)) < 1e-12


@pytest.mark.parametrize("lMax", [2, 3, 4])
def test_ab_arrays_batch(gpu, lMax):
    import qpms_amd
    rng = np.random.default_rng(7 + lMax)
    n = 300
    kr = rng.uniform(0.2, 80.0, n) * np.where(rng.random(n) < 0.2, 1 + 0.05j, 1.0)
    th = rng.uniform(0, np.pi, n)
    ph = rng.uniform(-np.pi, np.pi, n)
    c = qpms_amd.TransCalculator(lMax)
    o = oracle_ffi.OracleCalc(lMax)
    A, B = c.get_AB_arrays_batch(kr, th, ph, J=3)
    for i in range(0, n, 7):
        Ao, Bo = o.get_AB_arrays(kr[i], th[i], ph[i], J=3)
        assert relerr(A[i], Ao) < TOL and relerr(B[i], Bo) < TOL, i


def random_cluster(n, seed, box=(4e-6, 4e-6, 0.5e-6), dmin=120e-9):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, 1, 3) * np.array(box)
        if all(np.linalg.norm(p - q) >= dmin for q in pts):
            pts.append(p)
    return np.array(pts)


@pytest.mark.parametrize("lMax,npart", [(1, 5), (2, 16), (3, 37), (3, 70), (4, 65)])
def test_w_build_full(gpu, lMax, npart):
    import qpms_amd
    pos = random_cluster(npart, 1000 + npart)
    k = 1.52 * 2.0 * EV / C0
    c = qpms_amd.TransCalculator(lMax)
    W = c.build_W(pos, k)
    Wo = oracle_ffi.OracleCalc(lMax).build_W_rows(pos, k, 0, npart)
    nb = 2 * lMax * (lMax + 2)
    assert np.all(W[:nb, :nb] == 0)
    for r in range(npart):
        for cc in range(npart):
            blk = W[r * nb:(r + 1) * nb, cc * nb:(cc + 1) * nb]
            blko = Wo[r * nb:(r + 1) * nb, cc * nb:(cc + 1) * nb]
            if r == cc:
                assert np.all(blk == 0)
            else:
                assert relerr(blk, blko) < TOL, (r, cc)


def test_w_build_complex_k(gpu):
    import qpms_amd
    pos = random_cluster(20, 5)
    k = 1.52 * (2.0 + 0.003j) * EV / C0
    W = qpms_amd.TransCalculator(2).build_W(pos, k)
    Wo = oracle_ffi.OracleCalc(2).build_W_rows(pos, k, 0, 20)
    assert relerr(W, Wo) < TOL


def test_w_rows_api(gpu):
    import ctypes
    import qpms_amd
    lMax, npart = 3, 90
    pos = random_cluster(npart, 77)
    k = 1.52 * 1.8 * EV / C0
    c = qpms_amd.TransCalculator(lMax)
    nb = 2 * lMax * (lMax + 2)
    L = nb * npart
    r0, nr = 17, 11
    ld = L + 8   # padded leading dimension
    dpos = qpms_amd.DeviceBuffer.from_array(pos)
    dW = qpms_amd.DeviceBuffer(nr * nb * ld * 16)
    qpms_amd.lib().qpms_hip_build_translation_matrix_rows_dev(c.c, npart, dpos.ptr, qpms_amd.c_cplx.of(k), 3, r0, nr,
                                                             dW.ptr, ld, None)
    qpms_amd.lib().qpms_hip_sync()
    W = dW.to_array(np.complex128, (nr * nb, ld))[:, :L]
    Wo = oracle_ffi.OracleCalc(lMax).build_W_rows(pos, k, r0, nr)
    assert relerr(W, Wo) < TOL
    assert ctypes.sizeof(ctypes.c_double) == 8


def test_modeproblem(gpu):
    import qpms_amd
    lMax, npart = 2, 40
    pos = random_cluster(npart, 9)
    k = 1.52 * 2.05 * EV / C0
    nb = 2 * lMax * (lMax + 2)
    rng = np.random.default_rng(3)
    tdiag = (rng.normal(size=(2, nb)) + 1j * rng.normal(size=(2, nb))) * 0.1
    ttype = rng.integers(0, 2, npart)
    M = qpms_amd.TransCalculator(lMax).build_M(pos, ttype, tdiag, k)
    Mo = oracle_ffi.OracleCalc(lMax).build_M(pos, ttype, tdiag, k)
    assert relerr(M, Mo) < TOL


def test_mie(gpu):
    import ctypes
    import qpms_amd
    L = qpms_amd.lib()
    omega = 2.0 * EV
    for (eps_i, eps_e, a, lMax) in [(-12 + 1j, 2.31, 50e-9, 2), (-15 + 1.5j, 1.52 ** 2, 80e-9, 4),
                                    (4.0 + 0.01j, 1.0, 200e-9, 4)]:
        ki = np.sqrt(eps_i + 0j) * omega / C0
        ke = np.sqrt(eps_e + 0j) * omega / C0
        RVo, RHo = oracle_ffi.mie(lMax, a, ki, ke)
        L.qpms_vswf_set_spec_from_lMax.restype = ctypes.c_void_p
        L.qpms_vswf_set_spec_from_lMax.argtypes = [ctypes.c_int, ctypes.c_int]
        L.qpms_mie_coefficients_reflection.restype = ctypes.c_void_p
        L.qpms_mie_coefficients_reflection.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                                       qpms_amd.c_cplx, qpms_amd.c_cplx, qpms_amd.c_cplx,
                                                       qpms_amd.c_cplx, ctypes.c_int, ctypes.c_int]
        bs = L.qpms_vswf_set_spec_from_lMax(lMax, qpms_amd.NORM_DEFAULT)
        nb = 2 * lMax * (lMax + 2)
        out = np.empty(nb, np.complex128)
        L.qpms_mie_coefficients_reflection(out.ctypes.data, bs, a, qpms_amd.c_cplx.of(ki), qpms_amd.c_cplx.of(ke),
                                           qpms_amd.c_cplx.of(1), qpms_amd.c_cplx.of(1), 1, 3)
        ne = lMax * (lMax + 2)
        ls = np.concatenate([[l] * (2 * l + 1) for l in range(1, lMax + 1)])
        assert relerr(out[:ne], RVo[ls]) < TOL
        assert relerr(out[ne:], RHo[ls]) < TOL


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 100, 256, 600])
def test_lu_solve(gpu, n):
    import qpms_amd
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    LU, piv, info = qpms_amd.zgetrf(A)
    assert info == 0
    # P A = L U reconstructs A (pivot order may differ from the oracle at near-ties)
    Lm = np.tril(LU, -1) + np.eye(n)
    Um = np.triu(LU)
    PA = A.copy()
    for i in range(n):
        p = piv[i] - 1
        PA[[i, p]] = PA[[p, i]]
    assert relerr(Lm @ Um, PA) < 1e-12
    x = qpms_amd.zgetrs(LU, piv, b)
    xo = np.linalg.solve(A, b)
    assert np.max(np.abs(x - xo)) / np.max(np.abs(xo)) < 1e-9


def test_lu_singular_info(gpu):
    import qpms_amd
    A = np.zeros((40, 40), np.complex128)
    A[:, :10] = np.random.default_rng(1).normal(size=(40, 10))
    LU, piv, info = qpms_amd.zgetrf(A)
    _, _, infoo = oracle_ffi.zgetrf(A)
    assert info == infoo == 11


@pytest.mark.parametrize("n,nrhs", [(64, 1), (300, 4), (1000, 8)])
def test_mixed_solve(gpu, n, nrhs):
    import qpms_amd
    rng = np.random.default_rng(n + nrhs)
    A = np.eye(n) * 4 + (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    B = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
    X, it = qpms_amd.solve_mixed(A, B, tol=1e-15, maxit=20)
    Xo = np.linalg.solve(A, B)
    assert it != 0
    assert np.max(np.abs(X - Xo)) / np.max(np.abs(Xo)) < 1e-9
