"""GPU parity of the 2D Ewald sums (ewald.c) and the e32 coefficients
(translations.c:926-1094) against the oracle, on the hexagonal lattice of the
2D-periodic configuration: pitch 375 nm, shifts {0, +-(a/sqrt3, 0)}, lMax = 3
(Ewald order 7), eta = 2.5/a, real and complex k, Bloch vectors inside the
zone and within 1e-4 of the light cone. Tolerance: 1e-10 relative to the
largest coefficient.
"""
import ctypes

import numpy as np
import pytest

import ewald_common as ec
import oracle_ffi as of

pytestmark = pytest.mark.gpu
LE = 7
TOL = 1e-10


def relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def cases():
    k0 = ec.wavenumber(2.0)
    s = ec.A_LAT / np.sqrt(3)
    out = []
    for k in (k0, k0 * (1 + 1e-3j), ec.wavenumber(1.5), ec.wavenumber(2.3)):
        for beta in ((0.0, 0.0), (0.3 * 2 * np.pi / ec.A_LAT, 0.1 * 2 * np.pi / ec.A_LAT),
                     (k0.real * (1 + 1e-4), 0.0), (k0.real * (1 - 1e-4) * 0.6, k0.real * (1 - 1e-4) * 0.8)):
            for sh in ((0.0, 0.0), (s, 0.0), (-s, 0.0)):
                out.append((k, beta, sh))
    return out


@pytest.fixture(scope="module")
def ew(gpu):
    L = ec.lib_ewald()
    e3c = L.qpms_ewald3_constants_init(LE, -1)
    yield L, e3c
    L.qpms_ewald3_constants_free(e3c)


def test_pgen_matches_oracle_points(ew):
    import qpms_amd
    L, _ = ew
    L.PGen_next_cart2.restype = type("R", (ctypes.Structure,), {"_fields_": [("flags", ctypes.c_int),
                                                                              ("x", ctypes.c_double),
                                                                              ("y", ctypes.c_double)]})
    L.PGen_next_cart2.argtypes = [ctypes.POINTER(ec.PGen)]
    for (b1, b2, maxR) in [(ec.B1, ec.B2, 12.5 * ec.H_R), (ec.RB1, ec.RB2, 30 * 2 * np.pi / ec.A_LAT)]:
        g = L.PGen_xyWeb_new(qpms_amd.cart2_t(*b1), qpms_amd.cart2_t(*b2), 1e-13, qpms_amd.cart2_t(0, 0), 0, True,
                             maxR, False)
        pts = []
        while True:
            r = L.PGen_next_cart2(ctypes.byref(g))
            if not (r.flags & 2):
                break
            pts.append((r.x, r.y))
        ref = ec.xyweb(b1, b2, 0, True, maxR, False)
        assert np.array_equal(np.array(pts), ref)


@pytest.mark.parametrize("case", range(48))
def test_sigma_short_long(ew, case):
    import qpms_amd
    L, e3c = ew
    k, beta, sh = cases()[case]
    s0 = (sh == (0.0, 0.0))
    cart2, cart3 = qpms_amd.cart2_t, qpms_amd.cart3_t
    nsc = (LE + 1) ** 2
    # short range: R points with the origin excluded iff s = 0 (translations.c:981-987)
    g = L.PGen_xyWeb_new(cart2(*ec.B1), cart2(*ec.B2), 1e-13, cart2(0, 0), 0, not s0, 12.5 * ec.H_R, False)
    out = np.empty(nsc, np.complex128)
    err = np.empty(nsc)
    L.ewald3_sigma_short(out.ctypes.data, err.ctypes.data, e3c, ec.ETA, qpms_amd.c_cplx.of(k),
                         ec.LAT_2D_IN_3D_XYONLY, ctypes.byref(g), False, cart3(beta[0], beta[1], 0),
                         cart3(sh[0], sh[1], 0))
    R = ec.xyweb(ec.B1, ec.B2, 0, not s0, 12.5 * ec.H_R, False)
    ref = ec.oracle_sigma_short(LE, k, beta, sh, R)
    assert relerr(out, ref) < TOL
    # long range
    maxK = 30 * 2 * np.pi / ec.A_LAT
    g = L.PGen_xyWeb_new(cart2(*ec.RB1), cart2(*ec.RB2), 1e-13, cart2(0, 0), 0, True, maxK, False)
    L.ewald3_sigma_long(out.ctypes.data, None, e3c, ec.ETA, qpms_amd.c_cplx.of(k), ec.AREA,
                        ec.LAT_2D_IN_3D_XYONLY, ctypes.byref(g), False, cart3(beta[0], beta[1], 0),
                        cart3(sh[0], sh[1], 0))
    K = ec.xyweb(ec.RB1, ec.RB2, 0, True, maxK, False)
    refl = ec.oracle_sigma_long(LE, k, beta, sh, K)
    assert relerr(out, refl) < TOL


def test_sigma0(ew):
    import qpms_amd
    L, e3c = ew
    for k in (ec.wavenumber(2.0), ec.wavenumber(2.0) * (1 + 1e-3j), ec.wavenumber(1.6) * (1 - 2e-3j)):
        out = np.empty(1, np.complex128)
        L.ewald3_sigma0(out.ctypes.data, None, e3c, ec.ETA, qpms_amd.c_cplx.of(k))
        ref = ec.oracle_sigma0(k)
        assert abs(out[0] - ref) <= TOL * abs(ref)


def test_batch_many_triples(ew):
    import qpms_amd
    L, e3c = ew
    rng = np.random.default_rng(5)
    R = ec.xyweb(ec.B1, ec.B2, 0, True, 12.5 * ec.H_R, False)
    R = R[1:]  # origin excluded: the batch triples here all have s != 0 or want no origin
    K = ec.K_points()
    s = ec.A_LAT / np.sqrt(3)
    n = 40
    tr = np.zeros((n, 8))
    for i in range(n):
        k = ec.wavenumber(rng.uniform(1.5, 2.3))
        b = rng.uniform(-0.5, 0.5, 2) * 2 * np.pi / ec.A_LAT
        sh = [(0, 0), (s, 0), (-s, 0)][i % 3]
        tr[i] = [b[0], b[1], 0, k, 0, sh[0], sh[1], 0]
    R3 = np.concatenate([R, np.zeros((len(R), 1))], axis=1)
    dtr, dR, dK = (qpms_amd.DeviceBuffer.from_array(x) for x in (tr, R3, K))
    nsc = (LE + 1) ** 2
    dss, dsl = qpms_amd.DeviceBuffer(n * nsc * 16), qpms_amd.DeviceBuffer(n * nsc * 16)
    L.qpms_hip_ewald3_sigma_batch_dev(e3c, ec.ETA, ec.AREA, n, dtr.ptr, len(R3), dR.ptr, len(K), dK.ptr, dss.ptr,
                                      dsl.ptr, None)
    qpms_amd.lib().qpms_hip_sync()
    ss = dss.to_array(np.complex128, (n, nsc))
    sl = dsl.to_array(np.complex128, (n, nsc))
    for i in range(0, n, 5):
        k, b, sh = tr[i, 3], tr[i, :2], tr[i, 5:7]
        assert relerr(ss[i], ec.oracle_sigma_short(LE, k, b, sh, R)) < TOL, i
        assert relerr(sl[i], ec.oracle_sigma_long(LE, k, b, sh, K)) < TOL, i


def test_eta_independence(ew):
    """sigma_short + sigma_long (+ sigma0) does not depend on eta (the
    reference's tests/test_scalar_ewald32.c invariant), on the GPU."""
    import qpms_amd
    L, e3c = ew
    cart2, cart3 = qpms_amd.cart2_t, qpms_amd.cart3_t
    k = ec.wavenumber(2.0) * (1 + 0.05j)   # mild absorption
    nsc = (LE + 1) ** 2
    for sh in ((0.0, 0.0), (ec.A_LAT / np.sqrt(3), 0.0)):
        s0 = sh == (0.0, 0.0)
        tot = []
        for eta in (1.5 / ec.A_LAT, 2.5 / ec.A_LAT, 3.5 / ec.A_LAT):
            a = np.empty(nsc, np.complex128)
            b = np.empty(nsc, np.complex128)
            g = L.PGen_xyWeb_new(cart2(*ec.B1), cart2(*ec.B2), 1e-13, cart2(0, 0), 0, not s0, 40 * ec.H_R, False)
            L.ewald3_sigma_short(a.ctypes.data, None, e3c, eta, qpms_amd.c_cplx.of(k), ec.LAT_2D_IN_3D_XYONLY,
                                 ctypes.byref(g), False, cart3(0.4e6, 0.2e6, 0), cart3(sh[0], sh[1], 0))
            g = L.PGen_xyWeb_new(cart2(*ec.RB1), cart2(*ec.RB2), 1e-13, cart2(0, 0), 0, True,
                                 60 * 2 * np.pi / ec.A_LAT, False)
            L.ewald3_sigma_long(b.ctypes.data, None, e3c, eta, qpms_amd.c_cplx.of(k), ec.AREA,
                                ec.LAT_2D_IN_3D_XYONLY, ctypes.byref(g), False, cart3(0.4e6, 0.2e6, 0),
                                cart3(sh[0], sh[1], 0))
            t = a + b
            if s0:
                z = np.empty(1, np.complex128)
                L.ewald3_sigma0(z.ctypes.data, None, e3c, eta, qpms_amd.c_cplx.of(k))
                t[0] += z[0]
            tot.append(t)
        for t in tot[1:]:
            assert np.max(np.abs(t - tot[0])) <= 1e-10 + 1e-8 * np.max(np.abs(tot[0]))


def test_e32_ab_arrays(ew, gpu):
    import qpms_amd
    L, _ = ew
    lMax = 3
    calc = qpms_amd.TransCalculator(lMax)
    oc = of.OracleCalc(lMax)
    ne = calc.nelem
    k = ec.wavenumber(2.0)
    maxR, maxK = 12.5 * ec.H_R, 30 * 2 * np.pi / ec.A_LAT
    s = ec.A_LAT / np.sqrt(3)
    o = of.oracle()
    o.or_e32_combine.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    for beta in ((0.0, 0.0), (0.2e7, -0.1e7)):
        for sh in ((0.0, 0.0), (s, 0.0)):
            s0 = sh == (0.0, 0.0)
            A = np.empty((ne, ne), np.complex128)
            B = np.empty((ne, ne), np.complex128)
            L.qpms_trans_calculator_get_AB_arrays_e32_e(calc.c, A.ctypes.data, None, B.ctypes.data, None, ne, 1,
                                                        ec.ETA, qpms_amd.c_cplx.of(k), qpms_amd.cart2_t(*ec.B1),
                                                        qpms_amd.cart2_t(*ec.B2), qpms_amd.cart2_t(*beta),
                                                        qpms_amd.cart3_t(sh[0], sh[1], 0), maxR, maxK, 7)
            R = ec.xyweb(ec.B1, ec.B2, 0, not s0, maxR, False)
            K = ec.xyweb(ec.RB1, ec.RB2, 0, True, maxK, False)
            sig = ec.oracle_sigma_short(LE, k, beta, sh, R) + ec.oracle_sigma_long(LE, k, beta, sh, K)
            if s0:
                sig[0] += ec.oracle_sigma0(k)
            Ao = np.empty((ne, ne), np.complex128)
            Bo = np.empty((ne, ne), np.complex128)
            o.or_e32_combine(oc.c, sig.ctypes.data, 0, Ao.ctypes.data, Bo.ctypes.data)
            assert relerr(A, Ao) < TOL
            assert relerr(B, Bo) < TOL
