"""Pins of the CPU oracle (no GPU).

Each oracle function is checked against something that does not share its
code: golden vectors from mpmath / sympy / scipy (tests/golden/golden.npz,
written by tests/golden/make_fixtures.py), the reference's own GSL-free C
sources compiled from /root/reference by oracle/build_ref.sh (Gaunt
coefficients by Xu's recursion, the spiderweb lattice generator), and
invariants that a dropped term, a wrong sign or a transposed index breaks
(two independent translation formulas, parity under d -> -d, reciprocity of
the regular operator, eta-independence of the Ewald sum).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_ffi as of

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


def test_bessel_vs_mpmath():
    args, J, Y = GOLD["bessel_args"], GOLD["bessel_j"], GOLD["bessel_y"]
    for i, z in enumerate(args):
        j = of.sph_bessel(1, 17, z)
        y = of.sph_bessel(2, 17, z)
        h = of.sph_bessel(3, 17, z)
        assert np.max(np.abs(j - J[i]) / np.abs(J[i])) < 1e-13, z
        assert np.max(np.abs(y - Y[i]) / np.abs(Y[i])) < 1e-13, z
        assert np.max(np.abs(h - (J[i] + 1j * Y[i])) / np.abs(J[i] + 1j * Y[i])) < 1e-13, z


def test_hankel_large_imaginary_part_vs_mpmath():
    """h^(1) of complex argument where forming j + i y would cancel
    (|Im z| up to 35): the direct form must hold to 1e-13."""
    for z, H in zip(GOLD["hankel_args"], GOLD["hankel_h1"]):
        h = of.sph_bessel(3, 17, z)
        assert np.max(np.abs(h - H) / np.abs(H)) < 1e-13, z
        hm = of.sph_bessel(4, 17, np.conj(z))   # h^(2)(conj z) = conj h^(1)(z)
        assert np.max(np.abs(hm - np.conj(H)) / np.abs(H)) < 1e-13, z


def test_legendre_vs_sympy():
    for x, P in zip(GOLD["legendre_x"], GOLD["legendre_P"]):
        got = of.legendre_array(9, x, -1)
        scale = np.maximum(np.abs(P), 1e-300)
        ok = (np.abs(got - P) <= 1e-12 * scale) | (np.abs(P) < 1e-14)
        assert np.all(ok), x
        # csphase = +1 drops (-1)^m
        got1 = of.legendre_array(9, x, 1)
        m = np.concatenate([np.arange(l + 1) for l in range(10)])
        assert np.allclose(got1, got * (-1.0) ** m, rtol=1e-15, atol=0)


def test_wigner3j_vs_sympy():
    o = of.oracle()
    for j1, j2, j3, m1, m2, m3, v in GOLD["wigner3j"]:
        got = o.or_wigner3j(int(j1), int(j2), int(j3), int(m1), int(m2), int(m3))
        assert abs(got - v) < 1e-14 * max(1, abs(v))


def test_gaunt_vs_sympy_exact():
    o = of.oracle()
    for m, n, mu, nu, q, v in GOLD["gaunt"]:
        got = o.or_gaunt(int(m), int(n), int(mu), int(nu), int(q))
        assert abs(got - v) <= 1e-12 * max(1, abs(v)), (m, n, mu, nu, q)


def test_gaunt_vs_reference_xu_recursion():
    R = of.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    o = of.oracle()
    for n in range(1, 5):
        for nu in range(1, 5):
            for m in range(-n, n + 1):
                for mu in range(-nu, nu + 1):
                    qmax = min(n, nu, (n + nu - abs(m + mu)) // 2)
                    buf = np.zeros(qmax + 1)
                    err = ctypes.c_int()
                    R.gaunt_xu(m, n, mu, nu, qmax, buf.ctypes.data, ctypes.byref(err))
                    assert err.value == 0
                    mine = np.array([o.or_gaunt(m, n, mu, nu, q) for q in range(qmax + 1)])
                    # the translation multipliers use a_q / a_0
                    assert np.max(np.abs(buf / buf[0] - mine / mine[0])) < 1e-12, (m, n, mu, nu)


def test_xyweb_vs_reference_generator():
    R = of.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")

    class PGen(ctypes.Structure):
        _fields_ = [("c", ctypes.c_void_p), ("s", ctypes.c_void_p)]

    class cart2(ctypes.Structure):
        _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double)]

    class Ret(ctypes.Structure):
        _fields_ = [("flags", ctypes.c_int), ("x", ctypes.c_double), ("y", ctypes.c_double)]

    R.PGen_xyWeb_new.restype = PGen
    R.PGen_xyWeb_new.argtypes = [cart2, cart2, ctypes.c_double, cart2, ctypes.c_double, ctypes.c_bool,
                                 ctypes.c_double, ctypes.c_bool]
    R.PGen_xyWeb_next_cart2.restype = Ret
    R.PGen_xyWeb_next_cart2.argtypes = [ctypes.POINTER(PGen)]
    o = of.oracle()
    o.or_xyweb_points.restype = ctypes.c_int
    o.or_xyweb_points.argtypes = [ctypes.c_double] * 5 + [ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                                         ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    a = 375e-9
    cases = [((a, 0), (a / 2, a * np.sqrt(3) / 2), 0, False, 12.5 * a * np.sqrt(3) / 2, False),
             ((1, 0), (0, 1), 0, True, 20, False), ((1, 0), (0, 1), 0, True, 20, True),
             ((2, 0.3), (0.5, 1.7), 1.0, False, 9.0, False),
             ((2 * np.pi / a, -2 * np.pi / (a * np.sqrt(3))), (0, 4 * np.pi / (a * np.sqrt(3))), 0, True,
              30 * 2 * np.pi / a, False)]
    for b1, b2, minR, incmin, maxR, incmax in cases:
        g = R.PGen_xyWeb_new(cart2(*b1), cart2(*b2), 1e-13, cart2(0, 0), minR, incmin, maxR, incmax)
        ref = []
        while True:
            r = R.PGen_xyWeb_next_cart2(ctypes.byref(g))
            if not (r.flags & 2):
                break
            ref.append((r.x, r.y))
        ref = np.array(ref)
        buf = np.empty((200000, 2))
        n = o.or_xyweb_points(b1[0], b1[1], b2[0], b2[1], 1e-13, minR, int(incmin), maxR, int(incmax),
                              buf.ctypes.data, 200000)
        assert n == len(ref), (b1, b2, n, len(ref))
        assert np.array_equal(buf[:n], ref)


def test_sr_integral_vs_mpmath():
    o = of.oracle()
    o.or_sr_integral.restype = ctypes.c_double
    o.or_sr_integral.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    for R, k, n, eta, v in GOLD["sr_integrals"]:
        got = o.or_sr_integral(R, k, int(n), eta)
        assert abs(got - v) <= 1e-13 * abs(v), (R, k, n)


def test_gamma_inc_vs_mpmath_with_branches():
    o = of.oracle()
    o.or_gamma_inc.restype = of.c_cplx
    o.or_gamma_inc.argtypes = [ctypes.c_double, of.c_cplx, ctypes.c_int]
    for a, zr, zi, sgn, m, vr, vi in GOLD["gamma_inc"]:
        z = complex(zr, zi)
        if zi == 0 and zr < 0:
            z = complex(zr, -0.0 if sgn < 0 else 0.0)
        g = o.or_gamma_inc(a, of.c_cplx(z.real, z.imag), int(m))
        got, ref = complex(g.re, g.im), complex(vr, vi)
        assert abs(got - ref) <= 1e-12 * abs(ref), (a, z, m, got, ref)


# ---- translation coefficients: invariants ---------------------------------
def y2mn(y):
    n = int(np.sqrt(y + 1))
    return y - n * (n + 1) + 1, n


@pytest.mark.parametrize("lMax", [2, 3, 4])
def test_single_formula_matches_array(lMax):
    """Two independent formulas (translations.c:135-286 vs :313-698), as the
    reference's tests/single_translations_vs_calc.c checks."""
    o = of.oracle()
    c = of.OracleCalc(lMax)
    rng = np.random.default_rng(666)
    pts = [(2.0, 0.5, 0.5), (7.3, 0.0, 0.0), (0.8, np.pi, 0.3)] + \
          [(rng.uniform(0.5, 30), rng.uniform(0, np.pi), rng.uniform(-3, 3)) for _ in range(5)]
    for J in (1, 3):
        for kr, th, ph in pts:
            A, B = c.get_AB_arrays(kr, th, ph, J=J)
            atol = 1e-13 * max(np.max(np.abs(A)), np.max(np.abs(B)))  # analytic zeros carry rounding noise
            for y in range(c.nelem):
                m, n = y2mn(y)
                for yu in range(c.nelem):
                    mu, nu = y2mn(yu)
                    a = o.or_trans_single_A(c.norm, m, n, mu, nu, of.c_cplx.of(kr), th, ph, J)
                    b = o.or_trans_single_B(c.norm, m, n, mu, nu, of.c_cplx.of(kr), th, ph, J)
                    a, b = complex(a.re, a.im), complex(b.re, b.im)
                    assert abs(a - A[y, yu]) <= atol + 1e-8 * 0.5 * (abs(a) + abs(A[y, yu]))
                    assert abs(b - B[y, yu]) <= atol + 1e-8 * 0.5 * (abs(b) + abs(B[y, yu]))


@pytest.mark.parametrize("lMax", [2, 3])
def test_parity_under_inversion(lMax):
    """A(-d) = (-1)^{n+nu} A(d), B(-d) = (-1)^{n+nu+1} B(d) (inversion symmetry
    of the VSWFs: M ~ (-1)^n, N ~ (-1)^{n+1})."""
    c = of.OracleCalc(lMax)
    for kr, th, ph in [(3.1, 0.7, 0.4), (12.0, 2.2, -1.3)]:
        A, B = c.get_AB_arrays(kr, th, ph, J=3)
        Am, Bm = c.get_AB_arrays(kr, np.pi - th, ph + np.pi, J=3)
        for y in range(c.nelem):
            _, n = y2mn(y)
            for yu in range(c.nelem):
                _, nu = y2mn(yu)
                s = (-1) ** (n + nu)
                assert abs(Am[y, yu] - s * A[y, yu]) < 1e-12 * max(1, abs(A[y, yu]))
                assert abs(Bm[y, yu] + s * B[y, yu]) < 1e-12 * max(1, abs(B[y, yu]))


def test_regular_operator_reciprocity():
    """The regular operator of d followed by that of -d is the identity on the
    waves it keeps (truncation error decays fast for small kd)."""
    lMax = 4
    c = of.OracleCalc(lMax)
    kr, th, ph = 0.05, 0.9, 0.3
    A, B = c.get_AB_arrays(kr, th, ph, J=1)
    Am, Bm = c.get_AB_arrays(kr, np.pi - th, ph + np.pi, J=1)
    T = np.block([[A, B], [B, A]])
    Tm = np.block([[Am, Bm], [Bm, Am]])
    P = T @ Tm
    ne = c.nelem
    # restrict to n = 1, 2 (higher n lose accuracy from the truncated inner sum)
    keep = [i for i in range(ne) if y2mn(i)[1] <= 2]
    keep = keep + [ne + i for i in keep]
    assert np.max(np.abs(P[np.ix_(keep, keep)] - np.eye(len(keep)))) < 1e-6


def test_bessel_large_real_argument_vs_mpmath():
    """Real arguments far above the order take the Hankel finite sums
    (DLMF 10.49.6) instead of the downward recurrence."""
    import mpmath
    mpmath.mp.dps = 40
    for x in (1100.5, 5000.25, 1.0e6 + 0.125):
        j, y, h, hm = (of.sph_bessel(J, 9, x) for J in (1, 2, 3, 4))
        for n in range(10):
            f = mpmath.sqrt(mpmath.pi / (2 * mpmath.mpf(x)))
            jr = float(f * mpmath.besselj(n + 0.5, x))
            yr = float(f * mpmath.bessely(n + 0.5, x))
            s = np.hypot(jr, yr)
            assert abs(j[n] - jr) < 1e-13 * s and abs(y[n] - yr) < 1e-13 * s, (x, n)
            assert abs(h[n] - (jr + 1j * yr)) < 1e-13 * s and abs(hm[n] - (jr - 1j * yr)) < 1e-13 * s, (x, n)
