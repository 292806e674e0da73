"""Beyn integration contours (beyn.c:62-248) from the library's C ABI: node
positions and weights against the oracle's restatement, and the contour
integrals the trapezoid rule must reproduce (winding number of an interior
point, zero integral of analytic functions, enclosed area). Host-only code:
runs without a GPU."""
import ctypes

import numpy as np
import pytest

import oracle_beyn as ob


def lib():
    import qpms_amd
    L = qpms_amd.lib()
    c = qpms_amd.c_cplx
    L.beyn_contour_ellipse.restype = ctypes.c_void_p
    L.beyn_contour_ellipse.argtypes = [c, ctypes.c_double, ctypes.c_double, ctypes.c_size_t]
    L.beyn_contour_halfellipse.restype = ctypes.c_void_p
    L.beyn_contour_halfellipse.argtypes = [c, ctypes.c_double, ctypes.c_double, ctypes.c_size_t, ctypes.c_int]
    L.beyn_contour_kidney.restype = ctypes.c_void_p
    L.beyn_contour_kidney.argtypes = [c, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_size_t,
                                      ctypes.c_int]
    return L


def read_contour(p):
    """beyn_contour_t: size_t n; complex centre; inside_test pointer; z_dz[n][2] (beyn.h:30-43)."""
    n = ctypes.c_size_t.from_address(p).value
    raw = (ctypes.c_double * (4 * n)).from_address(p + 32)
    a = np.frombuffer(raw, np.float64).copy().view(np.complex128).reshape(n, 2)
    fn = ctypes.c_void_p.from_address(p + 24).value
    return a[:, 0], a[:, 1], fn


def free(p):
    ctypes.CDLL(None).free(ctypes.c_void_p(p))


def check_integrals(z, dz, inside_pt, area, tol, tol_w=None):
    """Closed contour: sum dz = 0, winding number 1 about an interior point,
    enclosed area 1/2 Im sum conj(z) dz; tol is the quadrature's relative
    accuracy (exponential for the smooth ellipse, O(h^2) with corners)."""
    perim = np.sum(np.abs(dz))
    assert abs(np.sum(dz)) < tol * perim
    w = np.sum(dz / (z - inside_pt)) / (2j * np.pi)
    assert abs(w - 1) < (tol_w if tol_w else 10 * tol)
    a = 0.5 * np.sum(np.conj(z - inside_pt) * dz).imag
    assert abs(a - area) < 10 * tol * area


def test_ellipse():
    import qpms_amd
    L = lib()
    p = L.beyn_contour_ellipse(qpms_amd.c_cplx.of(2.05 + 0.01j), 0.05, 0.005, 96)
    z, dz, fn = read_contour(p)
    zo, dzo = ob.contour_ellipse(2.05 + 0.01j, 0.05, 0.005, 96)
    assert np.allclose(z, zo, rtol=0, atol=1e-15) and np.allclose(dz, dzo, rtol=0, atol=1e-17)
    assert fn  # inside_test set
    free(p)
    check_integrals(zo, dzo, 2.05 + 0.01j, np.pi * 0.05 * 0.005, 1e-13, 1e-3)


@pytest.mark.parametrize("orient", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [31, 64])
def test_halfellipse(orient, n):
    import qpms_amd
    L = lib()
    c = 1.5 - 0.5j
    p = L.beyn_contour_halfellipse(qpms_amd.c_cplx.of(c), 0.4, 0.3, n, orient)
    z, dz, _ = read_contour(p)
    free(p)
    zo, dzo = ob.contour_halfellipse(c, 0.4, 0.3, n, orient)
    assert np.allclose(z, zo, rtol=0, atol=1e-15) and np.allclose(dz, dzo, rtol=0, atol=1e-15)
    f, l, h = ob._orient(orient, 0.4, 0.3)
    check_integrals(z, dz, c + f * 0.3j * h, np.pi * l * h / 2, (np.pi / (n - n // 2)) ** 2)


@pytest.mark.parametrize("orient", [0, 1, 2, 3])
@pytest.mark.parametrize("rounding", [0.0, 0.1, 0.3])
def test_kidney(orient, rounding):
    import qpms_amd
    L = lib()
    c = -0.2 + 1j
    n = 200
    p = L.beyn_contour_kidney(qpms_amd.c_cplx.of(c), 0.5, 0.25, rounding, n, orient)
    z, dz, _ = read_contour(p)
    free(p)
    zo, dzo = ob.contour_kidney(c, 0.5, 0.25, rounding, n, orient)
    assert len(zo) == n
    assert np.allclose(z, zo, rtol=0, atol=1e-14) and np.allclose(dz, dzo, rtol=0, atol=1e-14)
    f, l, h = ob._orient(orient, 0.5, 0.25)
    zz = np.concatenate([z, z[:1]])
    area_poly = 0.5 * np.sum((np.conj(zz[:-1]) * (zz[1:] - zz[:-1])).imag)
    check_integrals(z, dz, c + f * 0.5j * h, abs(area_poly), 2.0 / n)
    assert abs(area_poly) <= np.pi * l * h / 2 * (1 + 1e-9)
