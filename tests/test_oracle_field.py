"""Pins of the oracle's incident-field and Mie restatements against what the
mathematics fixes, independently of the oracle's own formulas:

* Mie reflection coefficients equal minus the Bohren–Huffman a_n, b_n
  (electric, magnetic) computed here from scipy's spherical Bessel functions;
* the power-normalised plane-wave expansion carries sum_m |a_lm|^2 =
  2 pi (2l+1) |E|^2 in each degree, for normal and oblique incidence;
* the expansion about one point, translated with the regular translation
  operator (already pinned against the reference's Gaunt coefficients), is
  the expansion about another point — this fixes every coefficient's phase;
* the extinction cross-section of one sphere through the full pipeline equals
  the Mie closed form (2 pi / k^2) sum (2n+1) Re(a_n + b_n), for any
  incidence direction.

The field of a coefficient vector at points (the scattered-field sum):

* regular waves with plane-wave coefficients reproduce the plane wave near
  the expansion point, to the truncation error (k d)^lMax;
* the reference's translation-based variant, restated with the translation
  arrays (pinned above), equals the direct VSWF sum of the coefficients times
  (-1)^(l+1) (electric) and (-1)^l (magnetic): the translation runs from the
  particle to the point with the vector reversed;
* far field: the forward amplitude gives the extinction cross-section of the
  coefficient formula (optical theorem), and the integral of |E_s|^2 over a
  far sphere gives the scattering cross-section, equal to the extinction for
  lossless spheres.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn

import oracle_ffi as of


def bohren_huffman(nmax, x, mr):
    out = []
    for n in range(1, nmax + 1):
        jx, jmx = spherical_jn(n, x), spherical_jn(n, mr * x)
        djx, djmx = spherical_jn(n, x, True), spherical_jn(n, mr * x, True)
        hx = jx + 1j * spherical_yn(n, x)
        dhx = djx + 1j * spherical_yn(n, x, True)
        psi, dpsi = x * jx, jx + x * djx
        psim, dpsim = mr * x * jmx, jmx + mr * x * djmx
        xi, dxi = x * hx, hx + x * dhx
        an = (mr * psim * dpsi - psi * dpsim) / (mr * psim * dxi - xi * dpsim)
        bn = (psim * dpsi - mr * psi * dpsim) / (psim * dxi - mr * xi * dpsim)
        out.append((an, bn))
    return np.array(out)


@pytest.mark.parametrize("eps_in,radius,eV", [(-12 + 1j, 50e-9, 2.0), (4.0 + 0.1j, 80e-9, 1.7), (-20 + 2j, 30e-9, 2.4)])
def test_mie_is_minus_bohren_huffman(eps_in, radius, eV):
    om = eV * of.EV
    ke = np.sqrt(2.31) * om / of.C0
    ki = np.sqrt(complex(eps_in)) * om / of.C0
    RV, RH = of.mie(5, radius, ki, ke)
    ab = bohren_huffman(5, ke * radius, ki / ke)
    assert np.allclose(RV[1:], -ab[:, 0], rtol=1e-11, atol=1e-16)
    assert np.allclose(RH[1:], -ab[:, 1], rtol=1e-11, atol=1e-16)


DIRS = [((0, 0, 1.0), (1, 0, 0)), ((0, 0, -1.0), (0, 1j, 0)), ((0.3, 0.4, np.sqrt(0.75)), None),
        ((1.0, 0, 0), (0, 0.6, 0.8j)), ((-0.2, 0.7, -np.sqrt(1 - 0.53)), None)]


def _dir(kd, E):
    kd = np.asarray(kd, float)
    if E is None:
        E = np.cross(kd, [0, 0, 1.0])
        E = E / np.linalg.norm(E)
    return kd, np.asarray(E, complex)


@pytest.mark.parametrize("case", range(len(DIRS)))
def test_planewave_power_per_degree(case):
    kd, E = _dir(*DIRS[case])
    lMax = 5
    a = of.planewave(lMax, [[0, 0, 0]], kd * 7.0, E)[0]
    ne = lMax * (lMax + 2)
    for l in range(1, lMax + 1):
        sl = slice(l * l - 1, (l + 1) ** 2 - 1)
        for blk in (a[:ne], a[ne:]):
            assert abs(np.sum(abs(blk[sl]) ** 2) - 2 * np.pi * (2 * l + 1) * np.vdot(E, E).real) < 1e-11 * (2 * l + 1)


@pytest.mark.parametrize("case", range(len(DIRS)))
def test_planewave_translation_consistency(case):
    """a(r1) = S_reg(r1 <- r0) a(r0): regular translation of the truncated
    source expansion (lMax 14) reproduces the low-degree coefficients."""
    kd, E = _dir(*DIRS[case])
    k = 1.3e7
    r0 = np.array([20e-9, -35e-9, 10e-9])
    r1 = np.array([-40e-9, 15e-9, 55e-9])
    Ls, Ld = 14, 3
    a0 = of.planewave(Ls, [r0], kd * k, E)[0]
    a1 = of.planewave(Ld, [r1], kd * k, E)[0]
    oc = of.OracleCalc(Ls)
    d = (r1 - r0) * k
    kr = np.linalg.norm(d)
    A, B = oc.get_AB_arrays(kr, np.arccos(d[2] / kr), np.arctan2(d[1], d[0]), J=1)
    nes, ned = Ls * (Ls + 2), Ld * (Ld + 2)
    el0, mg0 = a0[:nes], a0[nes:]
    el1 = A[:ned] @ el0 + B[:ned] @ mg0
    mg1 = B[:ned] @ el0 + A[:ned] @ mg0
    scale = np.max(np.abs(a1))
    assert np.max(np.abs(el1 - a1[:ned])) < 1e-10 * scale
    assert np.max(np.abs(mg1 - a1[ned:])) < 1e-10 * scale


@pytest.mark.parametrize("case", range(len(DIRS)))
def test_single_sphere_extinction_closed_form(case):
    kd, E = _dir(*DIRS[case])
    E = E / np.sqrt(np.vdot(E, E).real)
    lMax, radius, eps_in, eps_bg, eV = 4, 60e-9, -12 + 1j, 2.31, 2.0
    r = of.finite_cluster(np.array([[10e-9, 5e-9, -3e-9]]), lMax, radius, eps_in, eps_bg, eV, kd, E)
    k = r["k"].real
    ab = bohren_huffman(lMax, k * radius, np.sqrt(eps_in) / np.sqrt(eps_bg))
    ref = 2 * np.pi / k ** 2 * np.sum((2 * np.arange(1, lMax + 1) + 1) * (ab[:, 0] + ab[:, 1]).real)
    assert abs(r["sigma_ext"] - ref) < 1e-11 * abs(ref)


def test_cluster_optical_theorem_lossless():
    """Lossless spheres: extinction = scattering, sigma_sca = f^H (W_reg + I) f / k^2
    (misc/finiterectlat-scatter.py:98,110), with W_reg the REGULAR translation matrix."""
    pos = np.array([[i * 300e-9, j * 300e-9, 0] for i in range(2) for j in range(2)])
    lMax = 4
    r = of.finite_cluster(pos, lMax, 50e-9, 6.0, 2.31, 2.0, (0.0, 0.6, 0.8), (1.0, 0, 0))
    oc = of.OracleCalc(lMax)
    Wreg = oc.build_W_rows(pos, r["k"], 0, len(pos), J=1)
    f = r["f"]
    sca = np.vdot(f, (Wreg + np.eye(len(f))) @ f).real / r["k"].real ** 2
    assert abs(sca - r["sigma_ext"]) < 1e-6 * abs(r["sigma_ext"])


@pytest.mark.parametrize("lMax", [2, 3, 4])
def test_regular_field_is_the_plane_wave(lMax):
    rng = np.random.default_rng(lMax)
    k = 1.1e7
    kd, E = _dir(*DIRS[2])
    pos = np.array([[10e-9, -20e-9, 5e-9]])
    a = of.planewave(lMax, pos, k * kd, E)
    for d in (3e-9, 3e-10):
        u = rng.normal(size=(8, 3))
        pts = pos + d * u / np.linalg.norm(u, axis=1)[:, None]
        got = of.scattered_E(lMax, pos, k, a.ravel(), pts, J=1)
        ref = np.exp(1j * k * (pts @ kd))[:, None] * E[None, :]
        assert np.max(np.abs(got - ref)) < (k * d) ** lMax * np.max(np.abs(ref))
    # the expansion point itself (kr = 0: only the l = 1 regular waves survive)
    got = of.scattered_E(lMax, pos, k, a.ravel(), pos, J=1)
    ref = np.exp(1j * k * (pos @ kd))[:, None] * E[None, :]
    assert np.max(np.abs(got - ref)) < 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("J", [1, 3])
@pytest.mark.parametrize("lMax", [1, 2, 4])
def test_alt_variant_is_the_parity_reversed_direct_sum(lMax, J):
    rng = np.random.default_rng(10 * lMax + J)
    ne = lMax * (lMax + 2)
    k = 1.2e7 * (1 + 0.05j)
    pos = rng.normal(size=(3, 3)) * 300e-9
    cv = rng.normal(size=(3, 2 * ne)) + 1j * rng.normal(size=(3, 2 * ne))
    pts = np.concatenate([rng.normal(size=(5, 3)) * 600e-9, pos[:1] + [[0, 0, 150e-9]]])   # incl. the z axis
    alt = of.scattered_E_alt(lMax, pos, k, cv.ravel(), pts, J=J)
    ls = np.concatenate([[l] * (2 * l + 1) for l in range(1, lMax + 1)])
    cvp = cv.copy()
    cvp[:, :ne] *= (-1.0) ** (ls + 1)
    cvp[:, ne:] *= (-1.0) ** ls
    direct = of.scattered_E(lMax, pos, k, cvp.ravel(), pts, J=J)
    assert np.max(np.abs(direct - alt)) < 1e-12 * np.max(np.abs(alt))
    # and not the plain direct sum: the variant is not a rounding-level restatement of it
    plain = of.scattered_E(lMax, pos, k, cv.ravel(), pts, J=J)
    assert np.max(np.abs(plain - alt)) > 1e-3 * np.max(np.abs(alt))


def _far_sphere(nt=40, nphi=80):
    x, w = np.polynomial.legendre.leggauss(nt)
    th = np.arccos(x)
    ph = 2 * np.pi * np.arange(nphi) / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1).reshape(-1, 3)
    wts = (w[:, None] * np.full(nphi, 2 * np.pi / nphi)[None, :]).ravel()
    return dirs, wts


@pytest.mark.parametrize("case", [0, 2, 3])
def test_far_field_optical_theorem_single_sphere(case):
    kd, E = _dir(*DIRS[case])
    E = E / np.sqrt(np.vdot(E, E).real)
    lMax, radius, eV = 4, 60e-9, 2.0
    p = np.array([[10e-9, 5e-9, -3e-9]])
    for eps_in in (-12 + 1j, 6.0):
        r = of.finite_cluster(p, lMax, radius, eps_in, 2.31, eV, kd, E)
        k = r["k"].real
        R = 1e6 / k
        fwd = p + R * kd
        Es = of.scattered_E(lMax, p, k, r["f"], fwd)[0]
        X = -1j * k * R * Es * np.exp(-1j * k * (fwd[0] @ kd))
        ext = 4 * np.pi / k ** 2 * np.vdot(E, X).real
        assert abs(ext - r["sigma_ext"]) < 1e-5 * abs(r["sigma_ext"])
        if eps_in == 6.0:
            dirs, wts = _far_sphere()
            Ef = of.scattered_E(lMax, p, k, r["f"], p + R * dirs)
            sca = np.sum(wts * np.sum(np.abs(Ef) ** 2, axis=1)) * R ** 2
            assert abs(sca - r["sigma_ext"]) < 1e-5 * abs(r["sigma_ext"])


def test_far_field_cluster_lossless():
    pos = np.array([[i * 300e-9, j * 300e-9, 0] for i in range(2) for j in range(2)])
    kd, E = np.array([0.0, 0.6, 0.8]), np.array([1.0, 0, 0], complex)
    r = of.finite_cluster(pos, 3, 50e-9, 6.0, 2.31, 2.0, kd, E)
    k = r["k"].real
    R = 1e9 / k   # far enough that the cluster's extent (k d^2 / R) does not shift the phase
    X = -1j * k * R * of.scattered_E(3, pos, k, r["f"], [R * kd])[0] * np.exp(-1j * k * R)
    ext = 4 * np.pi / k ** 2 * np.vdot(E, X).real
    assert abs(ext - r["sigma_ext"]) < 1e-5 * abs(r["sigma_ext"])
    R = 1e6 / k
    dirs, wts = _far_sphere()
    Ef = of.scattered_E(3, pos, k, r["f"], R * dirs)
    sca = np.sum(wts * np.sum(np.abs(Ef) ** 2, axis=1)) * R ** 2
    assert abs(sca - r["sigma_ext"]) < 1e-5 * abs(r["sigma_ext"])
