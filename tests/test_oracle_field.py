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
