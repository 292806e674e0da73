"""Pin of the oracle's e32 chain (sigma short + long + sigma0, then the A/B
combination, translations.c:926-1094) against the mathematics: with an
absorbing background (k = k0 (1 + 0.1 i)) the lattice sum of finite-pair
translation operators converges absolutely, and

    A_e32(beta, s)[y, y'] = F[y, y'] * sum_R e^{i beta.(R - s)} A(k (R - s))[y, y']

(and the same for B), where F = (-1)^{|mu - m|} when mu - m < 0 and 1
otherwise. F is the reference's convention: its sigma carries
min1pow_m_neg(m) (ewald.c:425, 883) on negative orders while the finite
operator's multipliers use P^{|mu-m|} directly. B agrees with the direct sum
only with B's own term count (QPMS_HIP_E32_B_QMAX=corrected); the reference's
default bound (A's count, translations.c:1029) drops the last term where the
two counts differ, and those elements must then differ. A dropped Ewald term,
a wrong sign or index, or a wrong sigma0 breaks these equalities."""
import ctypes

import numpy as np
import pytest

import ewald_common as ec
import oracle_ffi as of


def e32_oracle(lMax, k, beta, s, mode):
    LE = 2 * lMax + 1
    ne = lMax * (lMax + 2)
    s0 = s[0] == 0 and s[1] == 0
    R = ec.xyweb(ec.B1, ec.B2, 0, not s0, 12.5 * ec.H_R, False)
    sig = ec.oracle_sigma_short(LE, k, beta, s, R) + ec.oracle_sigma_long(LE, k, beta, s, ec.K_points())
    if s0:
        o = of.oracle()
        o.or_sigma0.argtypes = [ctypes.c_int, ctypes.c_double, of.c_cplx, ctypes.c_void_p]
        z = np.empty(1, np.complex128)
        o.or_sigma0(LE, ec.ETA, of.c_cplx.of(k), z.ctypes.data)
        sig[0] += z[0]
    o = of.oracle()
    o.or_e32_combine.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    oc = of.OracleCalc(lMax)
    A = np.empty((ne, ne), np.complex128)
    B = np.empty((ne, ne), np.complex128)
    o.or_e32_combine(oc.c, sig.ctypes.data, mode, A.ctypes.data, B.ctypes.data)
    return A, B


def direct_sum(lMax, k, beta, s, rings=70):
    ne = lMax * (lMax + 2)
    oc = of.OracleCalc(lMax)
    SA = np.zeros((ne, ne), np.complex128)
    SB = np.zeros((ne, ne), np.complex128)
    for Rx, Ry in ec.xyweb(ec.B1, ec.B2, 0, True, rings * ec.H_R, False):
        d = np.array([Rx - s[0], Ry - s[1]])
        r = np.hypot(*d)
        if r == 0:
            continue
        A, B = oc.get_AB_arrays(k * r, np.pi / 2, np.arctan2(d[1], d[0]), J=3)
        ph = np.exp(1j * (beta[0] * d[0] + beta[1] * d[1]))
        SA += ph * A
        SB += ph * B
    return SA, SB


def sign_convention(lMax):
    ne = lMax * (lMax + 2)
    F = np.ones((ne, ne))
    for y in range(ne):
        n = int(np.sqrt(y + 1)); m = y - n * (n + 1) + 1
        for yu in range(ne):
            nu = int(np.sqrt(yu + 1)); mu = yu - nu * (nu + 1) + 1
            if mu - m < 0 and (mu - m) % 2:
                F[y, yu] = -1
    return F


@pytest.mark.parametrize("s", [(0.0, 0.0), (0.7 * ec.A_LAT / np.sqrt(3), 0.05e-7)])
def test_e32_equals_absorbing_lattice_sum(s):
    lMax = 2
    k = ec.wavenumber(2.0) * (1 + 0.1j)
    beta = (0.3e7, -0.1e7)
    SA, SB = direct_sum(lMax, k, beta, s)
    F = sign_convention(lMax)
    A1, B1 = e32_oracle(lMax, k, beta, s, 1)
    scale_a, scale_b = np.max(np.abs(SA)), np.max(np.abs(SB))
    assert np.max(np.abs(A1 - F * SA)) < 1e-9 * scale_a
    assert np.max(np.abs(B1 - F * SB)) < 1e-9 * scale_b
    # the reference's default B bound: equal where A's and B's term counts agree
    A0, B0 = e32_oracle(lMax, k, beta, s, 0)
    assert np.array_equal(A0, A1)
    oc = of.OracleCalc(lMax)
    ne = lMax * (lMax + 2)
    differs = np.zeros((ne, ne), bool)
    for y in range(ne):
        for yu in range(ne):
            ma, mb = oc.multipliers(y, yu)
            differs[y, yu] = len(mb) > len(ma)
    same = ~differs
    assert np.max(np.abs(B0[same] - (F * SB)[same])) < 1e-9 * scale_b
    if differs.any():
        assert np.max(np.abs(B0[differs] - (F * SB)[differs])) > 1e-6 * scale_b
