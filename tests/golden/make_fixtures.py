"""Generates the golden vectors under tests/golden/ from mpmath (40 digits),
sympy (exact rationals) and scipy — independent of the reference's code and of
this repository's implementations. Run: python tests/golden/make_fixtures.py
"""
import os

import mpmath as mp
import numpy as np
import sympy
from sympy.physics.wigner import wigner_3j

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 40


def sph_bessel_fixtures():
    xs = [1e-3, 0.3, 0.49, 0.51, 1.5, 3.0, 6.9, 7.1, 10.0, 17.5, 30.0, 64.0, 105.0]
    zs = [1.5 + 0.01j, 3 - 0.2j, 7.5 + 0.3j, 0.3 + 0.1j, 20 + 0.5j, 10.0 + 1e-6j, 0.9 + 0.9j]
    L = 17
    args = np.array([complex(x) for x in xs] + zs)
    J = np.empty((len(args), L + 1), np.complex128)
    Y = np.empty((len(args), L + 1), np.complex128)
    for i, z in enumerate(args):
        zz = mp.mpf(z.real) if z.imag == 0 else mp.mpc(z.real, z.imag)
        for l in range(L + 1):
            pre = mp.sqrt(mp.pi / (2 * zz))
            J[i, l] = complex(pre * mp.besselj(l + 0.5, zz))
            Y[i, l] = complex(pre * mp.bessely(l + 0.5, zz))
    return dict(bessel_args=args, bessel_j=J, bessel_y=Y)


def hankel_fixtures():
    """h^(1)_l(z) at complex z with large |Im z|, where j + i y cancels."""
    zs = [20 + 8j, 50 + 20j, 5 + 12j, 0.5 + 3j, 100 + 35j, 2 - 6j, 30 - 1j]
    L = 17
    H = np.empty((len(zs), L + 1), np.complex128)
    with mp.workdps(80):
        for i, z in enumerate(zs):
            zz = mp.mpc(z.real, z.imag)
            for l in range(L + 1):
                pre = mp.sqrt(mp.pi / (2 * zz))
                H[i, l] = complex(pre * (mp.besselj(l + 0.5, zz) + 1j * mp.bessely(l + 0.5, zz)))
    return dict(hankel_args=np.array(zs), hankel_h1=H)


def legendre_fixtures():
    # exact Ferrers functions with the Condon-Shortley phase at rational x
    xs = [sympy.Rational(-999, 1000), sympy.Rational(-3, 10), sympy.Integer(0), sympy.Rational(1, 5),
          sympy.Rational(7, 10), sympy.Rational(99, 100)]
    L = 9
    X = sympy.Symbol("x")
    out = np.empty((len(xs), (L + 1) * (L + 2) // 2))
    for i, x in enumerate(xs):
        for l in range(L + 1):
            for m in range(l + 1):
                # Rodrigues: P_l^m = (-1)^m (1-x^2)^{m/2} d^m/dx^m P_l(x)
                pl = sympy.legendre(l, X)
                v = (-1) ** m * (1 - x ** 2) ** sympy.Rational(m, 2) * sympy.diff(pl, X, m).subs(X, x)
                out[i, l * (l + 1) // 2 + m] = float(sympy.N(v, 30))
    return dict(legendre_x=np.array([float(x) for x in xs]), legendre_P=out)


def gaunt_fixtures():
    rows = []
    for n in range(1, 5):
        for nu in range(1, 5):
            for m in range(-n, n + 1):
                for mu in range(-nu, nu + 1):
                    qmax = min(n, nu, (n + nu - abs(m + mu)) // 2)
                    for q in range(qmax + 1):
                        p = n + nu - 2 * q
                        if p < abs(m + mu):
                            continue
                        f = sympy.sqrt(sympy.factorial(n + m) * sympy.factorial(nu + mu) * sympy.factorial(p - m - mu)
                                       / (sympy.factorial(n - m) * sympy.factorial(nu - mu)
                                          * sympy.factorial(p + m + mu)))
                        a = (-1) ** (m + mu) * (2 * p + 1) * f * wigner_3j(n, nu, p, m, mu, -m - mu) \
                            * wigner_3j(n, nu, p, 0, 0, 0)
                        rows.append((m, n, mu, nu, q, float(sympy.N(a, 30))))
    arr = np.array(rows)
    threej = []
    for j1 in range(0, 5):
        for j2 in range(0, 5):
            for j3 in range(abs(j1 - j2), j1 + j2 + 2):
                for m1 in range(-j1, j1 + 1):
                    for m2 in range(-j2, j2 + 1):
                        m3 = -m1 - m2
                        threej.append((j1, j2, j3, m1, m2, m3, float(wigner_3j(j1, j2, j3, m1, m2, m3))))
    return dict(gaunt=arr, wigner3j=np.array(threej))


def ewald_fixtures():
    a = 375e-9
    eta = 2.5 / a
    rows = []
    for (R, k) in [(a / np.sqrt(3), 5.78 / a), (a, 5.78 / a), (3 * a, 4.4 / a), (0.3 * a, 6.0 / a),
                   (2 * a, 6.1 / a), (1.5 * a, 3.0 / a)]:
        for n in range(8):
            f = lambda x: mp.exp(-R * R * x * x + k * k / (4 * x * x)) * x ** (2 * n)
            v = mp.quad(f, [eta, eta + 2 / R, eta + 10 / R, eta + 40 / R])
            rows.append((R, k, n, eta, float(v)))
    gam = []
    for a_ in [0.5, -0.5, -1.5, -2.5, -3.5]:
        for z, sgn in [(0.3, 1), (2.0, 1), (30.0, 1), (-0.7, 1), (-0.7, -1), (1 + 2j, 1), (-1 + 0.5j, 1),
                       (0.2 - 0.1j, 1), (-1.3, -1), (5 - 3j, 1), (-0.01, -1), (-0.01, 1)]:
            z = complex(z)
            if z.imag == 0 and z.real < 0:
                zs = mp.mpc(z.real, sgn * mp.mpf(10) ** -35)
            else:
                zs = mp.mpc(z.real, z.imag)
            for m in (0, -1):
                v = mp.gammainc(a_, zs)
                if m:
                    f = mp.exp(2j * m * mp.pi * a_)
                    v = f * v + (1 - f) * mp.gamma(a_)
                gam.append((a_, z.real, z.imag, sgn, m, complex(v).real, complex(v).imag))
    return dict(sr_integrals=np.array(rows), gamma_inc=np.array(gam))


if __name__ == "__main__":
    d = {}
    d.update(sph_bessel_fixtures())
    d.update(hankel_fixtures())
    d.update(legendre_fixtures())
    d.update(gaunt_fixtures())
    d.update(ewald_fixtures())
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **d)
    print("wrote", os.path.join(HERE, "golden.npz"), {k: v.shape for k, v in d.items()})
