// Device special functions for the 2D Ewald sums.
//
//  * Faddeeva w(z) = exp(-z^2) erfc(-iz), Im z >= 0, by Weideman's rational
//    expansion (J.A.C. Weideman, SIAM J. Numer. Anal. 31 (1994) 1497-1518):
//      w(z) = 2 p(Z) / (L - iz)^2 + 1 / (sqrt(pi) (L - iz)),  Z = (L + iz)/(L - iz),
//    p(Z) = sum_{m=1..N} a_m Z^{m-1}, N = 32, L = sqrt(N / sqrt 2). The a_m are
//    computed once on the host (weideman_coefficients) by the discrete Fourier
//    sum of Weideman's eq. (3.6) and live in constant memory.
//    Lower half plane: w(z) = 2 exp(-z^2) - w(-z).
//  * The short-range integrals of the Ewald sum (ewald.c:651-740 evaluates
//    them by adaptive quadrature)
//      I_n(R) = int_eta^inf exp(-R^2 xi^2 + k^2/(4 xi^2)) xi^{2n} d xi,
//    in closed form: with kappa = k/(2 eta), F = exp(-R^2 eta^2 + kappa^2),
//    wp = w(kappa + i R eta), wm = w(-kappa + i R eta)
//      I_0    = F sqrt(pi)/(4R) (wp + wm)
//      I_{-1} = F sqrt(pi)/(2ik) (wp - wm)
//    (Abramowitz & Stegun 7.4.33 with b = -ik/2 and its b-derivative), and the
//    integration-by-parts recurrence
//      I_n = [(2n-1) I_{n-1} - (k^2/2) I_{n-2} + eta^{2n-1} F] / (2 R^2).
//  * Gamma(a, z) for half-integer a = 1/2 - j (ewaldsf.c:44-205 regimes):
//    Legendre continued fraction for Re z >= 0, |z| > 0.5 (DLMF 8.9.2, modified
//    Lentz), otherwise the series Gamma(a) [1 - z^a e^{-z} sum z^k/Gamma(a+k+1)]
//    (DLMF 8.7.3); branch m: Gamma(a, z e^{2 pi i m}) = e^{2 pi i m a} Gamma(a,z)
//    + (1 - e^{2 pi i m a}) Gamma(a) (DLMF 8.2.10).
#pragma once
#include "dev_complex.hpp"

namespace qew {

constexpr int WEID_N = 32;

QPMS_HD inline cplx cexp2(cplx z) { return cexp_(z); }

// Faddeeva w(z) for Im z >= 0 given the Weideman coefficients a[0..N-1] (a_1..a_N).
QPMS_HD inline cplx faddeeva_uhp(cplx z, const double *a) {
  const double L = 4.7568284600108841;  // sqrt(32 / sqrt(2))
  const cplx iz = mk(-z.im, z.re);
  const cplx den = mk(L, 0.0) - iz;         // L - iz
  const cplx Z = (mk(L, 0.0) + iz) / den;   // (L + iz) / (L - iz)
  cplx p = mk(a[WEID_N - 1], 0.0);
  for (int m = WEID_N - 2; m >= 0; --m) p = p * Z + mk(a[m], 0.0);
  const cplx d2 = den * den;
  return 2.0 * p / d2 + mk(0.56418958354775628695, 0.0) / den;  // 1/sqrt(pi)
}

QPMS_HD inline cplx faddeeva(cplx z, const double *a) {
  if (z.im >= 0.0) return faddeeva_uhp(z, a);
  const cplx mz = -z;
  const cplx e = cexp_(-(z * z));
  return 2.0 * e - faddeeva_uhp(mz, a);
}

// I_n(R), n = 0..NMAX (complex k allowed), into out[0..NMAX].
template <int NMAX>
QPMS_HD inline void sr_integrals(double R, cplx k, double eta, const double *wa, cplx *out) {
  const cplx kap = k / (2.0 * eta);
  const cplx kap2 = kap * kap;
  const double Re = R * eta;
  const cplx F = cexp_(mk(-Re * Re + kap2.re, kap2.im));
  const cplx wp = faddeeva(mk(kap.re, kap.im + Re), wa);
  const cplx wm = faddeeva(mk(-kap.re, -kap.im + Re), wa);
  const double sqpi = 1.7724538509055160273;
  cplx Jm1 = (sqpi * (wp - wm)) / mk(-2.0 * k.im, 2.0 * k.re);  // / (2 i k)
  cplx J0 = (sqpi / (4.0 * R)) * (wp + wm);
  const cplx hk2 = 0.5 * (k * k);
  const double inv2R2 = 1.0 / (2.0 * R * R);
  out[0] = F * J0;
  double etap = eta;  // eta^{2n-1}
  cplx Jn2 = Jm1, Jn1 = J0;
#pragma unroll
  for (int n = 1; n <= NMAX; ++n) {
    const cplx Jn = ((double)(2 * n - 1) * Jn1 - hk2 * Jn2 + mk(etap, 0.0)) * inv2R2;
    out[n] = F * Jn;
    Jn2 = Jn1;
    Jn1 = Jn;
    etap *= eta * eta;
  }
}

// Gamma(a) for a = 1/2 - j.
QPMS_HD inline double gamma_half(int j) {
  double g = 1.7724538509055160273;  // Gamma(1/2)
  for (int i = 0; i < j; ++i) g /= (0.5 - (i + 1));  // Gamma(a-1) = Gamma(a)/(a-1)
  return g;
}

// Principal z^a for half-integer a = 1/2 - j: sqrt(z) * z^{-j}.
QPMS_HD inline cplx cpow_half(cplx z, int j) {
  cplx s = csqrt_(z);
  cplx zi = cinv(z);
  for (int i = 0; i < j; ++i) s = s * zi;
  return s;
}

// Principal log (branch cut on the negative real axis, sign of Im z decides).
QPMS_HD inline cplx clog_(cplx z) { return mk(log(hypot(z.re, z.im)), atan2(z.im, z.re)); }

// Gamma(1/2 - j, z) on branch m (m = 0 principal, m = -1 as used by the
// Ewald sums for Re z < 0 with Im z >= +0).
QPMS_HD inline cplx gamma_inc_half(int j, cplx z, int m) {
  const double a = 0.5 - j;
  cplx val;
  const double az = hypot(z.re, z.im);
  if (z.re >= 0.0 && az > 0.5) {
    // F(a,z) continued fraction (modified Lentz), Gamma = z^{a-1} e^{-z} F.
    const double tiny = 1e-300;
    cplx hn = mk(1.0, 0.0), Cn = mk(1.0 / tiny, 0.0), Dn = mk(1.0, 0.0);
    const cplx zi = cinv(z);
    for (int n = 2; n < 5000; ++n) {
      const cplx an = (n % 2) ? (0.5 * (n - 1)) * zi : (0.5 * n - a) * zi;
      Dn = mk(1.0, 0.0) + an * Dn;
      if (cabs2(Dn) < tiny * tiny) Dn = mk(tiny, 0.0);
      Cn = mk(1.0, 0.0) + an / Cn;
      if (cabs2(Cn) < tiny * tiny) Cn = mk(tiny, 0.0);
      Dn = cinv(Dn);
      const cplx delta = Cn * Dn;
      hn = hn * delta;
      if (cabs2(delta - mk(1.0, 0.0)) < 4.93e-32) break;  // |delta-1| < eps
    }
    const cplx lg = clog_(z);
    const cplx ex = (a - 1.0) * lg - z;
    val = hn * cexp_(ex);
  } else {
    // Gamma(a) [1 - z^a e^{-z} sum_k z^k / Gamma(a+k+1)]
    const double ga = gamma_half(j);
    double gak = ga * a;  // Gamma(a+1)
    cplx zk = mk(1.0, 0.0), sum = mk(0.0, 0.0);
    for (int kk = 0; kk < 200; ++kk) {
      const cplx term = zk / gak;
      sum = sum + term;
      if (kk > 4 && cabs2(term) < 1e-34 * cabs2(sum)) break;
      zk = zk * z;
      gak *= (a + kk + 1);
    }
    const cplx pre = cpow_half(z, j) * cexp_(-z);
    val = ga * (mk(1.0, 0.0) - pre * sum);
  }
  if (m != 0) {
    // e^{2 pi i m a} = (-1)^m for half-integer a
    const double f = (m % 2) ? -1.0 : 1.0;
    val = f * val + mk((1.0 - f) * gamma_half(j), 0.0);
  }
  return val;
}

// Gamma(1/2 - j, x) for real x > 0.5 on the principal branch: the continued
// fraction of gamma_inc_half in real arithmetic (every quantity is real
// there), for the evanescent orders of a real wavenumber. Returns the
// fraction F; Gamma = x^{-1/2-j} e^{-x} F.
QPMS_HD inline double gamma_inc_half_real_cf(int j, double x) {
  const double a = 0.5 - j;
  const double tiny = 1e-300;
  double hn = 1.0, Cn = 1.0 / tiny, Dn = 1.0;
  const double xi = 1.0 / x;
  for (int n = 2; n < 5000; ++n) {
    const double an = (n % 2) ? (0.5 * (n - 1)) * xi : (0.5 * n - a) * xi;
    Dn = 1.0 + an * Dn;
    if (fabs(Dn) < tiny) Dn = tiny;
    Cn = 1.0 + an / Cn;
    if (fabs(Cn) < tiny) Cn = tiny;
    Dn = 1.0 / Dn;
    const double delta = Cn * Dn;
    hn *= delta;
    if (fabs(delta - 1.0) < 2.2204460492503131e-16) break;
  }
  return hn;
}

// erfcx(y) = e^{y^2} erfc(y) = w(iy) for real y >= 0: Weideman's expansion on
// the imaginary axis, where Z = (L - y) / (L + y) is real and the whole
// evaluation is real arithmetic (32 FMAs, one division, no exponential).
QPMS_HD inline double erfcx_pos(double y, const double *a) {
  const double L = 4.7568284600108841;  // sqrt(32 / sqrt(2))
  const double d = 1.0 / (L + y);
  const double Z = (L - y) * d;
  double p = a[WEID_N - 1];
  for (int m = WEID_N - 2; m >= 0; --m) p = fma(p, Z, a[m]);
  return d * (2.0 * p * d + 0.56418958354775628695);  // + 1/(sqrt(pi) (L + y))
}

// e^{x} Gamma(1/2 - j, x) for real x > 0 and j = 0..J into E[0..J]:
//   E_0 = sqrt(pi) erfcx(sqrt x)       (Gamma(1/2, x) = sqrt(pi) erfc(sqrt x))
//   E_j = (E_{j-1} - x^{1/2-j}) / (1/2 - j)   (Gamma(a, x) = (Gamma(a+1, x) - x^a e^{-x}) / a).
// The recurrence loses relative digits as x grows (about a factor x / |a| per
// step: 1e-13 at x = 16, j = 4), exactly where Gamma itself carries e^{-x},
// so the absolute error of every term stays at the rounding of O(1) terms;
// for x <= 8 the relative error is below 6e-13 for j <= 4. This replaces a
// continued fraction that needs 100-340 dependent iterations near x = 1.
template <int J>
QPMS_HD inline void gamma_half_scaled(double x, const double *a, double (&E)[J + 1]) {
  const double sx = sqrt(x), xi = 1.0 / x;
  E[0] = 1.7724538509055160273 * erfcx_pos(sx, a);
  double xa = sx;  // x^{a}, a = 1/2 - j + 1 at step j
#pragma unroll
  for (int j = 1; j <= J; ++j) {
    xa *= xi;
    E[j] = (E[j - 1] - xa) / (0.5 - j);
  }
}

QPMS_HD inline double gamma_inc_half_real(int j, double x) {
  return gamma_inc_half_real_cf(j, x) * exp((-0.5 - j) * log(x) - x);
}

// clilgamma(z) (ewald.h:157-170): sqrt(z-1) sqrt(z+1) with the branches
// -pi/2 < arg(z+1) < 3pi/2 and -3pi/2 < arg(z-1) < pi/2.
QPMS_HD inline cplx clilgamma(cplx z) {
  cplx a1 = z - mk(1.0, 0.0), a2 = z + mk(1.0, 0.0);
  a2 = (a2.re < 0 && a2.im <= 0) ? -csqrt_(a2) : csqrt_(a2);
  a1 = (a1.re < 0 && a1.im >= 0) ? -csqrt_(a1) : csqrt_(a1);
  return a1 * a2;
}

}  // namespace qew
