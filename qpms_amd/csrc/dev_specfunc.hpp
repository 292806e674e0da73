// Special functions evaluated in registers by the kernels.
//
//  * Ferrers (unnormalised associated Legendre) functions P_l^m(x), m >= 0,
//    with the Condon–Shortley phase, index l(l+1)/2 + m. This is the array
//    the reference obtains from gsl_sf_legendre_array_e(GSL_SF_LEGENDRE_NONE,
//    lmax, x, -1, buf) (translations.c:652-653).
//  * Spherical Bessel j_l, y_l and Hankel h_l^(1,2) = j_l ± i y_l for l = 0..L
//    at real or complex argument (bessel.c:86-147 obtains them from GSL Steed
//    arrays and Amos). Method here:
//      - |z| < 0.5: power series for j (DLMF 10.53.1), upward recurrence for y;
//      - |z| > L:   upward recurrence for both (stable for l < |z|);
//      - otherwise: ratio r = j_{L+1}/j_L from the continued fraction of the
//        three-term recurrence (DLMF 10.51.1), evaluated bottom-up from a
//        fixed depth, j_L from the Wronskian j_{L+1}y_L - j_L y_{L+1} = 1/z^2
//        (DLMF 10.50.3), then downward recurrence for j.
//    y_0 = -cos z / z and y_1 = -cos z / z^2 - sin z / z start the upward
//    recurrence of y in all cases.
#pragma once
#include "dev_complex.hpp"

namespace qsf {

// --------------------------------------------------------------------------
// Legendre
// --------------------------------------------------------------------------
QPMS_HD constexpr int leg_idx(int l, int m) { return l * (l + 1) / 2 + m; }
QPMS_HD constexpr int leg_n(int lmax) { return (lmax + 1) * (lmax + 2) / 2; }

// P[leg_idx(l,m)] for 0 <= m <= l <= LMAX; s = sqrt(1 - x^2) >= 0.
template <int LMAX>
QPMS_HD inline void legendre_cs(double x, double s, double *P) {
  double pmm = 1.0;  // P_m^m
#pragma unroll
  for (int m = 0; m <= LMAX; ++m) {
    if (m > 0) pmm = -pmm * (2 * m - 1) * s;
    P[leg_idx(m, m)] = pmm;
    if (m + 1 <= LMAX) {
      double p1 = x * (2 * m + 1) * pmm;
      P[leg_idx(m + 1, m)] = p1;
      double p0 = pmm;
#pragma unroll
      for (int l = m + 2; l <= LMAX; ++l) {
        double pl = ((2 * l - 1) * x * p1 - (l + m - 1) * p0) / (l - m);
        P[leg_idx(l, m)] = pl;
        p0 = p1;
        p1 = pl;
      }
    }
  }
}

// Runtime-lmax version (generic kernels, host helpers).
QPMS_HD inline void legendre_cs_rt(int lmax, double x, double s, double *P) {
  double pmm = 1.0;
  for (int m = 0; m <= lmax; ++m) {
    if (m > 0) pmm = -pmm * (2 * m - 1) * s;
    P[leg_idx(m, m)] = pmm;
    if (m + 1 <= lmax) {
      double p1 = x * (2 * m + 1) * pmm;
      P[leg_idx(m + 1, m)] = p1;
      double p0 = pmm;
      for (int l = m + 2; l <= lmax; ++l) {
        double pl = ((2 * l - 1) * x * p1 - (l + m - 1) * p0) / (l - m);
        P[leg_idx(l, m)] = pl;
        p0 = p1;
        p1 = pl;
      }
    }
  }
}

// --------------------------------------------------------------------------
// Spherical Bessel functions, real argument x > 0
// --------------------------------------------------------------------------
// j[0..L], y[0..L]
template <int L>
QPMS_HD inline void sph_jy_real(double x, double *j, double *y) {
  const double s = sin(x), c = cos(x);
  const double xi = 1.0 / x;
  double yl1 = -c * xi;                 // y_0
  double yl = -c * xi * xi - s * xi;    // y_1
  y[0] = yl1;
  if (L >= 1) y[1] = yl;
  double ytop = yl, ytop1 = 0.0;        // y_L, y_{L+1}
#pragma unroll
  for (int l = 1; l <= L; ++l) {
    double ynext = (2 * l + 1) * xi * yl - yl1;
    if (l + 1 <= L) y[l + 1] = ynext;
    yl1 = yl;
    yl = ynext;
  }
  if (L == 0) { ytop = y[0]; ytop1 = yl; } else { ytop = y[L]; ytop1 = yl; }
  if (x < 0.5) {
    // Power series j_l(x) = x^l/(2l+1)!! * sum_k (-x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1))
    const double hx2 = -0.5 * x * x;
    double xl = 1.0;  // x^l / (2l+1)!!
#pragma unroll
    for (int l = 0; l <= L; ++l) {
      if (l > 0) xl *= x / (2 * l + 1);
      double term = 1.0, sum = 1.0;
#pragma unroll
      for (int k = 1; k <= 12; ++k) {
        term *= hx2 / (k * (2 * l + 2 * k + 1));
        sum += term;
      }
      j[l] = xl * sum;
    }
  } else if (x > (double)L) {
    double jl1 = s * xi;               // j_0
    double jl = s * xi * xi - c * xi;  // j_1
    j[0] = jl1;
    if (L >= 1) j[1] = jl;
#pragma unroll
    for (int l = 1; l < L; ++l) {
      double jn = (2 * l + 1) * xi * jl - jl1;
      j[l + 1] = jn;
      jl1 = jl;
      jl = jn;
    }
  } else {
    // r = j_{L+1}/j_L from r_l = 1 / ((2l+1)/x - r_{l+1}), depth 32 + x.
    const int depth = 32 + (int)x;
    double r = 0.0;
    for (int l = L + depth; l >= L + 1; --l) r = 1.0 / ((2 * l + 1) * xi - r);
    double jL = 1.0 / (x * x * (r * ytop - ytop1));
    double jp1 = r * jL;
    j[L] = jL;
    double jl = jL;
#pragma unroll
    for (int l = L; l >= 1; --l) {
      double jm = (2 * l + 1) * xi * jl - jp1;
      j[l - 1] = jm;
      jp1 = jl;
      jl = jm;
    }
  }
}

// --------------------------------------------------------------------------
// Spherical Bessel functions, complex argument
// --------------------------------------------------------------------------
template <int L>
QPMS_HD inline void sph_jy_cplx(cplx z, cplx *j, cplx *y) {
  const cplx s = csin_(z), c = ccos_(z);
  const cplx zi = cinv(z);
  const cplx zi2 = zi * zi;
  cplx yl1 = -(c * zi);
  cplx yl = -(c * zi2) - s * zi;
  y[0] = yl1;
  if (L >= 1) y[1] = yl;
#pragma unroll
  for (int l = 1; l <= L; ++l) {
    cplx ynext = (double)(2 * l + 1) * (zi * yl) - yl1;
    if (l + 1 <= L) y[l + 1] = ynext;
    yl1 = yl;
    yl = ynext;
  }
  const cplx ytop = y[L], ytop1 = yl;
  const double az = cabsd(z);
  if (az < 0.5) {
    const cplx hz2 = -0.5 * (z * z);
    cplx zl = mk(1.0, 0.0);
#pragma unroll
    for (int l = 0; l <= L; ++l) {
      if (l > 0) zl = zl * z / (double)(2 * l + 1);
      cplx term = mk(1.0, 0.0), sum = mk(1.0, 0.0);
#pragma unroll
      for (int k = 1; k <= 12; ++k) {
        term = term * hz2 / (double)(k * (2 * l + 2 * k + 1));
        sum = sum + term;
      }
      j[l] = zl * sum;
    }
  } else if (az > (double)L) {
    cplx jl1 = s * zi;
    cplx jl = s * zi2 - c * zi;
    j[0] = jl1;
    if (L >= 1) j[1] = jl;
#pragma unroll
    for (int l = 1; l < L; ++l) {
      cplx jn = (double)(2 * l + 1) * (zi * jl) - jl1;
      j[l + 1] = jn;
      jl1 = jl;
      jl = jn;
    }
  } else {
    const int depth = 32 + (int)az;
    cplx r = mk(0.0, 0.0);
    for (int l = L + depth; l >= L + 1; --l) r = cinv((double)(2 * l + 1) * zi - r);
    cplx jL = cinv((z * z) * (r * ytop - ytop1));
    cplx jp1 = r * jL;
    j[L] = jL;
    cplx jl = jL;
#pragma unroll
    for (int l = L; l >= 1; --l) {
      cplx jm = (double)(2 * l + 1) * (zi * jl) - jp1;
      j[l - 1] = jm;
      jp1 = jl;
      jl = jm;
    }
  }
}

// Spherical Hankel functions h^(1) (kind 3) or h^(2) (kind 4), l = 0..L, at
// nonzero x directly (DLMF 10.49.6, 10.49.1): h0 = -i e^{iz}/z,
// h1 = -e^{iz}(z + i)/z^2 (and the conjugate forms for h^(2)), then upward
// recurrence, which is stable for h. Forming j + i y would lose everything to
// cancellation once |Im z| is large; at real z the result agrees with
// j + i y to rounding relative to |h|.
template <int L>
QPMS_HD inline void sph_hankel_direct(int kind, cplx x, cplx *out) {
  const double sg = (kind == 3) ? 1.0 : -1.0;
  const cplx e = cexp_(mk(-sg * x.im, sg * x.re));   // e^{+-iz}
  const cplx zi = cinv(x);
  cplx h0 = mk(sg * (e * zi).im, -sg * (e * zi).re);  // -+i e^{+-iz}/z
  cplx h1 = -(e * (x + mk(0.0, sg)) * (zi * zi));
  out[0] = h0;
  if (L >= 1) out[1] = h1;
#pragma unroll
  for (int l = 1; l < L; ++l) {
    const cplx h2 = (double)(2 * l + 1) * (zi * h1) - h0;
    out[l + 1] = h2;
    h0 = h1;
    h1 = h2;
  }
}

// z_l for the requested kind (1 = j, 2 = y, 3 = h+, 4 = h-), l = 0..L.
// Real argument when im == 0 exactly (the reference branches the same way,
// bessel.c:87-88).
template <int L>
QPMS_HD inline void sph_bessel_fill(int kind, cplx x, cplx *out) {
  if (x.im == 0.0) {
    double j[L + 1], y[L + 1];
    if (x.re == 0.0) {
      // Only the regular kind is finite at the origin.
#pragma unroll
      for (int l = 0; l <= L; ++l) out[l] = (kind == 1) ? mk(l == 0 ? 1.0 : 0.0, 0.0) : mk(NAN, NAN);
      return;
    }
    sph_jy_real<L>(x.re, j, y);
    // real part j (y for kind 2), imaginary part +y (h+), -y (h-) or 0
#pragma unroll
    for (int l = 0; l <= L; ++l) {
      const double re = (kind == 2) ? y[l] : j[l];
      const double im = (kind == 3) ? y[l] : ((kind == 4) ? -y[l] : 0.0);
      out[l] = mk(re, im);
    }
  } else if (kind >= 3) {
    sph_hankel_direct<L>(kind, x, out);
  } else {
    cplx j[L + 1], y[L + 1];
    sph_jy_cplx<L>(x, j, y);
#pragma unroll
    for (int l = 0; l <= L; ++l) out[l] = (kind == 1) ? j[l] : y[l];
  }
}

}  // namespace qsf
