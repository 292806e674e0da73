// Minimal complex<double>/complex<float> value types for device code.
//
// The kernels keep complex numbers as two registers (re, im) and spell the
// arithmetic out, so that the compiler sees plain FMAs; there is no
// dependency on hip_complex or std::complex. The header also compiles with a
// plain host C++ compiler (QPMS_HD expands to nothing), which is how the unit
// harness exercises the same special-function code on the CPU.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define QPMS_HD __host__ __device__
#define QPMS_DEV __device__
#else
#include <cmath>
#define QPMS_HD
#define QPMS_DEV
#endif

#include <math.h>

struct cplx {
  double re, im;
};

QPMS_HD inline cplx mk(double r, double i = 0.0) { cplx z; z.re = r; z.im = i; return z; }
QPMS_HD inline cplx operator+(cplx a, cplx b) { return mk(a.re + b.re, a.im + b.im); }
QPMS_HD inline cplx operator-(cplx a, cplx b) { return mk(a.re - b.re, a.im - b.im); }
QPMS_HD inline cplx operator-(cplx a) { return mk(-a.re, -a.im); }
QPMS_HD inline cplx operator*(cplx a, cplx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
QPMS_HD inline cplx operator*(double s, cplx a) { return mk(s * a.re, s * a.im); }
QPMS_HD inline cplx operator*(cplx a, double s) { return mk(s * a.re, s * a.im); }
QPMS_HD inline cplx operator/(cplx a, double s) { return mk(a.re / s, a.im / s); }
QPMS_HD inline cplx conj(cplx a) { return mk(a.re, -a.im); }
QPMS_HD inline double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
QPMS_HD inline double cabsd(cplx a) { return hypot(a.re, a.im); }
// Smith's algorithm: no overflow for large |b|, correct to a few ulp.
QPMS_HD inline cplx operator/(cplx a, cplx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return mk((a.re + a.im * r) / d, (a.im - a.re * r) / d);
  } else {
    double r = b.re / b.im, d = b.re * r + b.im;
    return mk((a.re * r + a.im) / d, (a.im * r - a.re) / d);
  }
}
QPMS_HD inline cplx cinv(cplx b) { return mk(1.0, 0.0) / b; }
QPMS_HD inline cplx cexpi(double phi) { return mk(cos(phi), sin(phi)); }
QPMS_HD inline cplx cexp_(cplx z) {
  double e = exp(z.re);
  return mk(e * cos(z.im), e * sin(z.im));
}
QPMS_HD inline cplx csin_(cplx z) {
  return mk(sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im));
}
QPMS_HD inline cplx ccos_(cplx z) {
  return mk(cos(z.re) * cosh(z.im), -sin(z.re) * sinh(z.im));
}
// Principal square root (branch cut on the negative real axis, sign of the
// imaginary part follows the sign bit of Im z, as C99 csqrt).
QPMS_HD inline cplx csqrt_(cplx z) {
  if (z.re == 0.0 && z.im == 0.0) return mk(0.0, z.im);
  double t = sqrt(0.5 * (fabs(z.re) + hypot(z.re, z.im)));
  if (z.re >= 0.0) return mk(t, z.im / (2.0 * t));
  return mk(fabs(z.im) / (2.0 * t), copysign(t, z.im));
}
// Fused multiply-accumulate acc += a*b.
QPMS_HD inline void cfma(cplx &acc, cplx a, cplx b) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
}
