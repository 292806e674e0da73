// 2D-periodic Ewald lattice sums on gfx950 (ewald.c:260-920) and the e32
// combination into translation coefficients (translations.c:926-1094).
//
// One wave64 per (beta, k, s) "triple": lanes stride over the lattice points
// (R for the short-range part, K for the long-range part), each lane keeps the
// sigma_{n,m} partial sums in registers, and a shuffle reduction finishes the
// triple. The point lists are shared by all triples of a batch (one lattice),
// so they stream from L1/L2. All arithmetic is FP64; the short-range
// integrals use the Faddeeva closed form and the long-range part
// Gamma(1/2 - j, x^2) is evaluated in-kernel (dev_ewald.hpp).
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include "dev_ewald.hpp"
#include "dev_specfunc.hpp"
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef double v2d __attribute__((ext_vector_type(2)));

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

__constant__ double c_weid[qew::WEID_N];

namespace {

constexpr int cabsi(int a) { return a < 0 ? -a : a; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Short range (ewald.c:742-920):
//   sigma^S_{n,m} = sum_R -i (2/k)^{n+1}/sqrt(pi) |R-s|^n P_n^{|m|}(cos th)
//                   I_n(|R-s|) e^{i beta.(R-s)} e^{i m phi} (-1)^m [m < 0]
// pts: lattice points (x, y, z); the triple's shift s is subtracted.
// XY: all points and s in the plane z = 0: P_n^{|m|}(0) from leg0 and the
// odd n+m coefficients vanish. The factors that do not depend on the point,
// -i (2/k)^{n+1}/sqrt(pi), P_n^{|m|}(0) and the sign of odd negative m, are
// then applied to the sums, and only the (LE+1)(LE+2)/2 coefficients with
// n + m even are accumulated (ev_index).
// ERR: also the error estimate (2^-52 times the sum of |Re| + |Im| of the
// terms).
// ---------------------------------------------------------------------------
template <int LE> constexpr int ev_count() { return (LE + 1) * (LE + 2) / 2; }
__host__ __device__ constexpr int ev_index(int n, int m) { return n * (n + 1) / 2 + (m + n) / 2; }

template <int LE, bool XY, bool ERR>
__global__ __launch_bounds__(64, (XY && !ERR && LE <= 7) ? 2 : 1) void k_sigma_short(
    int ntrip, const double *__restrict__ tr, int npts, const double *__restrict__ pts, double eta,
    const double *__restrict__ leg0, int csphase, v2d *__restrict__ out, double *__restrict__ err) {
  constexpr int NSC = (LE + 1) * (LE + 1);
  constexpr int NA = XY ? ev_count<LE>() : NSC;
  const int t = blockIdx.x;
  if (t >= ntrip) return;
  const int lane = threadIdx.x;
  const double *T = tr + 8 * (size_t)t;  // bx by bz kre kim sx sy sz
  const cplx k = mk(T[3], T[4]);
  const bool kreal = (T[4] == 0.0);
  // -i (2/k)^{n+1} / sqrt(pi): per term (3D shifts) or on the sums (XY)
  const cplx tk = kreal ? mk(2.0 / T[3], 0.0) : 2.0 * cinv(k);
  cplx pre[XY ? 1 : LE + 1];
  if constexpr (!XY) {
    cplx p = tk;
#pragma unroll
    for (int n = 0; n <= LE; ++n) {
      pre[n] = mk(p.im, -p.re) * 0.56418958354775628695;
      p = p * tk;
    }
  }
  cplx acc[NA];
#pragma unroll
  for (int i = 0; i < NA; ++i) acc[i] = mk(0.0, 0.0);
  double aerr = 0.0;
  // Every short-range term carries exp(-R^2 eta^2 + Re kappa^2), kappa = k / (2 eta):
  // where that exponent is below -746 the factor is exactly 0 in double and
  // so is the term; such points are skipped (the sum is unchanged bit for bit).
  const double kap2re = (T[3] * T[3] - T[4] * T[4]) / (4.0 * eta * eta);
  for (int ip = lane; ip < npts; ip += 64) {
    const double rx = pts[3 * ip] - T[5], ry = pts[3 * ip + 1] - T[6], rz = pts[3 * ip + 2] - T[7];
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    // R = s: the lattice point coinciding with the particle (the origin for
    // s = 0, translations.c:981-987) carries no short-range term.
    if (r == 0.0) continue;
    if (r * r * eta * eta - kap2re > 746.0) continue;
    cplx In[LE + 1];
    qew::sr_integrals<LE>(r, k, eta, c_weid, In);
    const cplx eb = cexpi(T[0] * rx + T[1] * ry + T[2] * rz);
    // e^{i phi}, phi = atan2(ry, rx); on the z axis only m = 0 terms survive
    const double rho = sqrt(rx * rx + ry * ry);
    const cplx e1 = (rho != 0.0) ? mk(rx / rho, ry / rho) : mk(1.0, 0.0);
    if constexpr (XY) {
      // per n: b = |R-s|^n I_n e^{i beta.(R-s)} e^{-i n phi}, stepped by e^{2 i phi}
      const cplx e2 = e1 * e1, e1c = conj(e1);
      cplx ph = eb;   // e^{i beta.(R-s)} e^{-i n phi}
      double rn = 1.0, tn = 0.56418958354775628695 * cabsd(tk);   // |-i (2/k)^{n+1} / sqrt(pi)|
#pragma unroll
      for (int n = 0; n <= LE; ++n) {
        if (n > 0) ph = ph * e1c;
        cplx b = (rn * In[n]) * ph;
#pragma unroll
        for (int m = -n; m <= n; m += 2) {
          const int e = ev_index(n, m);
          acc[e] = acc[e] + b;
          if (ERR) aerr += fabs(leg0[n * (n + 1) / 2 + cabsi(m)]) * tn * (fabs(b.re) + fabs(b.im));
          b = b * e2;
        }
        rn *= r;
        if (ERR) tn *= cabsd(tk);
      }
    } else {
      double P[qsf::leg_n(LE)];
      const double x = rz / r;
      qsf::legendre_cs<LE>(x, sqrt((1.0 - x) * (1.0 + x)), P);
      cplx em[LE + 1];
      em[0] = mk(1.0, 0.0);
#pragma unroll
      for (int m = 1; m <= LE; ++m) em[m] = em[m - 1] * e1;
      double rn = 1.0;
#pragma unroll
      for (int n = 0; n <= LE; ++n) {
        const cplx base = (pre[n] * (rn * In[n])) * eb;
#pragma unroll
        for (int m = -n; m <= n; ++m) {
          const int am = cabsi(m);
          const double leg = ((csphase == 1) && (am & 1)) ? -P[n * (n + 1) / 2 + am] : P[n * (n + 1) / 2 + am];
          const cplx e = (m >= 0) ? em[am] : conj(em[am]);
          cplx s = (leg * base) * e;
          if (m < 0 && (am & 1)) s = -s;
          acc[n * (n + 1) + m] = acc[n * (n + 1) + m] + s;
          if (ERR) aerr += fabs(s.re) + fabs(s.im);
        }
        rn *= r;
      }
    }
  }
  cplx p = tk;   // (2/k)^{n+1}
#pragma unroll
  for (int n = 0; n <= LE; ++n) {
#pragma unroll
    for (int m = -n; m <= n; ++m) {
      cplx v;
      if constexpr (XY) {
        v = mk(0.0, 0.0);
        if (!((n + m) & 1)) {
          const int am = cabsi(m);
          const int e = ev_index(n, m);
          const double f = (m < 0 && (am & 1)) ? -leg0[n * (n + 1) / 2 + am] : leg0[n * (n + 1) / 2 + am];
          v = mk(wave_sum(acc[e].re), wave_sum(acc[e].im)) * (mk(p.im, -p.re) * (0.56418958354775628695 * f));
        }
      } else {
        const int i = n * (n + 1) + m;
        v = mk(wave_sum(acc[i].re), wave_sum(acc[i].im));
      }
      if (lane == 0) out[(size_t)t * NSC + n * (n + 1) + m] = v2d{v.re, v.im};
    }
    p = p * tk;
  }
  if (ERR) {
    aerr = wave_sum(aerr);
    if (lane == 0)
      for (int i = 0; i < NSC; ++i) err[(size_t)t * NSC + i] = 2.2204460492503131e-16 * aerr;
  }
}

// ---------------------------------------------------------------------------
// Long range, lattice in the xy plane, in-plane shift (ewald.c:286-503):
//   sigma^L_{n,m} = 1/(k^2 A) sum_K e^{i K.s} e^{i m alpha} P_n^{|m|}(0) (-1)^m[m<0]
//     sum_j (|beta+K|/k)^{n-2j} gamma^{2j-1} Gamma(1/2-j, x^2) c_{n,m,j},
//   gamma = clilgamma(|beta+K|/k), x = gamma k / (2 eta).
// c_{n,m,j} = -2 i^{n+1} sqrt(pi) q_{n,|m|,j} (ewald.c:141-145, the s1
// constants), q = (-1)^j a! b! / (j! (a-j)! (b-j)!), a, b = (n -+ |m|)/2: an
// integer. P_n^{|m|}(0) is a dyadic rational, so P_n^{|m|}(0) q_{n,|m|,j} is an
// exact double formed at compile time and the per-term arithmetic has no table
// loads. The factor that remains per coefficient,
//   leg0[n,|m|] / P_n^{|m|}(0) * (-2 sqrt(pi)) i^{n+1}
// (leg0 carries the Condon-Shortley phase the caller chose), is applied to the
// sums (long_out_factor).
// ---------------------------------------------------------------------------
// P_n^{am}(0) without the Condon-Shortley phase (n + am even):
// (-1)^{(n-am)/2} (n+am-1)!! / (n-am)!!
constexpr double ferrers_zero(int n, int am) {
  double num = 1.0, den = 1.0;
  for (int i = n + am - 1; i > 1; i -= 2) num *= i;
  for (int i = n - am; i > 1; i -= 2) den *= i;
  return (((n - am) / 2) & 1) ? -num / den : num / den;
}
// P_n^{am}(0) q_{n,am,j}
constexpr double long_coef_q(int n, int am, int j) {
  const int a = (n - am) / 2, b = (n + am) / 2;
  double f = 1.0;
  for (int i = 0; i < j; ++i) f *= (double)(a - i);              // a! / (a-j)!
  for (int i = 0; i < j; ++i) f = f * (double)(b - i) / (i + 1);  // times C(b, j), exact at each step
  return ferrers_zero(n, am) * ((j & 1) ? -f : f);
}
__device__ __forceinline__ cplx long_out_factor(int n, int am, const double *__restrict__ leg0) {
  const double r = leg0[n * (n + 1) / 2 + am] / ferrers_zero(n, am) * (-2.0 * 1.7724538509055160273);
  switch ((n + 1) & 3) {   // i^{n+1}
    case 0: return mk(r, 0.0);
    case 1: return mk(0.0, r);
    case 2: return mk(-r, 0.0);
    default: return mk(0.0, -r);
  }
}

// The terms of one K point times `phase` (e^{i K.s}, or 1), each without its
// coefficient's long_out_factor: emit(e, t) for the coefficients with n + m
// even, e = ev_index(n, m) (the others vanish for an in-plane shift), so a
// caller that accumulates needs no array of them. The m-independent products
// (|beta+K|/k)^{n-2j} gamma^{2j-1} Gamma(1/2-j, z) are formed once per (n, j),
// the j sums once per (n, |m|) (the sum of -m is that of m, negated for odd
// |m|), and the phase is folded into e^{i m alpha}.
// Returns false when every Gamma(1/2-j, z) underflows.
template <int LE, class Emit>
__device__ __forceinline__ bool long_point(const double *__restrict__ T, double Kx, double Ky, double eta,
                                           cplx phase, Emit emit) {
  constexpr int JM = LE / 2;
  const cplx k = mk(T[3], T[4]);
  const bool kreal = (T[4] == 0.0);
  const double bx = Kx + T[0], by = Ky + T[1];
  const double rb = sqrt(bx * bx + by * by);
  // gamma, x^2 and the incomplete-gamma branch (ewald.c:384-396)
  cplx gam, z;
  int branch;
  if (kreal) {
    const double tt = rb / k.re;
    if (tt >= 1.0) {
      gam = mk(sqrt(tt - 1.0) * sqrt(tt + 1.0), 0.0);
      const double x = gam.re * k.re / (2.0 * eta);
      z = mk(x * x, 0.0);
    } else {
      gam = mk(0.0, -(sqrt(1.0 - tt) * sqrt(1.0 + tt)));
      const double x = gam.im * k.re / (2.0 * eta);
      z = mk(-(x * x), -0.0);  // (i x)^2 lies just below the cut
    }
    branch = 0;
  } else {
    gam = qew::clilgamma(mk(rb, 0.0) / k);
    const cplx x = (gam * k) / (2.0 * eta);
    z = x * x;
    branch = (z.re < 0 && !signbit(z.im)) ? -1 : 0;
  }
  if (z.re > 740.0) return false;  // every Gamma(1/2-j, z) underflows
  cplx G[JM + 1];
  if (branch == 0 && z.im == 0.0 && z.re > 0.5) {
    // evanescent order of a real wavenumber: e^{x} Gamma(1/2 - j, x) from
    // sqrt(pi) erfcx(sqrt x) by the downward recurrence in the order, then e^{-x}
    double E[JM + 1];
    qew::gamma_half_scaled<JM>(z.re, c_weid, E);
    const double ex = exp(-z.re);
#pragma unroll
    for (int j = 0; j <= JM; ++j) G[j] = mk(E[j] * ex, 0.0);
  } else {
#pragma unroll
    for (int j = 0; j <= JM; ++j) G[j] = qew::gamma_inc_half(j, z, branch);
  }
  // gamma^{2j-1} Gamma(1/2-j, z)
  cplx gG[JM + 1];
  {
    const cplx g2 = gam * gam;
    cplx gp = cinv(gam);
#pragma unroll
    for (int j = 0; j <= JM; ++j) {
      gG[j] = gp * G[j];
      gp = (j == 0) ? gam : gp * g2;
    }
  }
  // (|beta+K|/k)^p, p = 0..LE
  const cplx rk = kreal ? mk(rb / k.re, 0.0) : mk(rb, 0.0) / k;
  cplx rkp[LE + 1];
  rkp[0] = mk(1.0, 0.0);
#pragma unroll
  for (int p = 1; p <= LE; ++p) rkp[p] = rkp[p - 1] * rk;
  // e^{i alpha} (alpha = atan2(by, bx), 0 where beta + K vanishes); per n
  // the factors phase e^{i m alpha}, m = -n, -n+2, ..., n, from
  // phase e^{-i n alpha} by steps of e^{2 i alpha} (three complex values live
  // instead of two tables of LE + 1)
  const cplx e1 = (rb != 0.0) ? mk(bx / rb, by / rb) : mk(1.0, 0.0);
  const cplx e2 = e1 * e1, e1c = conj(e1);
  cplx base = phase;   // phase e^{-i n alpha}
#pragma unroll
  for (int n = 0; n <= LE; ++n) {
    if (n > 0) base = base * e1c;
    cplx P[JM + 1];   // (|beta+K|/k)^{n-2j} gamma^{2j-1} Gamma(1/2-j, z)
#pragma unroll
    for (int j = 0; j <= n / 2; ++j) P[j] = rkp[n - 2 * j] * gG[j];
    cplx js[LE + 1];  // per |m| (n + |m| even): sum_j P_n^{|m|}(0) q_{n,|m|,j} P[j]
#pragma unroll
    for (int am = n & 1; am <= n; am += 2) {
      cplx a = P[0] * long_coef_q(n, am, 0);
#pragma unroll
      for (int j = 1; j <= (n - am) / 2; ++j) a = a + P[j] * long_coef_q(n, am, j);
      js[am] = a;
    }
    cplx em = base;
#pragma unroll
    for (int m = -n; m <= n; m += 2) {
      const int am = cabsi(m);
      const cplx t = (m < 0 && (am & 1)) ? -(js[am] * em) : js[am] * em;
      emit(ev_index(n, m), t);
      em = em * e2;
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Long range, lattice in the xy plane, in-plane shift (ewald.c:286-503):
//   sigma^L_{n,m} = 1/(k^2 A) sum_K e^{i K.s} e^{i m alpha} P_n^{|m|}(0) (-1)^m[m<0]
//     sum_j (|beta+K|/k)^{n-2j} gamma^{2j-1} Gamma(1/2-j, x^2) c_{n,m,j},
//   gamma = clilgamma(|beta+K|/k), x = gamma k / (2 eta).
// One wave per triple, lanes over K points; ERR: also the error estimate.
// ---------------------------------------------------------------------------
template <int LE, bool ERR>
__global__ __launch_bounds__(64, LE <= 7 ? 2 : 1) void k_sigma_long(int ntrip, const double *__restrict__ tr, int npts,
                                                    const double *__restrict__ Kp /* x, y */, double eta,
                                                    double area, const double *__restrict__ leg0,
                                                    v2d *__restrict__ out, double *__restrict__ err) {
  constexpr int NSC = (LE + 1) * (LE + 1);
  const int t = blockIdx.x;
  if (t >= ntrip) return;
  const int lane = threadIdx.x;
  const double *T = tr + 8 * (size_t)t;
  const cplx k = mk(T[3], T[4]);
  const bool kreal = (T[4] == 0.0);
  constexpr int NE = ev_count<LE>();
  cplx acc[NE];   // the coefficients with n + m even (ev_index); the others vanish
#pragma unroll
  for (int i = 0; i < NE; ++i) acc[i] = mk(0.0, 0.0);
  double aerr = 0.0;
  for (int ip = lane; ip < npts; ip += 64) {
    const double Kx = Kp[2 * ip], Ky = Kp[2 * ip + 1];
    long_point<LE>(T, Kx, Ky, eta, cexpi(Kx * T[5] + Ky * T[6]), [&](int e, cplx v) {
      acc[e] = acc[e] + v;
      if (ERR) aerr += fabs(v.re) + fabs(v.im);
    });
  }
  const cplx cf = kreal ? mk(1.0 / (k.re * k.re * area), 0.0) : cinv((k * k) * area);
#pragma unroll
  for (int n = 0; n <= LE; ++n)
#pragma unroll
    for (int m = -n; m <= n; ++m) {
      cplx v = mk(0.0, 0.0);
      if (!((n + m) & 1)) {
        const int e = ev_index(n, m);
        v = mk(wave_sum(acc[e].re), wave_sum(acc[e].im)) * (cf * long_out_factor(n, cabsi(m), leg0));
      }
      if (lane == 0) out[(size_t)t * NSC + n * (n + 1) + m] = v2d{v.re, v.im};
    }
  if (ERR) {
    // every out-factor has modulus 2 sqrt(pi) |leg0 / P(0)| = 2 sqrt(pi)
    aerr = wave_sum(aerr);
    if (lane == 0)
      for (int i = 0; i < NSC; ++i)
        err[(size_t)t * NSC + i] = 2.2204460492503131e-16 * aerr * (2.0 * 1.7724538509055160273) * cabsd(cf);
  }
}

// Long range for many shifts sharing (beta, k), as two passes: the terms of
// every K point once (k_long_coef, the same per-point arithmetic as
// k_sigma_long, scaled by 1/(k^2 A)), then per shift the phase sum
// sum_K e^{i K.s} F_K (k_long_shifts: one thread per shift, K points staged
// in LDS tiles shared by the block's shifts).
template <int LE>
__global__ __launch_bounds__(64) void k_long_coef(const double *__restrict__ T, int npts,
                                                   const double *__restrict__ Kp, double eta, double area,
                                                   const double *__restrict__ leg0, v2d *__restrict__ F) {
  constexpr int NE = ev_count<LE>();
  const int ip = blockIdx.x * 64 + threadIdx.x;
  if (ip >= npts) return;
  const cplx k = mk(T[3], T[4]);
  const cplx cf = (T[4] == 0.0) ? mk(1.0 / (k.re * k.re * area), 0.0) : cinv((k * k) * area);
  cplx f[NE];
  const bool live = long_point<LE>(T, Kp[2 * ip], Kp[2 * ip + 1], eta, mk(1.0, 0.0),
                                   [&](int e, cplx v) { f[e] = v; });
#pragma unroll
  for (int n = 0; n <= LE; ++n)
#pragma unroll
    for (int m = -n; m <= n; m += 2) {
      const int e = ev_index(n, m);
      const cplx v = live ? f[e] * (cf * long_out_factor(n, cabsi(m), leg0)) : mk(0.0, 0.0);
      F[(size_t)ip * NE + e] = v2d{v.re, v.im};
    }
}

constexpr int LS_KT = 16;   // K points per LDS tile
template <int LE>
__global__ __launch_bounds__(256) void k_long_shifts(int ntrip, const double *__restrict__ tr, int npts,
                                                      const double *__restrict__ Kp, const v2d *__restrict__ F,
                                                      v2d *__restrict__ out) {
  constexpr int NE = ev_count<LE>(), NSC = (LE + 1) * (LE + 1);
  __shared__ v2d sF[LS_KT * NE];
  __shared__ double sK[2 * LS_KT];
  const int t = blockIdx.x * 256 + threadIdx.x;
  const bool live = t < ntrip;
  const double sx = live ? tr[8 * (size_t)t + 5] : 0.0, sy = live ? tr[8 * (size_t)t + 6] : 0.0;
  cplx acc[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) acc[e] = mk(0.0, 0.0);
  for (int k0 = 0; k0 < npts; k0 += LS_KT) {
    const int nk = min(LS_KT, npts - k0);
    __syncthreads();
    for (int i = threadIdx.x; i < nk * NE; i += 256) sF[i] = F[(size_t)k0 * NE + i];
    if (threadIdx.x < 2 * nk) sK[threadIdx.x] = Kp[2 * k0 + threadIdx.x];
    __syncthreads();
    for (int q = 0; q < nk; ++q) {
      const cplx ph = cexpi(sK[2 * q] * sx + sK[2 * q + 1] * sy);
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const v2d f = sF[q * NE + e];
        acc[e] = acc[e] + ph * mk(f.x, f.y);
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int n = 0; n <= LE; ++n)
#pragma unroll
    for (int m = -n; m <= n; ++m) {
      const cplx v = ((n + m) & 1) ? mk(0.0, 0.0) : acc[ev_index(n, m)];
      out[(size_t)t * NSC + n * (n + 1) + m] = v2d{v.re, v.im};
    }
}

// sigma0 = Gamma(-1/2, -(k/2 eta)^2) P_0^0(0) / (2 sqrt(pi)) (ewald.c:260-273)
__global__ void k_sigma0(int n, const double *__restrict__ kk, double eta, double leg00, v2d *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx k = mk(kk[2 * i], kk[2 * i + 1]);
  cplx z;
  if (k.im == 0.0) {
    const double h = k.re / (2.0 * eta);
    z = mk(-(h * h), -0.0);  // -(k/2eta)^2 with Im = -0 for real k
  } else {
    const cplx h = k / (2.0 * eta);
    z = -(h * h);
  }
  const int branch = (z.re < 0 && !signbit(z.im)) ? -1 : 0;
  const cplx g = qew::gamma_inc_half(1, z, branch);
  const cplx v = g * (leg00 / 2.0 / 1.7724538509055160273);
  out[i] = v2d{v.re, v.im};
}

// ---------------------------------------------------------------------------
// e32 combine (translations.c:1024-1067): per triple, from sigma_total
//   A(y<-yu) = sum_{q<qA} mult_A[q] sigma[y_sc(mu-m, n+nu-2q)]
//   B(y<-yu) = sum_{q<qB'} mult_B[q] sigma[y_sc(mu-m, n+nu-2q+1)]
// qB' = qA (bq_mode 0, the reference's loop bound) or B's own count (mode 1).
// ---------------------------------------------------------------------------
__global__ void k_e32_combine(int ntrip, int nelem, int qslots, int nsc, const v2d *__restrict__ sig,
                              const v2d *__restrict__ mA, const v2d *__restrict__ mB,
                              const unsigned char *__restrict__ qA, const unsigned char *__restrict__ qB,
                              int bq_mode, v2d *__restrict__ A, v2d *__restrict__ B) {
  const int t = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrip || e >= nelem * nelem) return;
  const int y = e / nelem, yu = e % nelem;
  const int n = (int)sqrt((double)y + 1), m = y - n * (n + 1) + 1;
  const int nu = (int)sqrt((double)yu + 1), mu = yu - nu * (nu + 1) + 1;
  const v2d *s = sig + (size_t)t * nsc;
  const int na = qA[e], nb = bq_mode ? qB[e] : qA[e];
  cplx sa = mk(0, 0), sb = mk(0, 0);
  for (int q = 0; q < na; ++q) {
    const int p = n + nu - 2 * q;
    const v2d mq = mA[(size_t)e * qslots + q], sv = s[p * (p + 1) + mu - m];
    cfma(sa, mk(mq.x, mq.y), mk(sv.x, sv.y));
  }
  for (int q = 0; q < nb; ++q) {
    const int p = n + nu - 2 * q + 1;
    const v2d mq = mB[(size_t)e * qslots + q], sv = s[p * (p + 1) + mu - m];
    cfma(sb, mk(mq.x, mq.y), mk(sv.x, sv.y));
  }
  A[(size_t)t * nelem * nelem + e] = v2d{sa.re, sa.im};
  B[(size_t)t * nelem * nelem + e] = v2d{sb.re, sb.im};
}

// sigma_total = short + long (+ sigma0 on y = 0 when flag0[t])
__global__ void k_sigma_total(int ntrip, int nsc, const v2d *__restrict__ ss, const v2d *__restrict__ sl,
                              const v2d *__restrict__ s0, const int *__restrict__ flag0, int parts,
                              v2d *__restrict__ tot) {
  const int t = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrip || i >= nsc) return;
  v2d v = v2d{0, 0};
  if (parts & 2) v += ss[(size_t)t * nsc + i];
  if (parts & 1) v += sl[(size_t)t * nsc + i];
  if ((parts & 4) && i == 0 && flag0[t]) v += s0[t];
  tot[(size_t)t * nsc + i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Weideman coefficients (host): a_m, m = 1..N, of Weideman's eq. (3.6), from
// the discrete Fourier sum over t_k = L tan(theta_k / 2), theta_k = k pi / M.
// ---------------------------------------------------------------------------
static void weideman_coefficients(double *a) {
  const int N = qew::WEID_N, M = 2 * N, M2 = 2 * M;
  const double L = sqrt(N / sqrt(2.0));
  double f[2 * 2 * qew::WEID_N];
  f[0] = 0.0;
  for (int k = -M + 1; k <= M - 1; ++k) {
    const double th = k * M_PI / M, tt = L * tan(th / 2);
    f[k + M] = exp(-tt * tt) * (L * L + tt * tt);
  }
  double g[2 * 2 * qew::WEID_N];
  for (int i = 0; i < M2; ++i) g[i] = f[(i + M) % M2];  // fftshift
  for (int m = 1; m <= N; ++m) {
    long double s = 0;
    for (int i = 0; i < M2; ++i) s += (long double)g[i] * cosl(2.0L * M_PI * (long double)i * m / M2);
    a[m - 1] = (double)(s / M2);
  }
}

static std::once_flag g_weid_once;
static void ensure_weid() {
  std::call_once(g_weid_once, [] {
    double a[qew::WEID_N];
    weideman_coefficients(a);
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_weid), a, sizeof a, 0, hipMemcpyHostToDevice));
  });
}

extern "C" void qh_weideman_coefficients(double *a) { weideman_coefficients(a); }

extern "C" int qh_launch_sigma_short(int LE, int xy, int ntrip, const double *tr_dev, int npts, const double *pts_dev,
                                     double eta, const double *leg0_dev, int csphase, double *out_dev,
                                     double *err_dev, void *stream) {
  qh_stream(stream);
  ensure_weid();
  if (ntrip <= 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const dim3 g(ntrip), b(64);
#define QH_SS1(L, X, E) QH_LAUNCH((k_sigma_short<L, X, E>), g, b, 0, s, ntrip, tr_dev, npts, pts_dev, eta, leg0_dev, \
      csphase, (v2d *)out_dev, err_dev)
#define QH_SS(L) case L: \
    if (xy) { if (err_dev) QH_SS1(L, true, true); else QH_SS1(L, true, false); } \
    else { if (err_dev) QH_SS1(L, false, true); else QH_SS1(L, false, false); } \
    break;
  switch (LE) {
    QH_SS(3) QH_SS(5) QH_SS(7) QH_SS(9)
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "Ewald lMax %d not instantiated (3, 5, 7, 9)", LE);
  }
#undef QH_SS
#undef QH_SS1
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_sigma_long(int LE, int ntrip, const double *tr_dev, int npts, const double *K_dev,
                                    double eta, double area, const double *leg0_dev, double *out_dev,
                                    double *err_dev, void *stream) {
  if (ntrip <= 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  ensure_weid();   // the evanescent orders use the erfcx expansion
  const dim3 g(ntrip), b(64);
#define QH_SL(L) case L: if (err_dev) QH_LAUNCH((k_sigma_long<L, true>), g, b, 0, s, ntrip, tr_dev, npts, K_dev, eta, \
      area, leg0_dev, (v2d *)out_dev, err_dev); else QH_LAUNCH((k_sigma_long<L, false>), g, b, 0, s, ntrip, tr_dev, \
      npts, K_dev, eta, area, leg0_dev, (v2d *)out_dev, err_dev); break;
  switch (LE) {
    QH_SL(3) QH_SL(5) QH_SL(7) QH_SL(9)
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "Ewald lMax %d not instantiated (3, 5, 7, 9)", LE);
  }
#undef QH_SL
  HIPCHK(hipGetLastError());
  return 0;
}

// Long range for ntrip triples that all share (beta, k) (those of triple 0);
// F: scratch of npts * (LE+1)(LE+2)/2 complex values.
extern "C" size_t qh_sigma_long_shifts_scratch(int LE, int npts) { return (size_t)npts * (LE + 1) * (LE + 2) / 2 * 16; }
extern "C" int qh_launch_sigma_long_shifts(int LE, int ntrip, const double *tr_dev, int npts, const double *K_dev,
                                           double eta, double area, const double *leg0_dev, double *F_dev,
                                           double *out_dev, void *stream) {
  if (ntrip <= 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  ensure_weid();   // the evanescent orders use the erfcx expansion
  const dim3 gc((npts + 63) / 64), gs((ntrip + 255) / 256);
#define QH_LS(L) case L: \
    if (npts > 0) QH_LAUNCH(k_long_coef<L>, gc, dim3(64), 0, s, tr_dev, npts, K_dev, eta, area, leg0_dev, \
                            (v2d *)F_dev); \
    QH_LAUNCH(k_long_shifts<L>, gs, dim3(256), 0, s, ntrip, tr_dev, npts, K_dev, (const v2d *)F_dev, \
                       (v2d *)out_dev); break;
  switch (LE) {
    QH_LS(3) QH_LS(5) QH_LS(7) QH_LS(9)
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "Ewald lMax %d not instantiated (3, 5, 7, 9)", LE);
  }
#undef QH_LS
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_sigma0(int n, const double *k_dev, double eta, double leg00, double *out_dev, void *stream) {
  if (n <= 0) return 0;
  QH_LAUNCH(k_sigma0, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)qh_stream(stream), n, k_dev, eta, leg00,
                     (v2d *)out_dev);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_sigma_total(int ntrip, int nsc, const double *ss, const double *sl, const double *s0,
                                     const int *flag0, int parts, double *tot, void *stream) {
  if (ntrip <= 0) return 0;
  QH_LAUNCH(k_sigma_total, dim3((nsc + 63) / 64, ntrip), dim3(64), 0, (hipStream_t)qh_stream(stream), ntrip, nsc,
                     (const v2d *)ss, (const v2d *)sl, (const v2d *)s0, flag0, parts, (v2d *)tot);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_e32_combine(const qh_trans_tables *t, int ntrip, int nsc, const double *sig_dev, int bq_mode,
                                     double *A_dev, double *B_dev, void *stream) {
  if (ntrip <= 0) return 0;
  const int ne = t->nelem;
  QH_LAUNCH(k_e32_combine, dim3((ne * ne + 255) / 256, ntrip), dim3(256), 0, (hipStream_t)qh_stream(stream),
                     ntrip, ne, t->qslots, nsc, (const v2d *)sig_dev, (const v2d *)t->dA, (const v2d *)t->dB, t->dqA,
                     t->dqB, bq_mode, (v2d *)A_dev, (v2d *)B_dev);
  HIPCHK(hipGetLastError());
  return 0;
}
