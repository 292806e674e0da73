// Incident-field and T-application kernels of the scattering solve.
//
//  k_planewave   regular-VSWF expansion of a plane wave at each particle
//                position (vswf.c:445-530, legendre.c:54-110, normalisation.h:140-210),
//                default normalisation (power-normalised, Condon–Shortley phase).
//                One thread per (particle, y); output in the default basis
//                order: electric block (y = 0..nelem-1), then magnetic block.
//  k_apply_t     f_p = T_{type(p)} a_p for dense nb x nb T per particle type
//                (scatsystem.c:1996-2011 with qpms_apply_tmatrix), one thread
//                per output row.
//  k_cdot        -Re(a^H f) partial sums for the extinction cross-section
//                (misc/finiterectlat-scatter.py:108), one block per batch item.
//  k_scattered_e field of VSWF coefficient vectors at points: per (point,
//                particle) the VSWFs of qpms_vswf_fill_csph (vswf.c:230-288)
//                at k (point - particle), contracted with the particle's
//                coefficients (vswf.c:611-655), turned to cartesian
//                (vectors.h:272-291) and summed over particles
//                (scatsystem.c:2013-2037). One thread per point; when points
//                are few, the four waves of a workgroup share 64 points and
//                split the particles.
#include <hip/hip_runtime.h>
#include "dev_specfunc.hpp"
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef double v2d __attribute__((ext_vector_type(2)));

namespace {

// Ferrers P_l^m (CS phase included) and its x-derivative for l = 1..LMAX,
// m = 0..l, scaled to the spherical-harmonic normalisation divided by
// sqrt(l(l+1)) as qpms_pitau_fill does (legendre.c:89-99).
template <int LMAX>
__device__ void pitau_lm(int l, int m, double x, double st, double &leg, double &pi, double &tau) {
  constexpr double PI_ = 3.14159265358979323846;
  const int am = m < 0 ? -m : m;
  if (fabs(x) == 1.0) {
    // DLMF 14.8.2 limits (legendre.c:65-80)
    leg = 0; pi = 0; tau = 0;
    if (m == 0) leg = ((l & 1) ? x : 1.0) * sqrt((2 * l + 1) / (4 * PI_ * l * (l + 1)));
    if (am == 1) {
      const double fl = 0.25 * sqrt((2 * l + 1) / PI_);
      const double lpar = (l & 1) ? -1.0 : 1.0;
      const double cs = -1.0;
      pi = -((x > 0) ? -1.0 : lpar) * fl * cs;
      tau = ((m > 0) ? 1.0 : -1.0) * ((x > 0) ? 1.0 : lpar) * fl * cs;
    }
    return;
  }
  double P[(LMAX + 1) * (LMAX + 2) / 2];
  qsf::legendre_cs<LMAX>(x, st, P);
  const double Plm = P[qsf::leg_idx(l, am)];
  const double Pl1 = (l - 1 >= am) ? P[qsf::leg_idx(l - 1, am)] : 0.0;
  // (1 - x^2) dP_l^m/dx = (l + m) P_{l-1}^m - l x P_l^m (DLMF 14.10.5)
  const double dP = ((l + am) * Pl1 - l * x * Plm) / ((1.0 - x) * (1.0 + x));
  // sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) / sqrt(l(l+1))
  double fr = 1.0;
  for (int t = l - am + 1; t <= l + am; ++t) fr *= t;
  double N = sqrt((2 * l + 1) / (4 * PI_ * fr)) / sqrt((double)(l * (l + 1)));
  if (m < 0 && (am & 1)) N = -N;  // negative m: (-1)^m (legendre.c:25-33)
  leg = N * Plm;
  const double legder = N * dP;
  pi = m / st * leg;
  tau = -st * legder;
}

template <int LMAX>
__global__ void k_planewave(int npart, const double *__restrict__ pos, double kx, double ky, double kz,
                            const double *__restrict__ Ecart, v2d *__restrict__ out, long ld) {
  constexpr int NE = LMAX * (LMAX + 2);
  constexpr double PI_ = 3.14159265358979323846;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npart * NE) return;
  const int p = t / NE, y = t % NE;
  int l = 1;
  while ((l + 1) * (l + 1) - 1 <= y) ++l;  // y = l(l+1) + m - 1, |m| <= l
  const int m = y - l * (l + 1) + 1;
  // direction of Re k (cart2sph, vectors.h:134-140)
  const double kr = sqrt(kx * kx + ky * ky + kz * kz);
  const double theta = kr != 0.0 ? acos(kz / kr) : PI_ / 2;
  const double phi = atan2(ky, kx);
  const double ct = cos(theta), st = sin(theta), cf = cos(phi), sf = sin(phi);
  double leg, pi, tau;
  pitau_lm<LMAX>(l, m, ct, st, leg, pi, tau);
  // E in the local geographic basis (ccart2csphvec, vectors.h:306-327)
  const cplx Ex = mk(Ecart[0], Ecart[1]), Ey = mk(Ecart[2], Ecart[3]), Ez = mk(Ecart[4], Ecart[5]);
  const cplx Et = (ct * cf) * Ex + (ct * sf) * Ey + (-st) * Ez;
  const cplx Ef = (-sf) * Ex + cf * Ey;
  // dual waves: e^{-i m phi}, derivative/m = -i e^{-i m phi} (normalisation.h:140-205)
  const cplx eimf = cexpi(-m * phi);
  const cplx deimf = (m == 0) ? mk(0.0, 0.0) : mk(eimf.im, -eimf.re);
  // A1 = (0, pi d, -tau e), A2 = (0, tau e, pi d) (vswf.c:347-381)
  const cplx dot1 = (pi * deimf) * Et + (-tau * eimf) * Ef;
  const cplx dot2 = (tau * eimf) * Et + (pi * deimf) * Ef;
  // 4 pi i^l and -4 pi i^(l+1) (vswf.c:455-462)
  const cplx il = ((l & 3) == 0) ? mk(1, 0) : ((l & 3) == 1) ? mk(0, 1) : ((l & 3) == 2) ? mk(-1, 0) : mk(0, -1);
  const cplx il1 = mk(-il.im, il.re);
  const cplx mg = (4 * PI_) * il * dot1;
  const cplx el = (-4 * PI_) * il1 * dot2;
  // phase e^{i k.r} of the expansion point (vswf.c:486-488)
  const double ph = kx * pos[3 * p] + ky * pos[3 * p + 1] + kz * pos[3 * p + 2];
  const cplx f = cexpi(ph);
  const cplx e1 = el * f, m1 = mg * f;
  out[(size_t)p * ld + y] = v2d{e1.re, e1.im};
  out[(size_t)p * ld + NE + y] = v2d{m1.re, m1.im};
}

// out[p] = T[ttype[p]] a[p], T row-major nb x nb per type.
__global__ void k_apply_t(int npart, int nb, const int *__restrict__ ttype, const v2d *__restrict__ T,
                          const v2d *__restrict__ a, v2d *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npart * nb) return;
  const int p = t / nb, r = t % nb;
  const v2d *Tr = T + ((size_t)ttype[p] * nb + r) * nb;
  const v2d *ap = a + (size_t)p * nb;
  double sr = 0, si = 0;
  for (int c = 0; c < nb; ++c) {
    const v2d x = Tr[c], v = ap[c];
    sr += x.x * v.x - x.y * v.y;
    si += x.x * v.y + x.y * v.x;
  }
  out[(size_t)p * nb + r] = v2d{sr, si};
}

// partial[b] = -Re(sum_i conj(a_i) f_i) over n entries of batch item b.
__global__ void k_neg_re_cdot(long n, const v2d *__restrict__ a, long sa, const v2d *__restrict__ f, long sf,
                              double *__restrict__ out) {
  __shared__ double red[256];
  const v2d *ab = a + (size_t)blockIdx.x * sa, *fb = f + (size_t)blockIdx.x * sf;
  double s = 0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    const v2d x = ab[i], y = fb[i];
    s += x.x * y.x + x.y * y.y;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = -red[0];
}


// One particle's field at offset (dx, dy, dz) from it, added to E (cartesian).
// c: the particle's 2 NE coefficients (electric block, then magnetic);
// parity: multiply electric coefficients by (-1)^(l+1), magnetic by (-1)^l.
// HANKEL: kind is 3 or 4 and h is formed directly at every argument (the
// outgoing field's case; the origin is then a singular point, NaN).
template <int LMAX, bool HANKEL>
__device__ __forceinline__ void field_of_particle(double dx, double dy, double dz, cplx k, int kind,
                                                  const v2d *__restrict__ c, bool parity, cplx *E) {
  constexpr int NE = LMAX * (LMAX + 2);
  constexpr double PI_ = 3.14159265358979323846;
  // direction cosines of cart2sph's (theta, phi) (vectors.h:134-140): theta =
  // pi/2 at r = 0, phi = atan2(dy, dx) (also on the z axis)
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double rho = sqrt(dx * dx + dy * dy);
  const double ct = r != 0.0 ? dz / r : cos(PI_ / 2), st = r != 0.0 ? rho / r : 1.0;
  double cf, sf;
  if (rho != 0.0) {
    cf = dx / rho;
    sf = dy / rho;
  } else {
    const double phi = atan2(dy, dx);
    cf = cos(phi);
    sf = sin(phi);
  }
  const cplx kr = k * r;
  cplx z[LMAX + 1];
  const bool origin = (kr.re == 0.0 && kr.im == 0.0);
  if (HANKEL) {
    if (origin) {   // h_l is singular at the origin (as sph_bessel_fill gives it)
#pragma unroll
      for (int l = 0; l <= LMAX; ++l) z[l] = mk(NAN, NAN);
    } else {
      qsf::sph_hankel_direct<LMAX>(kind, kr, z);
    }
  } else {
    qsf::sph_bessel_fill<LMAX>(kind, kr, z);
  }
  const cplx krinv = origin ? mk(0.0, 0.0) : cinv(kr);
  const bool pole = (fabs(ct) == 1.0);   // legendre.c:65-80
  double P[qsf::leg_n(LMAX)];
  qsf::legendre_cs<LMAX>(ct, st, P);
  cplx Er = mk(0.0, 0.0), Et = mk(0.0, 0.0), Ef = mk(0.0, 0.0);
#pragma unroll
  for (int l = 1; l <= LMAX; ++l) {
    // z_l / kr (1/3 for the regular l = 1 wave at the origin) and z_{l-1} - l z_l / kr
    const cplx besfac = origin ? mk(l == 1 ? 1.0 / 3.0 : 0.0, 0.0) : z[l] * krinv;
    const cplx besder = z[l - 1] - (double)l * besfac;
    const double pe = (parity && !(l & 1)) ? -1.0 : 1.0;   // electric: (-1)^(l+1)
    const double pm = (parity && (l & 1)) ? -1.0 : 1.0;    // magnetic: (-1)^l
    // angular sums of degree l; the radial factors multiply them once per l:
    // N = (l(l+1) leg z/kr e, tau besder e, pi besder i e), M = (0, pi z i e, -tau z e)
    cplx sLegN = mk(0.0, 0.0), sTauN = mk(0.0, 0.0), sPiN = mk(0.0, 0.0), sPiM = mk(0.0, 0.0),
         sTauM = mk(0.0, 0.0);
    cplx em = mk(1.0, 0.0);   // e^{i m phi}
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      if (m > 0) em = em * mk(cf, sf);
      // normalised leg, pi, tau for +m (legendre.c:54-110, default normalisation)
      double leg, pi, tau;
      if (pole) {
        const double fl = 0.25 * sqrt((2 * l + 1) / PI_);
        const double lpar = (l & 1) ? -1.0 : 1.0;
        leg = (m == 0) ? ((l & 1) ? ct : 1.0) * sqrt((2 * l + 1) / (4 * PI_ * l * (l + 1))) : 0.0;
        pi = (m == 1) ? ((ct > 0) ? -1.0 : lpar) * fl : 0.0;     // -(...) * fl * csphase, csphase = -1
        tau = (m == 1) ? -((ct > 0) ? 1.0 : lpar) * fl : 0.0;
      } else {
        double fr = 1.0;
#pragma unroll
        for (int t = l - m + 1; t <= l + m; ++t) fr *= t;
        const double nlm = sqrt((2 * l + 1) / (4 * PI_ * fr)) / sqrt((double)(l * (l + 1)));
        const double Plm = P[qsf::leg_idx(l, m)];
        const double Pl1 = (l - 1 >= m) ? P[qsf::leg_idx(l - 1, m)] : 0.0;
        const double dP = ((l + m) * Pl1 - l * ct * Plm) / ((1.0 - ct) * (1.0 + ct));   // DLMF 14.10.5
        leg = nlm * Plm;
        pi = m / st * leg;
        tau = -st * nlm * dP;
      }
#pragma unroll
      for (int sgn = 0; sgn < (m == 0 ? 1 : 2); ++sgn) {
        // -m: leg and tau carry (-1)^m, pi (-1)^(m+1); e^{-i m phi} = conj
        const int mm = sgn ? -m : m;
        const double fm = (sgn && (m & 1)) ? -1.0 : 1.0;
        const double lg = fm * leg, ta = fm * tau, p_ = sgn ? -fm * pi : pi;
        const cplx e = sgn ? conj(em) : em;
        const int y = l * (l + 1) + mm - 1;
        const v2d cn = c[y], cm = c[NE + y];
        const cplx u = mk(pe * cn.x, pe * cn.y) * e, v = mk(pm * cm.x, pm * cm.y) * e;
        sLegN = sLegN + lg * u;
        sTauN = sTauN + ta * u;
        sPiN = sPiN + p_ * u;
        sPiM = sPiM + p_ * v;
        sTauM = sTauM + ta * v;
      }
    }
    const cplx zs = z[l] * sPiM, bs = besder * sPiN;
    Er = Er + (double)(l * (l + 1)) * (besfac * sLegN);
    Et = Et + besder * sTauN + mk(-zs.im, zs.re);
    Ef = Ef + mk(-bs.im, bs.re) - z[l] * sTauM;
  }
  E[0] = E[0] + (st * cf) * Er + (ct * cf) * Et - sf * Ef;
  E[1] = E[1] + (st * sf) * Er + (ct * sf) * Et + cf * Ef;
  E[2] = E[2] + ct * Er - st * Et;
}

// One thread per point, 64 points per wave; the S waves of a group share
// their 64 points and split the particles (wave w of the group takes
// p = w, w + S, ...), so a particle's position and coefficients are the same
// for the whole wave (scalar loads, read once per wave instead of once per
// lane) and points too few to fill the device still give S times as many
// waves. The S partial fields of a point are added in LDS in slice order
// (deterministic). Up to lMax 3 the kernel is held to two waves per SIMD
// (256 registers, a few values spilled), which measured 15 % faster than one.
template <int LMAX, int S, bool HANKEL>
__global__ __launch_bounds__(256, LMAX <= 3 ? 2 : 1) void k_scattered_e(int npts, const double *__restrict__ where, int npart,
                                                     const double *__restrict__ pos, double k_re, double k_im,
                                                     int kind, const v2d *__restrict__ cvf, int parity,
                                                     v2d *__restrict__ out) {
  static_assert(S == 1 || S == 2 || S == 4, "slices of a 4-wave workgroup");
  constexpr int NB = 2 * LMAX * (LMAX + 2), G = 4 / S;   // point groups per workgroup
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slice = __builtin_amdgcn_readfirstlane(wave % S), grp = wave / S;
  const int t = (blockIdx.x * G + grp) * 64 + lane;
  const bool live = t < npts;
  const cplx k = mk(k_re, k_im);
  cplx E[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
  if (live) {
    const double wx = where[3 * t], wy = where[3 * t + 1], wz = where[3 * t + 2];
    for (int p = slice; p < npart; p += S)
      field_of_particle<LMAX, HANKEL>(wx - pos[3 * p], wy - pos[3 * p + 1], wz - pos[3 * p + 2], k, kind,
                                      cvf + (size_t)p * NB, parity != 0, E);
  }
  if (S > 1) {
    __shared__ double red[4][6][64];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      red[wave][2 * q][lane] = E[q].re;
      red[wave][2 * q + 1][lane] = E[q].im;
    }
    __syncthreads();
    if (slice != 0) return;
#pragma unroll
    for (int w = 1; w < S; ++w)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        E[q].re += red[wave + w][2 * q][lane];
        E[q].im += red[wave + w][2 * q + 1][lane];
      }
  }
  if (!live) return;
#pragma unroll
  for (int q = 0; q < 3; ++q) out[3 * (size_t)t + q] = v2d{E[q].re, E[q].im};
}

}  // namespace

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

extern "C" int qh_launch_planewave(int lMax, size_t npart, const double *pos_dev, const double kvec[3],
                                   const double *E_dev, double *out_dev, long ld, void *stream) {
  if (npart == 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const int ne = lMax * (lMax + 2);
  const dim3 grid((unsigned)((npart * ne + 255) / 256)), block(256);
  v2d *o = (v2d *)out_dev;
  switch (lMax) {
    case 1: QH_LAUNCH(k_planewave<1>, grid, block, 0, s, (int)npart, pos_dev, kvec[0], kvec[1], kvec[2], E_dev, o, ld); break;
    case 2: QH_LAUNCH(k_planewave<2>, grid, block, 0, s, (int)npart, pos_dev, kvec[0], kvec[1], kvec[2], E_dev, o, ld); break;
    case 3: QH_LAUNCH(k_planewave<3>, grid, block, 0, s, (int)npart, pos_dev, kvec[0], kvec[1], kvec[2], E_dev, o, ld); break;
    case 4: QH_LAUNCH(k_planewave<4>, grid, block, 0, s, (int)npart, pos_dev, kvec[0], kvec[1], kvec[2], E_dev, o, ld); break;
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "lMax %d not instantiated (1..4)", lMax);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_apply_t(size_t npart, int nb, const int *ttype_dev, const double *T_dev, const double *a_dev,
                                 double *out_dev, void *stream) {
  if (npart == 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const dim3 grid((unsigned)((npart * nb + 255) / 256)), block(256);
  QH_LAUNCH(k_apply_t, grid, block, 0, s, (int)npart, nb, ttype_dev, (const v2d *)T_dev,
                     (const v2d *)a_dev, (v2d *)out_dev);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_neg_re_cdot(int batch, long n, const double *a_dev, long sa, const double *f_dev, long sf,
                                     double *out_dev, void *stream) {
  if (batch == 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  QH_LAUNCH(k_neg_re_cdot, dim3(batch), dim3(256), 0, s, n, (const v2d *)a_dev, sa, (const v2d *)f_dev, sf,
                     out_dev);
  HIPCHK(hipGetLastError());
  return 0;
}

// Points at where[3 npts], particles at pos[3 npart], coefficients cvf[npart][2 nelem(lMax)];
// out[3 npts] complex. parity != 0 multiplies electric coefficients by (-1)^(l+1)
// and magnetic ones by (-1)^l first.
extern "C" int qh_launch_scattered_e(int lMax, size_t npts, const double *where_dev, size_t npart,
                                     const double *pos_dev, double k_re, double k_im, int kind,
                                     const double *cvf_dev, int parity, double *out_dev, void *stream) {
  if (npts == 0) return 0;
  if (npts > (size_t)INT32_MAX || npart > (size_t)INT32_MAX)
    qpms_hip_die(__FILE__, __LINE__, __func__, "too many points (%zu) or particles (%zu)", npts, npart);
  hipStream_t s = (hipStream_t)qh_stream(stream);
  // points too few to fill the device one thread each (fewer than 262,144,
  // 4096 waves): the particles split over 4 waves per 64 points
  const int S = (npart >= 4 && npts < 262144) ? 4 : 1;
  const bool hankel = kind >= 3;
  const dim3 grid((unsigned)((npts + 64 * (4 / S) - 1) / (64 * (4 / S)))), block(256);
  const v2d *c = (const v2d *)cvf_dev;
  v2d *o = (v2d *)out_dev;
#define QH_SCE1(LM, SL, H) \
  QH_LAUNCH((k_scattered_e<LM, SL, H>), grid, block, 0, s, (int)npts, where_dev, (int)npart, pos_dev, k_re, \
                     k_im, kind, c, parity, o)
#define QH_SCE(LM)                                   \
  do {                                               \
    if (S == 4) {                                    \
      if (hankel) QH_SCE1(LM, 4, true);              \
      else QH_SCE1(LM, 4, false);                    \
    } else {                                         \
      if (hankel) QH_SCE1(LM, 1, true);              \
      else QH_SCE1(LM, 1, false);                    \
    }                                                \
  } while (0)
  switch (lMax) {
    case 1: QH_SCE(1); break;
    case 2: QH_SCE(2); break;
    case 3: QH_SCE(3); break;
    case 4: QH_SCE(4); break;
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "lMax %d not instantiated (1..4)", lMax);
  }
#undef QH_SCE
#undef QH_SCE1
  HIPCHK(hipGetLastError());
  return 0;
}
