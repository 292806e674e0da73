// VSWF translation coefficients on gfx950.
//
//  k_ab_batch<LMAX>   one lane per displacement point: the nelem x nelem A and
//                     B arrays (translations.c:610-698).
//  k_w_build<LMAX,M>  one wave64 per (destination particle R, 64 consecutive
//                     source particles): the rows of particle R of the
//                     translation matrix W (scatsystem.c:1180-1217) or of the
//                     mode-problem matrix M = I - T W for diagonal T
//                     (scatsystem.c:1283-1343).
//  k_mie              Lorentz–Mie reflection coefficients (tmatrices.c:500-561).
//
// Mapping (k_w_build): lane = source particle C. Every lane evaluates its own
// Ferrers P_p^{|d|}(cos θ), spherical Hankel h_p(kr) (p <= 2 LMAX + 1) and
// e^{i d φ} in registers; the (y, yu, q) loops are unrolled at compile time,
// so the multipliers (identical for all lanes) are read with scalar loads and
// every per-lane operand sits in a fixed VGPR. For each destination index y
// the lane's 2 x nelem results are staged in LDS (stride nb + 1 complex to
// keep the ds_write_b128 pattern free of bank conflicts) and the wave then
// writes the two matrix rows (E row: [A | B] per block, M row: [B | A]) as
// contiguous 1 KiB non-temporal stores. Every element of W is written exactly
// once: no memset pass.
#include <hip/hip_runtime.h>
#include <utility>
#include "dev_specfunc.hpp"
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef double v2d __attribute__((ext_vector_type(2)));

namespace {

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cabsi(int a) { return a < 0 ? -a : a; }
constexpr int isqrt_c(int v) { int r = 0; while ((r + 1) * (r + 1) <= v) ++r; return r; }
constexpr int y2n(int y) { return isqrt_c(y + 1); }
constexpr int y2m(int y) { return y - y2n(y) * (y2n(y) + 1) + 1; }
// Number of A / B terms for (m,n) <- (mu,nu) (gaunt.h:6-8, translations.c:113-115).
constexpr int qA_count(int m, int n, int mu, int nu) { return cmin(n, cmin(nu, (n + nu - cabsi(mu - m)) / 2)) + 1; }
constexpr int qB_count(int m, int n, int mu, int nu) { return cmin(n, cmin(nu, (n + nu + 1 - cabsi(mu - m)) / 2)) + 1; }

template <int LMAX>
struct Dims {
  static constexpr int NE = LMAX * (LMAX + 2);
  static constexpr int NB = 2 * NE;
  static constexpr int L2 = 2 * LMAX + 1;
  static constexpr int NLEG = (L2 + 1) * (L2 + 2) / 2;
  static constexpr int QS = LMAX + 1;
};

template <int LMAX>
struct PairData {
  double P[Dims<LMAX>::NLEG];
  cplx z[Dims<LMAX>::L2 + 1];
  cplx E[2 * LMAX + 1];
};

template <int LMAX>
__device__ __forceinline__ void pair_setup(cplx kr, double theta, double phi, int kind, PairData<LMAX> &pd) {
  constexpr int L2 = Dims<LMAX>::L2;
  const double x = cos(theta);
  const double s = sqrt((1.0 - x) * (1.0 + x));
  qsf::legendre_cs<L2>(x, s, pd.P);
  qsf::sph_bessel_fill<L2>(kind, kr, pd.z);
  const cplx e1 = cexpi(phi);
  pd.E[0] = mk(1.0, 0.0);
#pragma unroll
  for (int d = 1; d <= 2 * LMAX; ++d) pd.E[d] = pd.E[d - 1] * e1;
}

// Displacement (dest - src) → (kr, θ, φ) as cart2csph (vectors.h:359-365) then kr *= k.
template <int LMAX>
__device__ __forceinline__ void pair_setup_cart(double dx, double dy, double dz, cplx k, int kind,
                                                PairData<LMAX> &pd) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double theta = (r != 0.0) ? acos(dz / r) : M_PI_2;
  const double phi = atan2(dy, dx);
  pair_setup<LMAX>(mk(r * k.re, r * k.im), theta, phi, kind, pd);
}

// A(y <- yu), B(y <- yu) for compile-time Y, YU.
template <int LMAX, int Y, int YU>
__device__ __forceinline__ void ab_elem(const PairData<LMAX> &pd, const v2d *__restrict__ mA,
                                        const v2d *__restrict__ mB, cplx &A, cplx &B) {
  constexpr int NE = Dims<LMAX>::NE, QS = Dims<LMAX>::QS;
  constexpr int n = y2n(Y), m = y2m(Y), nu = y2n(YU), mu = y2m(YU);
  constexpr int d = mu - m, ad = cabsi(d);
  constexpr int QA = qA_count(m, n, mu, nu), QB = qB_count(m, n, mu, nu);
  constexpr int base = (Y * NE + YU) * QS;
  cplx sa = mk(0.0, 0.0), sb = mk(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < QA; ++q) {
    const int p = n + nu - 2 * q;
    const v2d mq = mA[base + q];
    const cplx pz = pd.P[qsf::leg_idx(p, ad)] * pd.z[p];
    cfma(sa, mk(mq.x, mq.y), pz);
  }
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int p = n + nu - 2 * q + 1;
    const v2d mq = mB[base + q];
    const cplx pz = pd.P[qsf::leg_idx(p, ad)] * pd.z[p];
    cfma(sb, mk(mq.x, mq.y), pz);
  }
  const cplx e = (d >= 0) ? pd.E[ad] : conj(pd.E[ad]);
  A = sa * e;
  B = sb * e;
}

template <int LMAX, int Y, int... YU>
__device__ __forceinline__ void ab_row(const PairData<LMAX> &pd, const v2d *__restrict__ mA,
                                       const v2d *__restrict__ mB, cplx *A, cplx *B,
                                       std::integer_sequence<int, YU...>) {
  (ab_elem<LMAX, Y, YU>(pd, mA, mB, A[YU], B[YU]), ...);
}

// ---------------------------------------------------------------------------
// AB arrays, one lane per point
// ---------------------------------------------------------------------------
template <int LMAX, int Y>
__device__ __forceinline__ void ab_batch_row(const PairData<LMAX> &pd, const v2d *__restrict__ mA,
                                             const v2d *__restrict__ mB, v2d *__restrict__ Aout,
                                             v2d *__restrict__ Bout) {
  constexpr int NE = Dims<LMAX>::NE;
  cplx A[NE], B[NE];
  ab_row<LMAX, Y>(pd, mA, mB, A, B, std::make_integer_sequence<int, NE>{});
#pragma unroll
  for (int yu = 0; yu < NE; ++yu) {
    Aout[Y * NE + yu] = v2d{A[yu].re, A[yu].im};
    Bout[Y * NE + yu] = v2d{B[yu].re, B[yu].im};
  }
}

template <int LMAX, int... Ys>
__device__ __forceinline__ void ab_batch_all(const PairData<LMAX> &pd, const v2d *__restrict__ mA,
                                             const v2d *__restrict__ mB, v2d *Aout, v2d *Bout,
                                             std::integer_sequence<int, Ys...>) {
  (ab_batch_row<LMAX, Ys>(pd, mA, mB, Aout, Bout), ...);
}

template <int LMAX>
__global__ __launch_bounds__(64) void k_ab_batch(int npts, const v2d *__restrict__ kr,
                                                 const double *__restrict__ th, const double *__restrict__ ph,
                                                 int kind, const v2d *__restrict__ mA,
                                                 const v2d *__restrict__ mB, v2d *__restrict__ A,
                                                 v2d *__restrict__ B) {
  constexpr int NE = Dims<LMAX>::NE;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  PairData<LMAX> pd;
  const v2d k = kr[i];
  pair_setup<LMAX>(mk(k.x, k.y), th[i], ph[i], kind, pd);
  ab_batch_all<LMAX>(pd, mA, mB, A + (size_t)i * NE * NE, B + (size_t)i * NE * NE,
                     std::make_integer_sequence<int, NE>{});
}

// ---------------------------------------------------------------------------
// W / M rows of one destination particle against 64 source particles
// ---------------------------------------------------------------------------
template <int LMAX, bool MODE_M, int Y>
__device__ __forceinline__ void w_row(const PairData<LMAX> &pd, const v2d *__restrict__ mA,
                                      const v2d *__restrict__ mB, v2d *lds, int lane, int R, int Rloc,
                                      int C0, int nvalid, const v2d *__restrict__ tdiagR,
                                      v2d *__restrict__ W, size_t ld) {
  constexpr int NE = Dims<LMAX>::NE, NB = Dims<LMAX>::NB, LS = NB + 1;
  cplx A[NE], B[NE];
  ab_row<LMAX, Y>(pd, mA, mB, A, B, std::make_integer_sequence<int, NE>{});
  __syncthreads();  // previous row's LDS reads are done
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    lds[lane * LS + k] = v2d{A[k].re, A[k].im};
    lds[lane * LS + NE + k] = v2d{B[k].re, B[k].im};
  }
  __syncthreads();
  cplx sE = mk(1.0, 0.0), sM = mk(1.0, 0.0);
  if (MODE_M) {
    const v2d tE = tdiagR[Y], tM = tdiagR[NE + Y];
    sE = mk(-tE.x, -tE.y);
    sM = mk(-tM.x, -tM.y);
  }
  const int pdiag = R - C0;  // block column of the diagonal block inside this segment (if 0..63)
  v2d *rowE = W + ((size_t)Rloc * NB + Y) * ld + (size_t)C0 * NB;
  v2d *rowM = W + ((size_t)Rloc * NB + NE + Y) * ld + (size_t)C0 * NB;
  for (int g = lane; g < nvalid; g += 64) {
    const int pair = g / NB;
    const int kk = g - pair * NB;
    const int kk2 = (kk < NE) ? kk + NE : kk - NE;
    v2d vE = lds[pair * LS + kk];
    v2d vM = lds[pair * LS + kk2];
    if (pair == pdiag) {
      const double one = MODE_M ? 1.0 : 0.0;
      vE = v2d{(kk == Y) ? one : 0.0, 0.0};
      vM = v2d{(kk == NE + Y) ? one : 0.0, 0.0};
    } else if (MODE_M) {
      const cplx a = sE * mk(vE.x, vE.y), b = sM * mk(vM.x, vM.y);
      vE = v2d{a.re, a.im};
      vM = v2d{b.re, b.im};
    }
    __builtin_nontemporal_store(vE, rowE + g);
    __builtin_nontemporal_store(vM, rowM + g);
  }
}

template <int LMAX, bool MODE_M, int... Ys>
__device__ __forceinline__ void w_rows(const PairData<LMAX> &pd, const v2d *__restrict__ mA,
                                       const v2d *__restrict__ mB, v2d *lds, int lane, int R, int Rloc,
                                       int C0, int nvalid, const v2d *__restrict__ tdiagR,
                                       v2d *__restrict__ W, size_t ld, std::integer_sequence<int, Ys...>) {
  (w_row<LMAX, MODE_M, Ys>(pd, mA, mB, lds, lane, R, Rloc, C0, nvalid, tdiagR, W, ld), ...);
}

// With `node` set (Beyn quadrature chunks replayed as one graph), the
// wavenumber is node[0..1] and the T diagonals are tdiag + node_td * nb for
// the node entry node = ntab + 8 (*cursor + slot): the launch itself does not
// depend on which node it builds.
template <int LMAX, bool MODE_M>
__global__ __launch_bounds__(64) void k_w_build(const double *__restrict__ pos, int npart, int row_p0,
                                                int nchunks, double k_re, double k_im, int kind,
                                                const v2d *__restrict__ mA, const v2d *__restrict__ mB,
                                                const int *__restrict__ ttype,
                                                const v2d *__restrict__ tdiag, v2d *__restrict__ W,
                                                size_t ld, const double *__restrict__ ntab,
                                                const int *__restrict__ cursor, int slot) {
  constexpr int NE = Dims<LMAX>::NE, NB = Dims<LMAX>::NB;
  __shared__ v2d lds[64 * (NB + 1)];
  const int lane = threadIdx.x;
  const int R = row_p0 + blockIdx.x / nchunks;
  const int C0 = (blockIdx.x % nchunks) * 64;
  const int C = C0 + lane;
  const int nsrc = min(64, npart - C0);
  const int nvalid = nsrc * NB;
  const double rx = pos[3 * R], ry = pos[3 * R + 1], rz = pos[3 * R + 2];
  double dx = 1.0, dy = 0.0, dz = 0.0;  // dummy for lanes without a valid off-diagonal source
  if (C < npart && C != R) {
    dx = rx - pos[3 * C];
    dy = ry - pos[3 * C + 1];
    dz = rz - pos[3 * C + 2];
  }
  if (ntab) {
    const double *nd = ntab + 8 * (size_t)(cursor[0] + slot);
    k_re = nd[0];
    k_im = nd[1];
    tdiag += (size_t)(cursor[0] + slot) * NB;
  }
  PairData<LMAX> pd;
  pair_setup_cart<LMAX>(dx, dy, dz, mk(k_re, k_im), kind, pd);
  const v2d *tdiagR = MODE_M ? tdiag + (size_t)ttype[R] * NB : nullptr;
  w_rows<LMAX, MODE_M>(pd, mA, mB, lds, lane, R, R - row_p0, C0, nvalid, tdiagR, W, ld,
                       std::make_integer_sequence<int, NE>{});
}

// ---------------------------------------------------------------------------
// Mie reflection coefficients: params per sphere (a, k_i, k_e, mu_i, mu_e) as
// 9 doubles; output RV[l], RH[l] for l = 0..lMax (complex), per sphere.
// ---------------------------------------------------------------------------
constexpr int MIE_L = 17;
__global__ void k_mie(int lMax, int nsph, const double *__restrict__ prm, v2d *__restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsph) return;
  const double *p = prm + 9 * s;
  const double a = p[0];
  const cplx ki = mk(p[1], p[2]), ke = mk(p[3], p[4]), mui = mk(p[5], p[6]), mue = mk(p[7], p[8]);
  const cplx xi = a * ki, xe = a * ke;
  const cplx eti = ki / mui, ete = ke / mue;
  cplx zi[MIE_L + 1], ze[MIE_L + 1], zs[MIE_L + 1];
  qsf::sph_bessel_fill<MIE_L>(1, xi, zi);
  qsf::sph_bessel_fill<MIE_L>(1, xe, ze);
  qsf::sph_bessel_fill<MIE_L>(3, xe, zs);
  for (int l = 0; l <= lMax && l < MIE_L; ++l) {
    // DLMF 10.51.2: z_l' = -z_{l+1} + (l/x) z_l
    const cplx dzi = -zi[l + 1] + (double)l * zi[l] / xi;
    const cplx dze = -ze[l + 1] + (double)l * ze[l] / xe;
    const cplx dzs = -zs[l + 1] + (double)l * zs[l] / xe;
    const cplx fi = zi[l] / xi + dzi, fs = zs[l] / xe + dzs, fe = ze[l] / xe + dze;
    const cplx RV = (-(eti * fe * zi[l]) + ete * ze[l] * fi) / (-(ete * fi * zs[l]) + eti * zi[l] * fs);
    const cplx RH = (-(ete * fe * zi[l]) + eti * ze[l] * fi) / (-(eti * fi * zs[l]) + ete * zi[l] * fs);
    out[(size_t)s * 2 * (lMax + 1) + l] = v2d{RV.re, RV.im};
    out[(size_t)s * 2 * (lMax + 1) + lMax + 1 + l] = v2d{RH.re, RH.im};
  }
}

}  // namespace

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

extern "C" int qh_launch_ab_batch(const qh_trans_tables *t, size_t npts, const double *kr_dev,
                                  const double *theta_dev, const double *phi_dev, int kind,
                                  double *A_dev, double *B_dev, void *stream) {
  if (npts == 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const dim3 grid((unsigned)((npts + 63) / 64)), block(64);
  const v2d *mA = (const v2d *)t->dA, *mB = (const v2d *)t->dB;
  const v2d *kr = (const v2d *)kr_dev;
  v2d *A = (v2d *)A_dev, *B = (v2d *)B_dev;
  switch (t->lMax) {
    case 1: QH_LAUNCH(k_ab_batch<1>, grid, block, 0, s, (int)npts, kr, theta_dev, phi_dev, kind, mA, mB, A, B); break;
    case 2: QH_LAUNCH(k_ab_batch<2>, grid, block, 0, s, (int)npts, kr, theta_dev, phi_dev, kind, mA, mB, A, B); break;
    case 3: QH_LAUNCH(k_ab_batch<3>, grid, block, 0, s, (int)npts, kr, theta_dev, phi_dev, kind, mA, mB, A, B); break;
    case 4: QH_LAUNCH(k_ab_batch<4>, grid, block, 0, s, (int)npts, kr, theta_dev, phi_dev, kind, mA, mB, A, B); break;
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "lMax %d not instantiated (1..4)", t->lMax);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

template <bool MODE_M>
static void launch_w(const qh_trans_tables *t, size_t npart, const double *pos_dev, double k_re,
                     double k_im, int kind, size_t row_p0, size_t row_np, const int *ttype,
                     const double *tdiag, double *W_dev, size_t ld, hipStream_t s,
                     const double *ntab = nullptr, const int *cursor = nullptr, int slot = 0) {
  const int nchunks = (int)((npart + 63) / 64);
  const size_t nblk = row_np * (size_t)nchunks;
  if (nblk == 0) return;
  if (nblk > 0x7fffffffu) qpms_hip_die(__FILE__, __LINE__, __func__, "grid too large");
  const dim3 grid((unsigned)nblk), block(64);
  const v2d *mA = (const v2d *)t->dA, *mB = (const v2d *)t->dB;
  const v2d *td = (const v2d *)tdiag;
  v2d *W = (v2d *)W_dev;
  switch (t->lMax) {
#define QH_W_CASE(LM) case LM: QH_LAUNCH((k_w_build<LM, MODE_M>), grid, block, 0, s, pos_dev, (int)npart, \
      (int)row_p0, nchunks, k_re, k_im, kind, mA, mB, ttype, td, W, ld, ntab, cursor, slot); break;
    QH_W_CASE(1) QH_W_CASE(2) QH_W_CASE(3) QH_W_CASE(4)
#undef QH_W_CASE
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "lMax %d not instantiated (1..4)", t->lMax);
  }
  HIPCHK(hipGetLastError());
}

extern "C" int qh_launch_w_build(const qh_trans_tables *t, size_t npart, const double *pos_dev,
                                 double k_re, double k_im, int kind, size_t row_p0, size_t row_np,
                                 double *W_dev, size_t ld, void *stream) {
  if (row_p0 + row_np > npart) qpms_hip_die(__FILE__, __LINE__, __func__, "row range exceeds particle count");
  if (ld < npart * (size_t)(2 * t->nelem)) qpms_hip_die(__FILE__, __LINE__, __func__, "leading dimension too small");
  launch_w<false>(t, npart, pos_dev, k_re, k_im, kind, row_p0, row_np, nullptr, nullptr, W_dev, ld,
                  (hipStream_t)qh_stream(stream));
  return 0;
}

extern "C" int qh_launch_modeproblem(const qh_trans_tables *t, size_t npart, const double *pos_dev,
                                     const int *ttype_dev, const double *tdiag_dev, double k_re,
                                     double k_im, double *M_dev, size_t ld, void *stream) {
  if (ld < npart * (size_t)(2 * t->nelem)) qpms_hip_die(__FILE__, __LINE__, __func__, "leading dimension too small");
  launch_w<true>(t, npart, pos_dev, k_re, k_im, QPMS_HANKEL_PLUS, 0, npart, ttype_dev, tdiag_dev, M_dev, ld,
                 (hipStream_t)qh_stream(stream));
  return 0;
}

// The mode problem of Beyn node *cursor + slot (wavenumber from the node
// table ntab, 8 doubles per node; T diagonals at tdiag + node * nb).
extern "C" int qh_launch_modeproblem_node(const qh_trans_tables *t, size_t npart, const double *pos_dev,
                                          const int *ttype_dev, const double *tdiag_dev, const double *ntab,
                                          const int *cursor, int slot, double *M_dev, size_t ld, void *stream) {
  if (ld < npart * (size_t)(2 * t->nelem)) qpms_hip_die(__FILE__, __LINE__, __func__, "leading dimension too small");
  launch_w<true>(t, npart, pos_dev, 0.0, 0.0, QPMS_HANKEL_PLUS, 0, npart, ttype_dev, tdiag_dev, M_dev, ld,
                 (hipStream_t)qh_stream(stream), ntab, cursor, slot);
  return 0;
}

extern "C" int qh_launch_mie(int lMax, int nsph, const double *params_dev, double *out_dev, void *stream) {
  if (nsph <= 0) return 0;
  if (lMax + 1 > MIE_L) qpms_hip_die(__FILE__, __LINE__, __func__, "lMax %d too large for Mie kernel", lMax);
  QH_LAUNCH(k_mie, dim3((nsph + 63) / 64), dim3(64), 0, (hipStream_t)qh_stream(stream), lMax, nsph,
                     params_dev, (v2d *)out_dev);
  HIPCHK(hipGetLastError());
  return 0;
}
