// Small kernels of the serving pipeline (serve.c): T-matrix diagonal from
// the Mie coefficients, T a, identity padding of batched matrices, batched
// norms / dot products, log-determinant from an LU, the Beyn probe matrix,
// and the periodic mode-problem assembly straight from the Ewald sums.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "dev_complex.hpp"
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef double v2d __attribute__((ext_vector_type(2)));

namespace {

// tdiag[i] of the default basis (electric block, then magnetic) from the Mie
// output {RV[0..lMax], RH[0..lMax]} (tmatrices.c:580-591).
__global__ void k_expand_tdiag(int lMax, int nsets, const v2d *__restrict__ R, v2d *__restrict__ td) {
  const int ne = lMax * (lMax + 2), nb = 2 * ne;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsets * nb) return;
  const int s = t / nb, i = t % nb;
  const int y = i % ne;
  int l = 1;
  while ((l + 1) * (l + 1) - 1 <= y) ++l;
  td[t] = R[(size_t)s * 2 * (lMax + 1) + (i < ne ? 0 : lMax + 1) + l];
}

// out[z][i] = tdiag[z][i % nb] * a[z][i] for i < L, 0 for L <= i < npad.
__global__ void k_tdiag_mul(int L, int nb, int npad, const v2d *__restrict__ td, long std_, const v2d *__restrict__ a,
                            long sa, v2d *__restrict__ out, long so) {
  const int z = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  v2d r = v2d{0, 0};
  if (i < L) {
    const v2d x = td[(size_t)z * std_ + i % nb], v = a[(size_t)z * sa + i];
    r = v2d{x.x * v.x - x.y * v.y, x.x * v.y + x.y * v.x};
  }
  out[(size_t)z * so + i] = r;
}

// Rows / columns L..npad-1 of an npad x npad matrix become the identity.
__global__ void k_pad_identity(int L, int npad, v2d *__restrict__ M, long sM) {
  const int z = blockIdx.z;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = (long)npad * npad;
  if (e >= tot) return;
  const int r = (int)(e / npad), c = (int)(e % npad);
  if (r < L && c < L) return;
  M[(size_t)z * sM + e] = (r == c) ? v2d{1, 0} : v2d{0, 0};
}

// sum |x_z|^2 over n entries in two deterministic passes: P partial sums
// per batch item (grid P x batch), then one block per item adds them.
__device__ __forceinline__ double block_sum256(double s, double *red) {
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}
__global__ void k_sumsq_part(long n, const v2d *__restrict__ x, long sx, int P, double *__restrict__ part) {
  __shared__ double red[256];
  const v2d *xb = x + (size_t)blockIdx.y * sx;
  double s = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)P * 256) s += xb[i].x * xb[i].x + xb[i].y * xb[i].y;
  s = block_sum256(s, red);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * P + blockIdx.x] = s;
}
__global__ void k_sumsq_fin(int P, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < P; i += 256) s += part[(size_t)blockIdx.x * P + i];
  s = block_sum256(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// log|det| and arg det of an LU with 1-based pivots: sum log|u_ii|,
// sum arg u_ii + pi * (number of interchanges), folded into (-pi, pi].
__global__ void k_logdet(int n, const v2d *__restrict__ LU, const int *__restrict__ ipiv, double *__restrict__ out) {
  __shared__ double sl[256], sa[256];
  double l = 0, a = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const v2d u = LU[(size_t)i * n + i];
    l += 0.5 * log(u.x * u.x + u.y * u.y);
    a += atan2(u.y, u.x);
    if (ipiv[i] != i + 1) a += 3.14159265358979323846;
  }
  sl[threadIdx.x] = l;
  sa[threadIdx.x] = a;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sl[threadIdx.x] += sl[threadIdx.x + w]; sa[threadIdx.x] += sa[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double tp = 6.28318530717958647692;
    double ph = fmod(sa[0], tp);
    if (ph > 3.14159265358979323846) ph -= tp;
    if (ph <= -3.14159265358979323846) ph += tp;
    out[0] = sl[0];
    out[1] = ph;
  }
}

}  // namespace

// Counter-based normal deviates: element i of the probe matrix from
// splitmix64(seed, i) and Box–Muller; the host twin is qh_probe_host.
__host__ __device__ static inline uint64_t qh_splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
__host__ __device__ static inline v2d qh_probe_value(uint64_t seed, uint64_t i) {
  const uint64_t a = qh_splitmix(seed * 0x2545F4914F6CDD1DULL + 2 * i + 1), b = qh_splitmix(a ^ (2 * i + 2));
  const double u1 = ((a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double r = sqrt(-2.0 * log(u1));
  return v2d{r * cos(6.28318530717958647692 * u2), r * sin(6.28318530717958647692 * u2)};
}

namespace {
__global__ void k_probe(long n, uint64_t seed, v2d *__restrict__ V) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) V[i] = qh_probe_value(seed, (uint64_t)i);
}

// Periodic mode problem straight from the Ewald sums: block (d, s) of
// M = I - T_d W(d <- s), W from sigma_total of the pair (t = d * np + s) by
// the e32 combination (translations.c:1024-1067) in the default basis
// (same type -> A, cross type -> B). One thread per element of the block.
__global__ void k_e32_assemble_M(int np, int nelem, int qslots, int nsc, const v2d *__restrict__ sig,
                                 const v2d *__restrict__ mA, const v2d *__restrict__ mB,
                                 const unsigned char *__restrict__ qA, const unsigned char *__restrict__ qB,
                                 int bq_mode, const int *__restrict__ ttype, const v2d *__restrict__ tdiag,
                                 v2d *__restrict__ M, long ld, int t0) {
  const int nb = 2 * nelem;
  const int t = t0 + blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * np || e >= nb * nb) return;
  const int d = t / np, s = t % np;
  const int i = e / nb, j = e % nb;
  const int y = i % nelem, yu = j % nelem;
  const bool same = (i < nelem) == (j < nelem);
  const int n = (int)sqrt((double)y + 1), m = y - n * (n + 1) + 1;
  const int nu = (int)sqrt((double)yu + 1), mu = yu - nu * (nu + 1) + 1;
  const v2d *sg = sig + (size_t)t * nsc;
  const int ee = y * nelem + yu;
  cplx w = mk(0, 0);
  if (same) {
    for (int q = 0; q < qA[ee]; ++q) {
      const int p = n + nu - 2 * q;
      const v2d mq = mA[(size_t)ee * qslots + q], sv = sg[p * (p + 1) + mu - m];
      cfma(w, mk(mq.x, mq.y), mk(sv.x, sv.y));
    }
  } else {
    const int nbq = bq_mode ? qB[ee] : qA[ee];
    for (int q = 0; q < nbq; ++q) {
      const int p = n + nu - 2 * q + 1;
      const v2d mq = mB[(size_t)ee * qslots + q], sv = sg[p * (p + 1) + mu - m];
      cfma(w, mk(mq.x, mq.y), mk(sv.x, sv.y));
    }
  }
  const v2d tt = tdiag[(size_t)ttype[d] * nb + i];
  v2d r = v2d{-(tt.x * w.re - tt.y * w.im), -(tt.x * w.im + tt.y * w.re)};
  if (d == s && i == j) r.x += 1.0;
  M[(size_t)(d * nb + i) * ld + (size_t)s * nb + j] = r;
}

// Ewald triples {beta_x, beta_y, 0, Re k, Im k, s = pos_d - pos_s} of every
// ordered particle pair t = d * np + s.
__global__ void k_pair_triples(int np, const double *__restrict__ pos, double bx, double by, double kre, double kim,
                               double *__restrict__ tr) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)np * np) return;
  const int d = (int)(t / np), s = (int)(t % np);
  double *o = tr + 8 * t;
  o[0] = bx; o[1] = by; o[2] = 0; o[3] = kre; o[4] = kim;
  o[5] = pos[3 * d] - pos[3 * s];
  o[6] = pos[3 * d + 1] - pos[3 * s + 1];
  o[7] = pos[3 * d + 2] - pos[3 * s + 2];
}

// sigma_total = short + long, plus sigma0 on the (0,0) coefficient of the
// pairs with d == s (zero shift).
__global__ void k_pair_sigma_total(int np, int nsc, const v2d *__restrict__ ss, const v2d *__restrict__ sl,
                                   const v2d *__restrict__ s0, v2d *__restrict__ tot) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)np * np * nsc;
  if (e >= n) return;
  const long t = e / nsc;
  const int i = (int)(e % nsc);
  v2d v = ss[e] + sl[e];
  if (i == 0 && (t / np) == (t % np)) v += s0[0];
  tot[e] = v;
}
}  // namespace

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

// Clears `bytes` (a multiple of 4) at a 16-byte-aligned address: 16-byte
// stores over the bulk, then the 4-byte words of the tail. A kernel rather
// than hipMemsetAsync, so the serving loop's launch sequences stay plain
// kernel nodes when a caller captures them in a graph.
__global__ __launch_bounds__(256) void k_zero(uint4 *__restrict__ p, long n16, unsigned *__restrict__ tail, int ntail) {
  const long stride = (long)gridDim.x * 256;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) p[i] = uint4{0, 0, 0, 0};
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0u;
}

extern "C" int qh_zero(void *dst, size_t bytes, void *stream) {
  if (bytes == 0) return 0;
  if (((uintptr_t)dst & 15) || (bytes & 3))
    qpms_hip_die(__FILE__, __LINE__, __func__, "qh_zero: %zu bytes at %p (needs 16-byte alignment, 4-byte multiple)",
                 bytes, dst);
  const long n16 = (long)(bytes / 16);
  const int ntail = (int)((bytes % 16) / 4);
  long blocks = (n16 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  QH_LAUNCH(k_zero, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)qh_stream(stream), (uint4 *)dst, n16,
                     (unsigned *)((char *)dst + n16 * 16), ntail);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" void qh_probe_host(long n, unsigned long long seed, double *V) {
  for (long i = 0; i < n; ++i) {
    const v2d v = qh_probe_value(seed, (uint64_t)i);
    V[2 * i] = v.x;
    V[2 * i + 1] = v.y;
  }
}

extern "C" int qh_launch_probe(long n, unsigned long long seed, double *V, void *stream) {
  if (n <= 0) return 0;
  QH_LAUNCH(k_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (uint64_t)seed, (v2d *)V);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_expand_tdiag(int lMax, int nsets, const double *R, double *td, void *stream) {
  const int nb = 2 * lMax * (lMax + 2);
  QH_LAUNCH(k_expand_tdiag, dim3((nsets * nb + 255) / 256), dim3(256), 0, (hipStream_t)qh_stream(stream),
                     lMax, nsets, (const v2d *)R, (v2d *)td);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_tdiag_mul(int batch, int L, int nb, int npad, const double *td, long std_, const double *a,
                                   long sa, double *out, long so, void *stream) {
  QH_LAUNCH(k_tdiag_mul, dim3((npad + 255) / 256, batch), dim3(256), 0, (hipStream_t)qh_stream(stream), L, nb,
                     npad, (const v2d *)td, std_, (const v2d *)a, sa, (v2d *)out, so);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_pad_identity(int batch, int L, int npad, double *M, long sM, void *stream) {
  if (L >= npad) return 0;
  const long tot = (long)npad * npad;
  QH_LAUNCH(k_pad_identity, dim3((unsigned)((tot + 255) / 256), 1, batch), dim3(256), 0,
                     (hipStream_t)qh_stream(stream), L, npad, (v2d *)M, sM);
  HIPCHK(hipGetLastError());
  return 0;
}

// part: batch * QH_SUMSQ_PARTS doubles of scratch.
#define QH_SUMSQ_PARTS 512
extern "C" int qh_launch_sumsq_batched(int batch, long n, const double *x, long sx, double *part, double *out,
                                       void *stream) {
  const long p = (n + 4095) / 4096;
  const int P = (int)(p < 1 ? 1 : (p > QH_SUMSQ_PARTS ? QH_SUMSQ_PARTS : p));
  hipStream_t s = (hipStream_t)qh_stream(stream);
  QH_LAUNCH(k_sumsq_part, dim3(P, batch), dim3(256), 0, s, n, (const v2d *)x, sx, P, part);
  QH_LAUNCH(k_sumsq_fin, dim3(batch), dim3(256), 0, s, P, (const double *)part, out);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_logdet(int n, const double *LU, const int *ipiv, double *out, void *stream) {
  QH_LAUNCH(k_logdet, dim3(1), dim3(256), 0, (hipStream_t)qh_stream(stream), n, (const v2d *)LU, ipiv, out);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_e32_assemble_M(const qh_trans_tables *t, int np, int nsc, const double *sig, int bq_mode,
                                        const int *ttype, const double *tdiag, double *M, long ld, void *stream) {
  const int nb = 2 * t->nelem;
  const long npairs = (long)np * np;
  for (long t0 = 0; t0 < npairs; t0 += 65535) {
    const int cnt = (int)((npairs - t0 < 65535) ? npairs - t0 : 65535);
    QH_LAUNCH(k_e32_assemble_M, dim3((nb * nb + 255) / 256, cnt), dim3(256), 0,
                       (hipStream_t)qh_stream(stream), np, t->nelem, t->qslots, nsc, (const v2d *)sig,
                       (const v2d *)t->dA, (const v2d *)t->dB, t->dqA, t->dqB, bq_mode, ttype, (const v2d *)tdiag,
                       (v2d *)M, ld, (int)t0);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_pair_triples(int np, const double *pos, double bx, double by, double kre, double kim,
                                      double *tr, void *stream) {
  const long n = (long)np * np;
  QH_LAUNCH(k_pair_triples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream),
                     np, pos, bx, by, kre, kim, tr);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_launch_pair_sigma_total(int np, int nsc, const double *ss, const double *sl, const double *s0,
                                          double *tot, void *stream) {
  const long n = (long)np * np * nsc;
  QH_LAUNCH(k_pair_sigma_total, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)qh_stream(stream), np, nsc, (const v2d *)ss, (const v2d *)sl, (const v2d *)s0,
                     (v2d *)tot);
  HIPCHK(hipGetLastError());
  return 0;
}
