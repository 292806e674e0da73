// Beyn contour eigensolver kernels (beyn.c:345-606), FP64 on gfx950.
//
//  k_beyn_accum     A0 += dz X, A1 += dz (z - z0) X (+ the coarse sums on even
//                   nodes) for the m x l block X = M(z)^{-1} V.
//  k_jacobi_round   one round of the one-sided (Hestenes) Jacobi SVD of the
//                   m x l matrix A0, stored transposed (l rows of length m):
//                   one workgroup per column pair reduces the 2 x 2 Gram block,
//                   builds the complex rotation that diagonalises it and rotates
//                   the two columns of U and of V. The l/2 pairs of a
//                   round-robin round are independent.
//  k_ctrans_gemm    C = Ut[0:K] conj . A (K x l), i.e. V0^H A1.
//  k_small_eig      single workgroup: Hessenberg reduction, complex shifted QR
//                   to Schur form, eigenvectors by back-substitution, for the
//                   K x K reduced matrix (K <= 64), everything in LDS.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include "dev_complex.hpp"
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef double v2d __attribute__((ext_vector_type(2)));

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

namespace {

__device__ __forceinline__ cplx ld2(const v2d v) { return mk(v.x, v.y); }
__device__ __forceinline__ v2d st2(cplx c) { return v2d{c.re, c.im}; }

__global__ void k_beyn_accum(long n, const v2d *__restrict__ X, double dzr, double dzi, double wr, double wi,
                             int coarse, v2d *__restrict__ A0, v2d *__restrict__ A1, v2d *__restrict__ A0c,
                             v2d *__restrict__ A1c) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx x = ld2(X[i]);
  const cplx a = mk(dzr, dzi) * x;          // dz X
  const cplx b = mk(wr, wi) * a;            // (z - z0) dz X
  A0[i] = st2(ld2(A0[i]) + a);
  A1[i] = st2(ld2(A1[i]) + b);
  if (coarse) {
    A0c[i] = st2(ld2(A0c[i]) + 2.0 * a);
    A1c[i] = st2(ld2(A1c[i]) + 2.0 * b);
  }
}

// The same accumulation for node *cursor + slot of a node table (8 doubles
// per node: k re, k im, dz re, dz im, (z - z0) re, (z - z0) im, coarse flag,
// unused), so that a chunk's launches do not depend on its first node.
__global__ void k_beyn_accum_node(long n, const v2d *__restrict__ X, const double *__restrict__ ntab,
                                  const int *__restrict__ cursor, int slot, v2d *__restrict__ A0,
                                  v2d *__restrict__ A1, v2d *__restrict__ A0c, v2d *__restrict__ A1c) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *nd = ntab + 8 * (size_t)(cursor[0] + slot);
  const cplx x = ld2(X[i]);
  const cplx a = mk(nd[2], nd[3]) * x;
  const cplx b = mk(nd[4], nd[5]) * a;
  A0[i] = st2(ld2(A0[i]) + a);
  A1[i] = st2(ld2(A1[i]) + b);
  if (nd[6] != 0.0) {
    A0c[i] = st2(ld2(A0c[i]) + 2.0 * a);
    A1c[i] = st2(ld2(A1c[i]) + 2.0 * b);
  }
}

// *cursor += by (the next chunk's first node), one thread.
__global__ void k_cursor_advance(int *cursor, int by) {
  if (threadIdx.x == 0) cursor[0] += by;
}

// out (l x m) = transpose of in (m x l)
__global__ void k_transpose(int m, int l, const v2d *__restrict__ in, v2d *__restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * l) return;
  const int i = (int)(e / l), j = (int)(e % l);
  out[(long)j * m + i] = in[e];
}

__device__ double block_sum(double v, double *sh) {
  const int t = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  double r = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  __syncthreads();
  return r;
}

// One round of one-sided Jacobi. Ut: l rows (columns of U) of length m;
// V: l x l row-major (row j = coordinates of column j of V... stored as
// Vt: l rows, Vt[j][.] = column j of V). pairs: 2*npairs column indices.
// offmax[0]: running max of |gamma| / sqrt(alpha beta) (as bits, >= 0).
__global__ __launch_bounds__(256) void k_jacobi_round(int m, int l, v2d *__restrict__ Ut, v2d *__restrict__ Vt,
                                                      const int *__restrict__ pairs, unsigned long long *offmax) {
  __shared__ double sh[8];
  const int p = pairs[2 * blockIdx.x], q = pairs[2 * blockIdx.x + 1];
  if (p < 0 || q < 0) return;
  v2d *up = Ut + (long)p * m, *uq = Ut + (long)q * m;
  double a = 0, b = 0, gr = 0, gi = 0;  // a = |u_p|^2, b = |u_q|^2, g = u_p^H u_q
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const v2d x = up[i], y = uq[i];
    a += x.x * x.x + x.y * x.y;
    b += y.x * y.x + y.y * y.y;
    gr += x.x * y.x + x.y * y.y;
    gi += x.x * y.y - x.y * y.x;
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  gr = block_sum(gr, sh);
  gi = block_sum(gi, sh);
  const double g = hypot(gr, gi);
  if (g == 0.0 || g <= 1e-300) return;
  const double rel = g / sqrt(a * b);
  if (threadIdx.x == 0) atomicMax(offmax, (unsigned long long)__double_as_longlong(rel));
  if (rel < 1e-16) return;
  // Rotation diagonalising [[a, g], [conj g, b]]: with e = g/|g|,
  // zeta = (b - a) / (2|g|), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
  // c = 1/sqrt(1+t^2), s = c t:
  //   u_p' = c u_p - s conj(e) ... as columns: [u_p' u_q'] = [u_p u_q] [[c, s e], [-s conj(e), c]]
  const double zeta = (b - a) / (2.0 * g);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
  const cplx e = mk(gr / g, gi / g);
  const cplx se = s * e, sec = s * conj(e);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const cplx x = ld2(up[i]), y = ld2(uq[i]);
    up[i] = st2(c * x - sec * y);
    uq[i] = st2(se * x + c * y);
  }
  v2d *vp = Vt + (long)p * l, *vq = Vt + (long)q * l;
  for (int i = threadIdx.x; i < l; i += blockDim.x) {
    const cplx x = ld2(vp[i]), y = ld2(vq[i]);
    vp[i] = st2(c * x - sec * y);
    vq[i] = st2(se * x + c * y);
  }
}

// norms of the l rows of Ut
__global__ __launch_bounds__(256) void k_row_norms(int m, const v2d *__restrict__ Ut, double *__restrict__ nrm) {
  __shared__ double sh[8];
  const v2d *u = Ut + (long)blockIdx.x * m;
  double a = 0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) a += u[i].x * u[i].x + u[i].y * u[i].y;
  a = block_sum(a, sh);
  if (threadIdx.x == 0) nrm[blockIdx.x] = sqrt(a);
}

// C[k][j] = sum_i conj(Ut[k][i]) A[i][j], k < K, j < l (A row-major m x l)
__global__ void k_ctrans_gemm(int m, int l, int K, const v2d *__restrict__ Ut, const v2d *__restrict__ A,
                              v2d *__restrict__ C) {
  const int k = blockIdx.x, j = threadIdx.x;
  if (k >= K || j >= l) return;
  const v2d *u = Ut + (long)k * m;
  double re = 0, im = 0;
  for (int i = 0; i < m; ++i) {
    const v2d x = u[i], y = A[(long)i * l + j];
    re += x.x * y.x + x.y * y.y;
    im += x.x * y.y - x.y * y.x;
  }
  C[(long)k * l + j] = v2d{re, im};
}

// Y[k][i] = sum_j S[j][k] Ut[j][i]  (rows of Y = columns of V0 S), k < K
__global__ void k_v0s(int m, int K, const v2d *__restrict__ S, const v2d *__restrict__ Ut, v2d *__restrict__ Y) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)K * m) return;
  const int k = (int)(e / m), i = (int)(e % m);
  cplx acc = mk(0, 0);
  for (int j = 0; j < K; ++j) acc = acc + ld2(S[j * K + k]) * ld2(Ut[(long)j * m + i]);
  Y[e] = st2(acc);
}

// B[i][j] = sum_r T[i][r] Vt[j][r] / sigma[j]   (i, j < K): V0^H A1 W0 Sigma^{-1}
__global__ void k_bmat(int K, int l, const v2d *__restrict__ T, const v2d *__restrict__ Vt,
                       const double *__restrict__ sigma, v2d *__restrict__ B) {
  const int i = blockIdx.x, j = threadIdx.x;
  if (i >= K || j >= K) return;
  cplx acc = mk(0, 0);
  for (int r = 0; r < l; ++r) acc = acc + ld2(T[i * l + r]) * ld2(Vt[j * l + r]);
  B[i * K + j] = st2(acc / sigma[j]);
}

// ---------------------------------------------------------------------------
// Small non-Hermitian eigenproblem, one workgroup of 64 threads.
// In: B (K x K row-major). Out: lam (K), S (K x K, column k = eigenvector k,
// unit 2-norm, largest-modulus component real and positive as zgeev).
// ---------------------------------------------------------------------------
constexpr int KMAX = 64;

__global__ __launch_bounds__(64) void k_small_eig(int K, const v2d *__restrict__ Bin, v2d *__restrict__ lam_out,
                                                  v2d *__restrict__ S_out, int *__restrict__ status) {
  __shared__ cplx H[KMAX][KMAX + 1];
  __shared__ cplx Q[KMAX][KMAX + 1];
  __shared__ cplx vh[KMAX];
  __shared__ cplx sc[4];
  __shared__ int sint[4];
  const int t = threadIdx.x;
  for (int i = 0; i < K; ++i) {
    if (t < K) {
      H[i][t] = ld2(Bin[i * K + t]);
      Q[i][t] = mk(i == t ? 1.0 : 0.0, 0.0);
    }
  }
  __syncthreads();
  // 1. Householder reduction to upper Hessenberg form: H = Q^H B Q
  for (int j = 0; j < K - 2; ++j) {
    if (t == 0) {
      double nrm2 = 0;
      for (int i = j + 1; i < K; ++i) nrm2 += cabs2(H[i][j]);
      const double nrm = sqrt(nrm2);
      const cplx x0 = H[j + 1][j];
      const double ax0 = cabsd(x0);
      if (nrm == 0.0) {
        sint[0] = 0;
      } else {
        sint[0] = 1;
        const cplx ph = (ax0 > 0) ? x0 / ax0 : mk(1.0, 0.0);
        const cplx alpha = -(nrm * ph);   // new H[j+1][j]
        // v = x - alpha e1, normalised so that P = I - 2 v v^H / (v^H v)
        vh[j + 1] = x0 - alpha;
        for (int i = j + 2; i < K; ++i) vh[i] = H[i][j];
        double vn = 0;
        for (int i = j + 1; i < K; ++i) vn += cabs2(vh[i]);
        sc[0] = mk(2.0 / vn, 0.0);
      }
    }
    __syncthreads();
    if (sint[0]) {
      const double beta = sc[0].re;
      // H <- P H : columns c handled by thread c
      if (t < K) {
        cplx d = mk(0, 0);
        for (int i = j + 1; i < K; ++i) d = d + conj(vh[i]) * H[i][t];
        d = beta * d;
        for (int i = j + 1; i < K; ++i) H[i][t] = H[i][t] - vh[i] * d;
      }
      __syncthreads();
      // H <- H P, Q <- Q P : rows r handled by thread r
      if (t < K) {
        cplx d = mk(0, 0), e = mk(0, 0);
        for (int i = j + 1; i < K; ++i) {
          d = d + H[t][i] * vh[i];
          e = e + Q[t][i] * vh[i];
        }
        d = beta * d;
        e = beta * e;
        for (int i = j + 1; i < K; ++i) {
          H[t][i] = H[t][i] - d * conj(vh[i]);
          Q[t][i] = Q[t][i] - e * conj(vh[i]);
        }
      }
      __syncthreads();
      if (t > j + 1 && t < K) H[t][j] = mk(0.0, 0.0);
      __syncthreads();
    }
  }
  // 2. Complex single-shift QR iterations with Wilkinson shifts and deflation.
  int hi = K - 1, iters = 0, fail = 0;
  while (hi > 0) {
    // find lo: the start of the unreduced block ending at hi
    if (t == 0) {
      int lo = hi;
      while (lo > 0) {
        const double s = cabsd(H[lo - 1][lo - 1]) + cabsd(H[lo][lo]);
        if (cabsd(H[lo][lo - 1]) <= 1.1102230246251565e-16 * (s > 0 ? s : 1e-300)) {
          H[lo][lo - 1] = mk(0.0, 0.0);
          break;
        }
        --lo;
      }
      sint[1] = lo;
    }
    __syncthreads();
    const int lo = sint[1];
    if (lo == hi) { --hi; iters = 0; continue; }
    if (++iters > 100 * K) { fail = 1; break; }
    if (t == 0) {
      // Wilkinson shift: eigenvalue of the trailing 2x2 closer to H[hi][hi]
      const cplx a = H[hi - 1][hi - 1], b = H[hi - 1][hi], c = H[hi][hi - 1], d = H[hi][hi];
      const cplx tr2 = 0.5 * (a + d);
      const cplx det = a * d - b * c;
      const cplx disc = csqrt_(tr2 * tr2 - det);
      const cplx l1 = tr2 + disc, l2 = tr2 - disc;
      cplx mu = (cabs2(l1 - d) < cabs2(l2 - d)) ? l1 : l2;
      if (iters % 11 == 10) mu = d + mk(cabsd(c), 0.0);  // exceptional shift
      sc[1] = mu;
    }
    __syncthreads();
    const cplx mu = sc[1];
    // implicit QR sweep with Givens rotations on rows/columns lo..hi
    for (int kk = lo; kk < hi; ++kk) {
      if (t == 0) {
        cplx x, y;
        if (kk == lo) { x = H[lo][lo] - mu; y = H[lo + 1][lo]; }
        else { x = H[kk][kk - 1]; y = H[kk + 1][kk - 1]; }
        const double r = hypot(cabsd(x), cabsd(y));
        cplx cg, sg;
        if (r == 0.0) { cg = mk(1.0, 0.0); sg = mk(0.0, 0.0); }
        else { cg = x / r; sg = y / r; }
        // G = [[conj(cg), conj(sg)], [-sg, cg]] maps (x, y) -> (r, 0)
        sc[2] = cg;
        sc[3] = sg;
      }
      __syncthreads();
      const cplx cg = sc[2], sg = sc[3];
      // rows kk, kk+1 (columns t)
      if (t < K) {
        const cplx h1 = H[kk][t], h2 = H[kk + 1][t];
        H[kk][t] = conj(cg) * h1 + conj(sg) * h2;
        H[kk + 1][t] = -(sg * h1) + cg * h2;
      }
      __syncthreads();
      // columns kk, kk+1 (rows t): H <- H G^H, Q <- Q G^H
      if (t < K) {
        const cplx h1 = H[t][kk], h2 = H[t][kk + 1];
        H[t][kk] = cg * h1 + sg * h2;
        H[t][kk + 1] = -(conj(sg) * h1) + conj(cg) * h2;
        const cplx q1 = Q[t][kk], q2 = Q[t][kk + 1];
        Q[t][kk] = cg * q1 + sg * q2;
        Q[t][kk + 1] = -(conj(sg) * q1) + conj(cg) * q2;
      }
      __syncthreads();
    }
  }
  // 3. Eigenvectors of the triangular Schur factor T = H: (T - lam_k) y = 0
  //    with y_k = 1, then x = Q y, normalised (thread k).
  if (t < K) lam_out[t] = st2(H[t][t]);
  __syncthreads();
  if (t < K) {
    const int k = t;
    const cplx lk = H[k][k];
    cplx y[KMAX];
    for (int j = 0; j < K; ++j) y[j] = mk(0, 0);
    y[k] = mk(1.0, 0.0);
    double tnorm = 0;
    for (int i = 0; i < K; ++i)
      for (int j = i; j < K; ++j) tnorm = fmax(tnorm, cabsd(H[i][j]));
    const double small = fmax(tnorm * 2.220446049250313e-16, 1e-300);
    for (int j = k - 1; j >= 0; --j) {
      cplx s = mk(0, 0);
      for (int i = j + 1; i <= k; ++i) s = s + H[j][i] * y[i];
      cplx d = H[j][j] - lk;
      if (cabsd(d) < small) d = mk(small, 0.0);
      y[j] = -(s / d);
    }
    cplx x[KMAX];
    double nrm = 0, best = -1;
    int ib = 0;
    for (int i = 0; i < K; ++i) {
      cplx acc = mk(0, 0);
      for (int j = 0; j <= k; ++j) acc = acc + Q[i][j] * y[j];
      x[i] = acc;
      const double a2 = cabs2(acc);
      nrm += a2;
      if (a2 > best) { best = a2; ib = i; }
    }
    nrm = sqrt(nrm);
    const cplx ph = conj(x[ib]) / cabsd(x[ib]);
    for (int i = 0; i < K; ++i) S_out[i * K + k] = st2((x[i] * ph) / nrm);
  }
  if (t == 0) status[0] = fail;
}

}  // namespace

extern "C" int qh_beyn_accum(long n, const double *X, double dzr, double dzi, double wr, double wi, int coarse,
                             double *A0, double *A1, double *A0c, double *A1c, void *stream) {
  QH_LAUNCH(k_beyn_accum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (const v2d *)X, dzr, dzi, wr, wi, coarse, (v2d *)A0, (v2d *)A1, (v2d *)A0c, (v2d *)A1c);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_beyn_accum_node(long n, const double *X, const double *ntab, const int *cursor, int slot,
                                  double *A0, double *A1, double *A0c, double *A1c, void *stream) {
  QH_LAUNCH(k_beyn_accum_node, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
            (const v2d *)X, ntab, cursor, slot, (v2d *)A0, (v2d *)A1, (v2d *)A0c, (v2d *)A1c);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_cursor_advance(int *cursor, int by, void *stream) {
  QH_LAUNCH(k_cursor_advance, dim3(1), dim3(64), 0, (hipStream_t)qh_stream(stream), cursor, by);
  HIPCHK(hipGetLastError());
  return 0;
}

// One-sided Jacobi SVD of A (m x l, row-major, device). Outputs: Ut (l x m)
// left singular vectors as rows (sorted by decreasing sigma), sigma (l, host),
// Vt (l x l) rows = right singular vectors (column j of V as row j).
// Device scratch of qh_jacobi_svd (bytes), carved from `ws` when given.
static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }
extern "C" size_t qh_jacobi_svd_ws_bytes(int m, int l) {
  const int n = (l % 2) ? l + 1 : l;
  return a256((size_t)m * l * sizeof(v2d)) + a256((size_t)l * l * sizeof(v2d)) + 256 + a256((size_t)l * 8) +
         a256((size_t)(n - 1) * (n / 2) * 2 * sizeof(int));
}

extern "C" int qh_jacobi_svd(int m, int l, const double *A, double *Ut_out, double *sigma_host, double *Vt_out,
                             void *stream, void *ws) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  char *w = (char *)ws;
  auto take = [&](size_t b) -> void * {
    if (!ws) return qh_dmalloc(b);
    void *r = w;
    w += a256(b);
    return r;
  };
  v2d *Ut = (v2d *)take((size_t)m * l * sizeof(v2d));
  v2d *Vt = (v2d *)take((size_t)l * l * sizeof(v2d));
  unsigned long long *doff = (unsigned long long *)take(256);
  double *dn = (double *)take((size_t)l * sizeof(double));
  QH_LAUNCH(k_transpose, dim3((unsigned)(((long)m * l + 255) / 256)), dim3(256), 0, s, m, l, (const v2d *)A, Ut);
  v2d *I = (v2d *)calloc((size_t)l * l, sizeof(v2d));
  for (int i = 0; i < l; ++i) I[i * l + i] = v2d{1.0, 0.0};
  HIPCHK(hipMemcpyAsync(Vt, I, (size_t)l * l * sizeof(v2d), hipMemcpyHostToDevice, s));
  // round-robin pairing (circle method) with a dummy column for odd l; all
  // rounds of a sweep are uploaded once
  const int n = (l % 2) ? l + 1 : l;
  const int npr = n / 2;
  int *order = (int *)malloc(n * sizeof(int));
  for (int i = 0; i < n; ++i) order[i] = i;
  int *pr = (int *)malloc((size_t)(n - 1) * npr * 2 * sizeof(int));
  for (int r = 0; r < n - 1; ++r) {
    for (int k = 0; k < npr; ++k) {
      int a = order[k], b = order[n - 1 - k];
      pr[(r * npr + k) * 2] = (a >= l) ? -1 : a;
      pr[(r * npr + k) * 2 + 1] = (b >= l) ? -1 : b;
    }
    const int last = order[n - 1];
    for (int i = n - 1; i > 1; --i) order[i] = order[i - 1];
    order[1] = last;
  }
  int *dpr = (int *)take((size_t)(n - 1) * npr * 2 * sizeof(int));
  HIPCHK(hipMemcpyAsync(dpr, pr, (size_t)(n - 1) * npr * 2 * sizeof(int), hipMemcpyHostToDevice, s));
  for (int sweep = 0; sweep < 40; ++sweep) {
    HIPCHK(hipMemsetAsync(doff, 0, sizeof(unsigned long long), s));
    for (int r = 0; r < n - 1; ++r)
      QH_LAUNCH(k_jacobi_round, dim3(npr), dim3(256), 0, s, m, l, Ut, Vt, dpr + (size_t)r * npr * 2, doff);
    unsigned long long off = 0;
    HIPCHK(hipMemcpyAsync(&off, doff, sizeof off, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double offd;
    memcpy(&offd, &off, sizeof offd);
    if (offd < 1e-15) break;
  }
  if (!ws) qh_dfree(dpr);
  QH_LAUNCH(k_row_norms, dim3(l), dim3(256), 0, s, m, (const v2d *)Ut, dn);
  double *nrm = (double *)malloc(l * sizeof(double));
  HIPCHK(hipMemcpyAsync(nrm, dn, l * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // sort by decreasing sigma; normalise the U rows on the host copy
  int *idx = (int *)malloc(l * sizeof(int));
  for (int i = 0; i < l; ++i) idx[i] = i;
  for (int i = 1; i < l; ++i) {
    int v = idx[i], j = i;
    while (j > 0 && nrm[idx[j - 1]] < nrm[v]) { idx[j] = idx[j - 1]; --j; }
    idx[j] = v;
  }
  v2d *hU = (v2d *)malloc((size_t)m * l * sizeof(v2d));
  v2d *hV = (v2d *)malloc((size_t)l * l * sizeof(v2d));
  HIPCHK(hipMemcpyAsync(hU, Ut, (size_t)m * l * sizeof(v2d), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(hV, Vt, (size_t)l * l * sizeof(v2d), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  v2d *oU = (v2d *)malloc((size_t)m * l * sizeof(v2d));
  v2d *oV = (v2d *)malloc((size_t)l * l * sizeof(v2d));
  for (int r = 0; r < l; ++r) {
    const int j = idx[r];
    sigma_host[r] = nrm[j];
    const double inv = nrm[j] > 0 ? 1.0 / nrm[j] : 0.0;
    for (int i = 0; i < m; ++i) oU[(long)r * m + i] = hU[(long)j * m + i] * inv;
    for (int i = 0; i < l; ++i) oV[(long)r * l + i] = hV[(long)j * l + i];
  }
  HIPCHK(hipMemcpyAsync(Ut_out, oU, (size_t)m * l * sizeof(v2d), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(Vt_out, oV, (size_t)l * l * sizeof(v2d), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  free(I); free(order); free(pr); free(nrm); free(idx); free(hU); free(hV); free(oU); free(oV);
  if (!ws) { qh_dfree(Ut); qh_dfree(Vt); qh_dfree(doff); qh_dfree(dn); }
  return 0;
}

extern "C" int qh_ctrans_gemm(int m, int l, int K, const double *Ut, const double *A, double *C, void *stream) {
  if (K <= 0) return 0;
  QH_LAUNCH(k_ctrans_gemm, dim3(K), dim3(l), 0, (hipStream_t)qh_stream(stream), m, l, K, (const v2d *)Ut,
                     (const v2d *)A, (v2d *)C);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_v0s(int m, int K, const double *S, const double *Ut, double *Y, void *stream) {
  if (K <= 0) return 0;
  QH_LAUNCH(k_v0s, dim3((unsigned)(((long)K * m + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream),
                     m, K, (const v2d *)S, (const v2d *)Ut, (v2d *)Y);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_bmat(int K, int l, const double *T, const double *Vt, const double *sigma, double *B,
                       void *stream) {
  QH_LAUNCH(k_bmat, dim3(K), dim3(64), 0, (hipStream_t)qh_stream(stream), K, l, (const v2d *)T,
                     (const v2d *)Vt, sigma, (v2d *)B);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int qh_small_eig(int K, const double *B, double *lam, double *S, void *stream, void *ws) {
  if (K <= 0) return 0;
  if (K > KMAX) qpms_hip_die(__FILE__, __LINE__, __func__, "reduced eigenproblem of order %d exceeds %d", K, KMAX);
  hipStream_t s = (hipStream_t)qh_stream(stream);
  int *st = ws ? (int *)ws : (int *)qh_dmalloc(sizeof(int));
  QH_LAUNCH(k_small_eig, dim3(1), dim3(64), 0, s, K, (const v2d *)B, (v2d *)lam, (v2d *)S, st);
  HIPCHK(hipGetLastError());
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, st, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (!ws) qh_dfree(st);
  return h;
}
