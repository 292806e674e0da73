// Dense complex linear algebra for the mode problem on gfx950: blocked
// right-looking LU with partial pivoting (zgetrf / cgetrf semantics,
// row-major, 1-based ipiv), triangular solves, and the FP32-LU + FP64
// iterative-refinement solver.
//
// Reference call sites: scatsystem.c:2214 (LAPACKE_zgetrf), 2270 (zgetrs),
// beyn.c:556-561 (zgetrf + zgetrs with l right-hand sides).
//
// Structure per panel of NBW = 32 columns:
//   k_panel     one workgroup factors the tall panel column by column
//               (pivot = first max |Re|+|Im|, as izamax), keeps a row
//               permutation array up to date;
//   k_laswp     applies the panel's row swaps to the columns left and right
//               of the panel (one thread per column, coalesced rows);
//   k_trsm_l    U12 = L11^{-1} A12 (one thread per column, L11 in LDS);
//   k_gemm_sub  A22 -= L21 U12 — the O(n^3) part. FP64 uses
//               v_mfma_f64_16x16x4f64, FP32 uses v_mfma_f32_16x16x4f32; the
//               complex product is split into four real MFMA chains on
//               de-interleaved LDS tiles.
// Every kernel takes a batch index in blockIdx.z with per-matrix strides, so
// the Beyn quadrature nodes factor in one launch sequence.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>
#include "qpms_internal.h"

typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

namespace {

constexpr int NBW = 32;  // panel width

template <class T> struct C2;
template <> struct C2<double> { typedef v2d t; };
template <> struct C2<float> { typedef v2f t; };

template <class T> __device__ __forceinline__ typename C2<T>::t cmul(typename C2<T>::t a, typename C2<T>::t b) {
  return typename C2<T>::t{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T> __device__ __forceinline__ typename C2<T>::t cdiv(typename C2<T>::t a, typename C2<T>::t b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    T r = b.y / b.x, d = b.x + b.y * r;
    return typename C2<T>::t{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  } else {
    T r = b.x / b.y, d = b.x * r + b.y;
    return typename C2<T>::t{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
  }
}

// ---------------------------------------------------------------------------
// Panel factorisation: rows k0..n-1, columns k0..k0+w-1 of matrix z.
// perm (if non-null) holds the current row permutation (perm[i] = original
// row now at position i) and is updated with the panel's swaps.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_panel(int n, int k0, int w, typename C2<T>::t *A, long lda, long sA,
                                               int *ipiv, long sP, int *perm, int *info) {
  typedef typename C2<T>::t cT;
  A += (long)blockIdx.z * sA;
  ipiv += (long)blockIdx.z * sP;
  if (perm) perm += (long)blockIdx.z * sP;
  info += blockIdx.z;
  __shared__ T sval[256];
  __shared__ int sidx[256];
  __shared__ cT prow[NBW];
  const int t = threadIdx.x;
  for (int j = 0; j < w; ++j) {
    const int col = k0 + j;
    // pivot search: first index of the maximum |Re| + |Im|
    T best = -1;
    int bi = n;
    for (int i = col + t; i < n; i += 256) {
      cT a = A[(long)i * lda + col];
      T v = fabs(a.x) + fabs(a.y);
      if (v > best) { best = v; bi = i; }
    }
    sval[t] = best;
    sidx[t] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        T v2 = sval[t + s];
        int i2 = sidx[t + s];
        if (v2 > sval[t] || (v2 == sval[t] && i2 < sidx[t])) { sval[t] = v2; sidx[t] = i2; }
      }
      __syncthreads();
    }
    const int p = sidx[0];
    const T pmax = sval[0];
    __syncthreads();
    if (t == 0) {
      ipiv[col] = p + 1;
      if (pmax == 0 && *info == 0) *info = col + 1;
      if (perm && p != col) { int a = perm[col]; perm[col] = perm[p]; perm[p] = a; }
    }
    if (p != col && pmax != 0) {
      for (int c = t; c < w; c += 256) {
        cT a = A[(long)col * lda + k0 + c], b = A[(long)p * lda + k0 + c];
        A[(long)col * lda + k0 + c] = b;
        A[(long)p * lda + k0 + c] = a;
      }
    }
    __syncthreads();
    if (t < w) prow[t] = A[(long)col * lda + k0 + t];
    __syncthreads();
    if (pmax != 0) {
      const cT d = prow[j];
      for (int i = col + 1 + t; i < n; i += 256) {
        cT *row = A + (long)i * lda + k0;
        const cT l = cdiv<T>(row[j], d);
        row[j] = l;
        for (int c = j + 1; c < w; ++c) {
          const cT u = prow[c];
          row[c] = cT{row[c].x - (l.x * u.x - l.y * u.y), row[c].y - (l.x * u.y + l.y * u.x)};
        }
      }
    }
    __syncthreads();
  }
}

// Row swaps of rows k0..k0+w-1 (pivots ipiv[k0..]) on columns outside the panel.
template <class T>
__global__ __launch_bounds__(256) void k_laswp(int n, int k0, int w, typename C2<T>::t *A, long lda, long sA,
                                               const int *ipiv, long sP) {
  typedef typename C2<T>::t cT;
  A += (long)blockIdx.z * sA;
  ipiv += (long)blockIdx.z * sP;
  const int c0 = blockIdx.x * 256 + threadIdx.x;
  const int ncols = n - w;  // columns outside the panel
  if (c0 >= ncols) return;
  const int c = (c0 < k0) ? c0 : c0 + w;
  for (int j = 0; j < w; ++j) {
    const int r = k0 + j, p = ipiv[r] - 1;
    if (p != r) {
      cT a = A[(long)r * lda + c], b = A[(long)p * lda + c];
      A[(long)r * lda + c] = b;
      A[(long)p * lda + c] = a;
    }
  }
}

// A12 <- L11^{-1} A12 with L11 unit lower triangular (w x w at (k0, k0)).
// Generic form: rows r0..r0+w of B (ldb) against the triangle at L (ldl).
template <class T, bool UNIT_LOWER>
__global__ __launch_bounds__(256) void k_trsm(int w, int ncols, const typename C2<T>::t *L, long ldl, long sL,
                                              typename C2<T>::t *B, long ldb, long sB) {
  typedef typename C2<T>::t cT;
  L += (long)blockIdx.z * sL;
  B += (long)blockIdx.z * sB;
  __shared__ cT sl[NBW * NBW];
  for (int i = threadIdx.x; i < w * w; i += 256) sl[i] = L[(long)(i / w) * ldl + (i % w)];
  __syncthreads();
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncols) return;
  cT x[NBW];
  if (UNIT_LOWER) {
#pragma unroll
    for (int r = 0; r < NBW; ++r) {
      if (r < w) {
        cT v = B[(long)r * ldb + c];
#pragma unroll
        for (int s = 0; s < r; ++s) {
          const cT l = sl[r * w + s];
          v = cT{v.x - (l.x * x[s].x - l.y * x[s].y), v.y - (l.x * x[s].y + l.y * x[s].x)};
        }
        x[r] = v;
        B[(long)r * ldb + c] = v;
      }
    }
  } else {  // upper, non-unit: backward substitution
#pragma unroll
    for (int rr = NBW - 1; rr >= 0; --rr) {
      if (rr < w) {
        const int r = rr;
        cT v = B[(long)r * ldb + c];
#pragma unroll
        for (int s = r + 1; s < NBW; ++s) {
          if (s < w) {
            const cT u = sl[r * w + s];
            v = cT{v.x - (u.x * x[s].x - u.y * x[s].y), v.y - (u.x * x[s].y + u.y * x[s].x)};
          }
        }
        v = cdiv<T>(v, sl[r * w + r]);
        x[r] = v;
        B[(long)r * ldb + c] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// C[M x N] -= A[M x K] * B[K x N] (complex, row-major), MFMA.
// Block: 256 threads = 4 waves, C tile 64 x 64, each wave a 32 x 32 quadrant
// made of 2 x 2 MFMA 16x16 tiles, K step 16 (= 4 MFMA k-steps of 4).
// LDS tiles are de-interleaved: Ar, Ai [64][16+1], Br, Bi [16][64+1].
// Complex: Cr -= Ar Br - Ai Bi ; Ci -= Ar Bi + Ai Br  → accumulate
//   Pr = Ar Br + (-Ai) Bi and Pi = Ar Bi + Ai Br, then C -= P.
// MFMA 16x16x4 lane layout: a = A[l%16][l/16], b = B[l/16][l%16],
// acc[i] = D[MF<T>::row(l/16, i)][l%16].
// ---------------------------------------------------------------------------
template <class T> struct MF;
// Output row of accumulator element i held by lane group lk (= lane / 16):
// the f64 16x16x4 instruction interleaves the four lane groups row by row,
// the f32 one gives each lane group four consecutive rows.
template <> struct MF<double> {
  typedef v4d acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int i) { return lk + 4 * i; }
};
template <> struct MF<float> {
  typedef v4f acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int i) { return 4 * lk + i; }
};

constexpr int GT = 64, GK = 16;

template <class T>
__global__ __launch_bounds__(256) void k_gemm_sub(int M, int N, int K, const typename C2<T>::t *A, long lda, long sA,
                                                  const typename C2<T>::t *B, long ldb, long sB,
                                                  typename C2<T>::t *C, long ldc, long sC) {
  typedef typename C2<T>::t cT;
  typedef typename MF<T>::acc acc_t;
  A += (long)blockIdx.z * sA;
  B += (long)blockIdx.z * sB;
  C += (long)blockIdx.z * sC;
  __shared__ T Ar[GT][GK + 1], Ai[GT][GK + 1], Br[GK][GT + 1], Bi[GK][GT + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;  // wave quadrant
  acc_t pr[2][2], pi[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) { pr[a][b] = acc_t{0, 0, 0, 0}; pi[a][b] = acc_t{0, 0, 0, 0}; }
  const int lr = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += GK) {
    // load A tile 64 x 16 (1024 complex, 4 per thread) and B tile 16 x 64
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int ar = idx >> 4, ak = idx & 15;
      cT va = cT{0, 0};
      if (m0 + ar < M && k0 + ak < K) va = A[(long)(m0 + ar) * lda + k0 + ak];
      Ar[ar][ak] = va.x;
      Ai[ar][ak] = va.y;
      const int bk = idx >> 6, bn = idx & 63;
      cT vb = cT{0, 0};
      if (k0 + bk < K && n0 + bn < N) vb = B[(long)(k0 + bk) * ldb + n0 + bn];
      Br[bk][bn] = vb.x;
      Bi[bk][bn] = vb.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const T ar = Ar[wm + a * 16 + lr][kk + lk], ai = Ai[wm + a * 16 + lr][kk + lk];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const T br = Br[kk + lk][wn + b * 16 + lr], bi = Bi[kk + lk][wn + b * 16 + lr];
          pr[a][b] = MF<T>::mma(ar, br, pr[a][b]);
          pr[a][b] = MF<T>::mma(-ai, bi, pr[a][b]);
          pi[a][b] = MF<T>::mma(ar, bi, pi[a][b]);
          pi[a][b] = MF<T>::mma(ai, br, pi[a][b]);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = m0 + wm + a * 16 + MF<T>::row(lk, i), col = n0 + wn + b * 16 + lr;
        if (row < M && col < N) {
          cT c = C[(long)row * ldc + col];
          c.x -= pr[a][b][i];
          c.y -= pi[a][b][i];
          C[(long)row * ldc + col] = c;
        }
      }
}

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
__global__ void k_iota(int n, int *p, long sP) {
  p += (long)blockIdx.z * sP;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = i;
}

// Y[i][:] = X[perm[i]][:]
template <class T>
__global__ void k_gather_rows(int n, int ncols, const int *perm, long sP, const typename C2<T>::t *X, long ldx,
                              long sX, typename C2<T>::t *Y, long ldy, long sY) {
  perm += (long)blockIdx.z * sP;
  X += (long)blockIdx.z * sX;
  Y += (long)blockIdx.z * sY;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)n * ncols) return;
  const int i = (int)(e / ncols), c = (int)(e % ncols);
  Y[(long)i * ldy + c] = X[(long)perm[i] * ldx + c];
}

// perm from LAPACK-style pivots (single thread per matrix; used only when a
// caller hands in pivots without the permutation).
__global__ void k_perm_from_ipiv(int n, const int *ipiv, int *perm) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int i = 0; i < n; ++i) {
    const int p = ipiv[i] - 1;
    if (p != i) { int a = perm[i]; perm[i] = perm[p]; perm[p] = a; }
  }
}

__global__ void k_d2f(long n, const v2d *x, v2f *y) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = v2f{(float)x[i].x, (float)x[i].y};
}
__global__ void k_f2d_add(long n, const v2f *d, v2d *x) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = v2d{x[i].x + (double)d[i].x, x[i].y + (double)d[i].y};
}
__global__ void k_copy2d(long n, const v2d *x, v2d *y) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = x[i];
}
// Sum of |z|^2 into out[blockIdx.x] (partial sums, finished on the host).
__global__ void k_sumsq(long n, const v2d *x, double *out) {
  __shared__ double s[256];
  double a = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    a += x[i].x * x[i].x + x[i].y * x[i].y;
  s[threadIdx.x] = a;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// M = I - T W for one particle's block row: T (nr x nr), W rows (nr x L), the
// identity lands at columns col0..col0+nr-1.
__global__ void k_block_tw(int nr, int L, const v2d *T, const v2d *W, v2d *M, int col0) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= L) return;
  double re = (j - col0 == i) ? 1.0 : 0.0, im = 0.0;
  for (int k = 0; k < nr; ++k) {
    const v2d t = T[(long)i * nr + k], w = W[(long)k * L + j];
    re -= t.x * w.x - t.y * w.y;
    im -= t.x * w.y + t.y * w.x;
  }
  M[(long)i * L + j] = v2d{re, im};
}

}  // namespace

// ===========================================================================
// Host drivers
// ===========================================================================
template <class T>
static void getrf_batched(int n, int batch, typename C2<T>::t *A, long lda, long sA, int *ipiv, long sP,
                          int *perm, int *dinfo, hipStream_t s) {
  const dim3 b256(256);
  if (perm) hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256, 1, batch), b256, 0, s, n, perm, sP);
  for (int k0 = 0; k0 < n; k0 += NBW) {
    const int w = (n - k0 < NBW) ? n - k0 : NBW;
    hipLaunchKernelGGL(k_panel<T>, dim3(1, 1, batch), b256, 0, s, n, k0, w, A, lda, sA, ipiv, sP, perm, dinfo);
    const int nout = n - w;
    if (nout > 0)
      hipLaunchKernelGGL(k_laswp<T>, dim3((nout + 255) / 256, 1, batch), b256, 0, s, n, k0, w, A, lda, sA, ipiv, sP);
    const int nr = n - k0 - w;
    if (nr > 0) {
      typename C2<T>::t *A11 = A + (long)k0 * lda + k0;
      typename C2<T>::t *A12 = A + (long)k0 * lda + k0 + w;
      typename C2<T>::t *A21 = A + (long)(k0 + w) * lda + k0;
      typename C2<T>::t *A22 = A + (long)(k0 + w) * lda + k0 + w;
      hipLaunchKernelGGL((k_trsm<T, true>), dim3((nr + 255) / 256, 1, batch), b256, 0, s, w, nr, A11, lda, sA, A12,
                         lda, sA);
      hipLaunchKernelGGL(k_gemm_sub<T>, dim3((nr + GT - 1) / GT, (nr + GT - 1) / GT, batch), b256, 0, s, nr, nr, w,
                         A21, lda, sA, A12, lda, sA, A22, lda, sA);
    }
  }
  HIPCHK(hipGetLastError());
}

// B (n x nrhs, ldb) <- A^{-1} B given LU and the row permutation.
template <class T>
static void getrs_batched(int n, int nrhs, int batch, const typename C2<T>::t *LU, long lda, long sA,
                          const int *perm, long sP, typename C2<T>::t *B, long ldb, long sB,
                          typename C2<T>::t *work, long sW, hipStream_t s) {
  typedef typename C2<T>::t cT;
  const dim3 b256(256);
  const long tot = (long)n * nrhs;
  // work <- P B, then forward / backward substitution in place on work, copy back.
  hipLaunchKernelGGL(k_gather_rows<T>, dim3((unsigned)((tot + 255) / 256), 1, batch), b256, 0, s, n, nrhs, perm, sP,
                     (const cT *)B, ldb, sB, work, (long)nrhs, sW);
  for (int r0 = 0; r0 < n; r0 += NBW) {
    const int w = (n - r0 < NBW) ? n - r0 : NBW;
    hipLaunchKernelGGL((k_trsm<T, true>), dim3((nrhs + 255) / 256, 1, batch), b256, 0, s, w, nrhs,
                       LU + (long)r0 * lda + r0, lda, sA, work + (long)r0 * nrhs, (long)nrhs, sW);
    const int rem = n - r0 - w;
    if (rem > 0)
      hipLaunchKernelGGL(k_gemm_sub<T>, dim3((nrhs + GT - 1) / GT, (rem + GT - 1) / GT, batch), b256, 0, s, rem, nrhs, w,
                         LU + (long)(r0 + w) * lda + r0, lda, sA, (const cT *)(work + (long)r0 * nrhs), (long)nrhs, sW,
                         work + (long)(r0 + w) * nrhs, (long)nrhs, sW);
  }
  const int nblk = (n + NBW - 1) / NBW;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int r0 = bi * NBW;
    const int w = (n - r0 < NBW) ? n - r0 : NBW;
    hipLaunchKernelGGL((k_trsm<T, false>), dim3((nrhs + 255) / 256, 1, batch), b256, 0, s, w, nrhs,
                       LU + (long)r0 * lda + r0, lda, sA, work + (long)r0 * nrhs, (long)nrhs, sW);
    if (r0 > 0)
      hipLaunchKernelGGL(k_gemm_sub<T>, dim3((nrhs + GT - 1) / GT, (r0 + GT - 1) / GT, batch), b256, 0, s, r0, nrhs, w,
                         LU + r0, lda, sA, (const cT *)(work + (long)r0 * nrhs), (long)nrhs, sW, work, (long)nrhs, sW);
  }
  // copy back (B may have a different leading dimension)
  for (int z = 0; z < batch; ++z)
    HIPCHK(hipMemcpy2DAsync(B + (long)z * sB, ldb * sizeof(cT), work + (long)z * sW, (long)nrhs * sizeof(cT),
                            (long)nrhs * sizeof(cT), n, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipGetLastError());
}

extern "C" int qpms_hip_zgetrf_dev(size_t n, _Complex double *A_dev, int *ipiv_dev, void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  int *dinfo = (int *)qh_dmalloc(sizeof(int));
  HIPCHK(hipMemsetAsync(dinfo, 0, sizeof(int), s));
  getrf_batched<double>((int)n, 1, (v2d *)A_dev, (long)n, 0, ipiv_dev, 0, nullptr, dinfo, s);
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  qh_dfree(dinfo);
  return info;
}

extern "C" int qpms_hip_zgetrs_dev(size_t n, size_t nrhs, const _Complex double *LU_dev, const int *ipiv_dev,
                                   _Complex double *B_dev, void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  int *perm = (int *)qh_dmalloc(n * sizeof(int));
  v2d *work = (v2d *)qh_dmalloc(n * nrhs * sizeof(v2d));
  hipLaunchKernelGGL(k_perm_from_ipiv, dim3(1), dim3(64), 0, s, (int)n, ipiv_dev, perm);
  getrs_batched<double>((int)n, (int)nrhs, 1, (const v2d *)LU_dev, (long)n, 0, perm, 0, (v2d *)B_dev, (long)nrhs, 0,
                        work, 0, s);
  HIPCHK(hipStreamSynchronize(s));
  qh_dfree(perm);
  qh_dfree(work);
  return 0;
}

extern "C" int qh_launch_block_tw(size_t nr, size_t L, const double *T, const double *W, double *M, size_t col0,
                                  void *stream) {
  hipLaunchKernelGGL(k_block_tw, dim3((unsigned)((L + 255) / 256), (unsigned)nr), dim3(256), 0,
                     (hipStream_t)qh_stream(stream), (int)nr, (int)L, (const v2d *)T, (const v2d *)W, (v2d *)M,
                     (int)col0);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Batched LU / solve entry points used by Beyn and the serving loop.
// ---------------------------------------------------------------------------
extern "C" int qh_getrf_batched_f64(int n, int batch, double *A, long sA, int *ipiv, int *perm, int *info,
                                    void *stream) {
  getrf_batched<double>(n, batch, (v2d *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream));
  return 0;
}
extern "C" int qh_getrs_batched_f64(int n, int nrhs, int batch, const double *LU, long sA, const int *perm,
                                    double *B, long sB, double *work, void *stream) {
  getrs_batched<double>(n, nrhs, batch, (const v2d *)LU, n, sA, perm, n, (v2d *)B, nrhs, sB, (v2d *)work,
                        (long)n * nrhs, (hipStream_t)qh_stream(stream));
  return 0;
}
extern "C" int qh_getrf_batched_f32(int n, int batch, float *A, long sA, int *ipiv, int *perm, int *info,
                                    void *stream) {
  getrf_batched<float>(n, batch, (v2f *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream));
  return 0;
}
extern "C" int qh_getrs_batched_f32(int n, int nrhs, int batch, const float *LU, long sA, const int *perm,
                                    float *B, long sB, float *work, void *stream) {
  getrs_batched<float>(n, nrhs, batch, (const v2f *)LU, n, sA, perm, n, (v2f *)B, nrhs, sB, (v2f *)work,
                       (long)n * nrhs, (hipStream_t)qh_stream(stream));
  return 0;
}
// C -= A B (FP64 complex, batched, row-major, packed leading dimensions)
extern "C" int qh_gemm_sub_f64(int M, int N, int K, int batch, const double *A, long lda, long sA, const double *B,
                               long ldb, long sB, double *C, long ldc, long sC, void *stream) {
  hipLaunchKernelGGL(k_gemm_sub<double>, dim3((N + GT - 1) / GT, (M + GT - 1) / GT, batch), dim3(256), 0,
                     (hipStream_t)qh_stream(stream), M, N, K, (const v2d *)A, lda, sA, (const v2d *)B, ldb, sB,
                     (v2d *)C, ldc, sC);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int qh_gemm_sub_f32(int M, int N, int K, int batch, const float *A, long lda, long sA, const float *B,
                               long ldb, long sB, float *C, long ldc, long sC, void *stream) {
  hipLaunchKernelGGL(k_gemm_sub<float>, dim3((N + GT - 1) / GT, (M + GT - 1) / GT, batch), dim3(256), 0,
                     (hipStream_t)qh_stream(stream), M, N, K, (const v2f *)A, lda, sA, (const v2f *)B, ldb, sB,
                     (v2f *)C, ldc, sC);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int qh_d2f(long n, const double *x, float *y, void *stream) {
  hipLaunchKernelGGL(k_d2f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (const v2d *)x, (v2f *)y);
  return 0;
}
extern "C" int qh_f2d_add(long n, const float *d, double *x, void *stream) {
  hipLaunchKernelGGL(k_f2d_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (const v2f *)d, (v2d *)x);
  return 0;
}
extern "C" int qh_copy_c64(long n, const double *x, double *y, void *stream) {
  hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (const v2d *)x, (v2d *)y);
  return 0;
}
// Frobenius norm of n complex values (synchronises).
extern "C" double qh_norm2_c64(long n, const double *x, void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const int nb = 512;
  double *part = (double *)qh_dmalloc(nb * sizeof(double));
  hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, s, n, (const v2d *)x, part);
  double h[nb];
  HIPCHK(hipMemcpyAsync(h, part, sizeof h, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  qh_dfree(part);
  double t = 0;
  for (int i = 0; i < nb; ++i) t += h[i];
  return sqrt(t);
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 LU on MFMA + FP64 iterative refinement.
// ---------------------------------------------------------------------------
extern "C" int qpms_hip_solve_mixed_dev(size_t n_, size_t nrhs_, const _Complex double *A_dev,
                                        _Complex double *B_dev, _Complex double *X_dev, double tol, int maxit,
                                        void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const int n = (int)n_, nrhs = (int)nrhs_;
  const long nn = (long)n * n, nb = (long)n * nrhs;
  float *A32 = (float *)qh_dmalloc(nn * sizeof(v2f));
  int *ipiv = (int *)qh_dmalloc(n * sizeof(int));
  int *perm = (int *)qh_dmalloc(n * sizeof(int));
  int *dinfo = (int *)qh_dmalloc(sizeof(int));
  double *R = (double *)qh_dmalloc(nb * sizeof(v2d));
  float *R32 = (float *)qh_dmalloc(nb * sizeof(v2f));
  float *W32 = (float *)qh_dmalloc(nb * sizeof(v2f));
  HIPCHK(hipMemsetAsync(dinfo, 0, sizeof(int), s));
  qh_d2f(nn, (const double *)A_dev, A32, s);
  getrf_batched<float>(n, 1, (v2f *)A32, n, 0, ipiv, 0, perm, dinfo, s);
  HIPCHK(hipMemsetAsync(X_dev, 0, nb * sizeof(v2d), s));
  const double normA = qh_norm2_c64(nn, (const double *)A_dev, s);
  const double normB = qh_norm2_c64(nb, (const double *)B_dev, s);
  int it = 0, ok = 0;
  double prev = INFINITY;
  qh_copy_c64(nb, (const double *)B_dev, R, s);
  for (it = 0; it < maxit; ++it) {
    qh_d2f(nb, R, R32, s);
    getrs_batched<float>(n, nrhs, 1, (const v2f *)A32, n, 0, perm, 0, (v2f *)R32, nrhs, 0, (v2f *)W32, 0, s);
    qh_f2d_add(nb, R32, (double *)X_dev, s);
    // R = B - A X
    qh_copy_c64(nb, (const double *)B_dev, R, s);
    qh_gemm_sub_f64(n, nrhs, n, 1, (const double *)A_dev, n, 0, (const double *)X_dev, nrhs, 0, R, nrhs, 0, s);
    const double nr = qh_norm2_c64(nb, R, s);
    const double nx = qh_norm2_c64(nb, (const double *)X_dev, s);
    const double rel = nr / (normA * nx + normB * 1e-300);
    if (rel <= tol) { ok = 1; ++it; break; }
    if (!(rel < 0.5 * prev)) break;  // stalled
    prev = rel;
  }
  qh_dfree(A32); qh_dfree(ipiv); qh_dfree(perm); qh_dfree(dinfo); qh_dfree(R); qh_dfree(R32); qh_dfree(W32);
  if (ok) return it;
  // Refinement stalled: FP64 LU fallback on the GPU.
  double *A64 = (double *)qh_dmalloc(nn * sizeof(v2d));
  int *ip = (int *)qh_dmalloc(n * sizeof(int));
  int *pm = (int *)qh_dmalloc(n * sizeof(int));
  int *inf = (int *)qh_dmalloc(sizeof(int));
  double *work = (double *)qh_dmalloc(nb * sizeof(v2d));
  HIPCHK(hipMemsetAsync(inf, 0, sizeof(int), s));
  qh_copy_c64(nn, (const double *)A_dev, A64, s);
  qh_copy_c64(nb, (const double *)B_dev, (double *)X_dev, s);
  getrf_batched<double>(n, 1, (v2d *)A64, n, 0, ip, 0, pm, inf, s);
  getrs_batched<double>(n, nrhs, 1, (const v2d *)A64, n, 0, pm, 0, (v2d *)X_dev, nrhs, 0, (v2d *)work, 0, s);
  HIPCHK(hipStreamSynchronize(s));
  qh_dfree(A64); qh_dfree(ip); qh_dfree(pm); qh_dfree(inf); qh_dfree(work);
  return -1;
}
