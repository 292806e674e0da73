// Dense complex linear algebra for the mode problem on gfx950: blocked
// right-looking LU with partial pivoting (zgetrf / cgetrf semantics,
// row-major, 1-based ipiv), triangular solves, and the FP32-LU + FP64
// iterative-refinement solver.
//
// Reference call sites: scatsystem.c:2214 (LAPACKE_zgetrf), 2270 (zgetrs),
// beyn.c:556-561 (zgetrf + zgetrs with l right-hand sides).
//
// LU structure (getrf_batched): outer blocks of LU_NB = 256 columns.
//   k_tslu_*      factor a tall strip of PW columns (64 in FP32, 32 in
//                 FP64) with tournament pivoting: candidate rows chosen per
//                 chunk in LDS, merged pairwise, the winners moved to the top
//                 (recorded as zgetrf-style interchanges), then L21 by a
//                 triangular solve; no workgroup waits for another;
//   k_laswp       applies the strip's row swaps to every other column;
//   k_trsm        U = L^{-1} B for a PW-row block (one thread per column,
//                 the unit-lower triangle in LDS);
//   k_gemm_sub    C -= A B: the rank-PW updates inside an outer block and
//                 the rank-256 trailing update A22 -= L21 U12 (the O(n^3)
//                 part). FP64 uses v_mfma_f64_16x16x4f64, FP32
//                 v_mfma_f32_16x16x4f32; the complex product is four real
//                 MFMA chains on de-interleaved LDS tiles.
// Every kernel takes a batch index in blockIdx.z with per-matrix strides, so
// the Beyn quadrature nodes factor in one launch sequence.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <rocblas/rocblas.h>
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
template <class T, bool UNIT_LOWER, int WM = NBW>
__global__ __launch_bounds__(256) void k_trsm(int w, int ncols, const typename C2<T>::t *L, long ldl, long sL,
                                              typename C2<T>::t *B, long ldb, long sB) {
  typedef typename C2<T>::t cT;
  constexpr int NBW = WM;
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
// ---------------------------------------------------------------------------
// Tall-strip panel: LU with tournament pivoting (CALU / TSLU, Grigori,
// Demmel & Xiang 2011). The strip (rows c0..n-1, columns c0..c0+w-1,
// w <= PW) is split into chunks of RPB rows:
//   k_tslu_local  each workgroup runs partial-pivoting elimination on its
//                 chunk in LDS only to choose w candidate rows, and writes
//                 their indices and original values;
//   k_tslu_merge  pairs of candidate sets are merged the same way, level by
//                 level, until one set of w rows remains; the order in which
//                 the last merge chose them is the pivot order;
//   k_tslu_top    one workgroup per matrix moves those rows to the top as w
//                 row interchanges (ipiv, perm), factors the w x w block
//                 without further pivoting and sets info on a zero pivot;
//   k_tslu_l21    every other strip row gets L21 = A21 U11^{-1}.
// No workgroup ever waits for another one; each step is its own launch.
// ---------------------------------------------------------------------------
constexpr int PANEL_LDS = 64 * 1024;

template <class T> struct PanelW;
template <> struct PanelW<float> { static constexpr int W = 32; };
template <> struct PanelW<double> { static constexpr int W = 32; };

template <class T>
static int panel_rows_per_block() { return PANEL_LDS / (PanelW<T>::W * (int)sizeof(typename C2<T>::t)); }

// Per matrix: two candidate buffers of Gmax sets x PW rows (index + PW values).
template <class T>
struct PanelWS {
  static size_t set_bytes() {
    constexpr int PW = PanelW<T>::W;
    return (size_t)PW * 4 + (size_t)PW * PW * sizeof(typename C2<T>::t);
  }
  static size_t bytes(int G) { return 2 * (size_t)G * ((set_bytes() + 255) / 256 * 256) + 256; }
};

// Partial-pivoting selection on X[nrows][PW] (LDS, eliminated in place):
// the first min(w, nrows) pivot rows in order into sel[]. key[r] orders ties
// (smaller key wins, as izamax's first maximum). Per column: a wave-shuffle
// arg-max, one LDS combine of the four waves, the scaling of the column and
// the rank-1 update of the unused rows (three barriers).
template <class T>
__device__ int gepp_select(typename C2<T>::t *X, int nrows, int w, const int *key, int *sel, int *used,
                           T (*wv)[4], int (*wr)[4]) {
  typedef typename C2<T>::t cT;
  constexpr int PW = PanelW<T>::W;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int r = t; r < nrows; r += 256) used[r] = 0;
  __syncthreads();
  const int ns = min(w, nrows);
  for (int j = 0; j < ns; ++j) {
    const int par = j & 1;
    T best = -1;
    int bk = 0x7fffffff, br = -1;
    for (int r = t; r < nrows; r += 256) {
      if (used[r]) continue;
      const cT v = X[r * PW + j];
      const T a = fabs(v.x) + fabs(v.y);
      if (a > best || (a == best && key[r] < bk)) { best = a; bk = key[r]; br = r; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T v2 = __shfl_xor(best, off);
      const int k2 = __shfl_xor(bk, off), r2 = __shfl_xor(br, off);
      if (v2 > best || (v2 == best && k2 < bk)) { best = v2; bk = k2; br = r2; }
    }
    if (lane == 0) { wv[par][wave] = best; wr[par][wave] = br; }
    __syncthreads();
    T pv = -1;
    int p = -1, pk = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const T v = wv[par][q];
      const int r = wr[par][q];
      const int k = (r >= 0) ? key[r] : 0x7fffffff;
      if (v > pv || (v == pv && k < pk)) { pv = v; p = r; pk = k; }
    }
    if (p < 0) {   // only NaNs left: the first unused row
      p = 0;
      while (p < nrows - 1 && used[p]) ++p;
    }
    if (pv > 0) {
      const cT d = X[p * PW + j];
      for (int r = t; r < nrows; r += 256)
        if (!used[r] && r != p) X[r * PW + j] = cdiv<T>(X[r * PW + j], d);
    }
    __syncthreads();
    if (t == 0) { sel[j] = p; used[p] = 1; }
    if (pv > 0) {
      const int nc = w - j - 1;
      for (int e = t; e < nrows * nc; e += 256) {
        const int r = e / nc, c = j + 1 + e % nc;
        if (!used[r] && r != p) {
          const cT l = X[r * PW + j], u = X[p * PW + c];
          cT &x = X[r * PW + c];
          x = cT{x.x - (l.x * u.x - l.y * u.y), x.y - (l.x * u.y + l.y * u.x)};
        }
      }
    }
    __syncthreads();
  }
  return ns;
}

template <class T>
__global__ __launch_bounds__(256) void k_tslu_local(int n, int c0, int w, int RPB, const typename C2<T>::t *A, long lda,
                                                    long sA, char *ws, long sW, long setb) {
  typedef typename C2<T>::t cT;
  constexpr int PW = PanelW<T>::W;
  extern __shared__ char smem_raw[];
  cT *X = reinterpret_cast<cT *>(smem_raw);
  __shared__ int key[512], used[512], sel[PW], wr[2][4];
  __shared__ T wv[2][4];
  const int z = blockIdx.z, g = blockIdx.x, t = threadIdx.x;
  A += (long)z * sA;
  char *set = ws + (long)z * sW + (long)g * setb;
  int *oidx = reinterpret_cast<int *>(set);
  cT *oval = reinterpret_cast<cT *>(set + PW * 4);
  const int m = n - c0, r0 = g * RPB, nr = max(0, min(RPB, m - r0));
  for (int e = t; e < nr * PW; e += 256) {
    const int r = e / PW, c = e % PW;
    X[e] = (c < w) ? A[(long)(c0 + r0 + r) * lda + c0 + c] : cT{0, 0};
  }
  for (int r = t; r < nr; r += 256) key[r] = r0 + r;
  __syncthreads();
  const int ns = gepp_select<T>(X, nr, w, key, sel, used, wv, wr);
  for (int i = t; i < PW; i += 256) oidx[i] = (i < ns) ? r0 + sel[i] : -1;
  for (int e = t; e < ns * PW; e += 256) {
    const int i = e / PW, c = e % PW;
    oval[e] = (c < w) ? A[(long)(c0 + r0 + sel[i]) * lda + c0 + c] : cT{0, 0};
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_tslu_merge(int w, int nin, const char *wsin, char *wsout, long sW,
                                                    long setb) {
  typedef typename C2<T>::t cT;
  constexpr int PW = PanelW<T>::W;
  extern __shared__ char smem_raw[];
  cT *X = reinterpret_cast<cT *>(smem_raw);   // [2 PW][PW]
  __shared__ int key[2 * PW], used[2 * PW], sel[PW], src[2 * PW], wr[2][4];
  __shared__ T wv[2][4];
  __shared__ int nrows_s;
  const int z = blockIdx.z, b = blockIdx.x, t = threadIdx.x;
  const char *in = wsin + (long)z * sW;
  char *out = wsout + (long)z * sW + (long)b * setb;
  if (t == 0) {
    int nr = 0;
    for (int q = 0; q < 2; ++q) {
      const int si = 2 * b + q;
      if (si >= nin) break;
      const int *idx = reinterpret_cast<const int *>(in + (long)si * setb);
      for (int i = 0; i < PW; ++i)
        if (idx[i] >= 0) { key[nr] = idx[i]; src[nr] = si * PW + i; ++nr; }
    }
    nrows_s = nr;
  }
  __syncthreads();
  const int nr = nrows_s;
  for (int e = t; e < nr * PW; e += 256) {
    const int r = e / PW, c = e % PW;
    const int si = src[r] / PW, i = src[r] % PW;
    X[e] = reinterpret_cast<const cT *>(in + (long)si * setb + PW * 4)[i * PW + c];
  }
  __syncthreads();
  const int ns = gepp_select<T>(X, nr, w, key, sel, used, wv, wr);
  int *oidx = reinterpret_cast<int *>(out);
  cT *oval = reinterpret_cast<cT *>(out + PW * 4);
  for (int i = t; i < PW; i += 256) oidx[i] = (i < ns) ? key[sel[i]] : -1;
  for (int e = t; e < ns * PW; e += 256) {
    const int i = e / PW, c = e % PW;
    const int si = src[sel[i]] / PW, ii = src[sel[i]] % PW;
    oval[e] = reinterpret_cast<const cT *>(in + (long)si * setb + PW * 4)[ii * PW + c];
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_tslu_top(int n, int c0, int w, typename C2<T>::t *A, long lda, long sA,
                                                  const char *wsfin, long sW, int *ipiv, long sP, int *perm,
                                                  int *info) {
  typedef typename C2<T>::t cT;
  constexpr int PW = PanelW<T>::W;
  __shared__ int tgt[PW], pos_of[2 * PW], row_of[2 * PW], ids[2 * PW], swp[PW];
  __shared__ cT U[PW * PW];
  __shared__ int nids;
  const int z = blockIdx.z, t = threadIdx.x;
  A += (long)z * sA;
  ipiv += (long)z * sP;
  if (perm) perm += (long)z * sP;
  info += z;
  const int *fidx = reinterpret_cast<const int *>(wsfin + (long)z * sW);
  if (t < w) tgt[t] = fidx[t];
  __syncthreads();
  for (int i = 0; i < w; ++i)
    if (tgt[i] < 0 || tgt[i] >= n - c0) {   // cannot happen for n - c0 >= w; never index with it
      if (t == 0) *info = -1;
      return;
    }
  // interchange sequence: strip rows involved are 0..w-1 and the targets
  if (t == 0) {
    int k = 0;
    for (int i = 0; i < w; ++i) { ids[k] = i; pos_of[k] = i; row_of[k] = i; ++k; }
    for (int i = 0; i < w; ++i) {
      int f = -1;
      for (int q = 0; q < k; ++q) if (ids[q] == tgt[i]) { f = q; break; }
      if (f < 0) { ids[k] = tgt[i]; pos_of[k] = tgt[i]; row_of[k] = tgt[i]; ++k; }
    }
    // ids[q]: original strip row; pos_of[q]: its current position
    for (int i = 0; i < w; ++i) {
      int qt = 0;
      for (int q = 0; q < k; ++q) if (ids[q] == tgt[i]) { qt = q; break; }
      const int p = pos_of[qt];
      swp[i] = p;
      if (p != i) {
        int qi = 0;   // entry whose current position is i
        for (int q = 0; q < k; ++q) if (pos_of[q] == i) { qi = q; break; }
        pos_of[qi] = p;
        pos_of[qt] = i;
      }
    }
    nids = k;
  }
  __syncthreads();
  for (int i = 0; i < w; ++i) {
    const int p = swp[i];
    if (p != i) {
      for (int c = t; c < w; c += 256) {
        const cT a = A[(long)(c0 + i) * lda + c0 + c], b = A[(long)(c0 + p) * lda + c0 + c];
        A[(long)(c0 + i) * lda + c0 + c] = b;
        A[(long)(c0 + p) * lda + c0 + c] = a;
      }
    }
    if (t == 0) {
      ipiv[c0 + i] = c0 + p + 1;
      if (perm && p != i) { const int a = perm[c0 + i]; perm[c0 + i] = perm[c0 + p]; perm[c0 + p] = a; }
    }
    __syncthreads();
  }
  // unpivoted LU of the top w x w block
  for (int e = t; e < w * w; e += 256) U[(e / w) * PW + e % w] = A[(long)(c0 + e / w) * lda + c0 + e % w];
  __syncthreads();
  for (int j = 0; j < w; ++j) {
    const cT d = U[j * PW + j];
    const bool zero = (d.x == 0 && d.y == 0);
    if (t == 0 && zero && *info == 0) *info = c0 + j + 1;
    if (!zero) {
      for (int r = j + 1 + t; r < w; r += 256) U[r * PW + j] = cdiv<T>(U[r * PW + j], d);
      __syncthreads();
      const int nc = w - j - 1;
      for (int e = t; e < nc * nc; e += 256) {
        const int r = j + 1 + e / nc, c = j + 1 + e % nc;
        const cT l = U[r * PW + j], u = U[j * PW + c];
        cT &x = U[r * PW + c];
        x = cT{x.x - (l.x * u.x - l.y * u.y), x.y - (l.x * u.y + l.y * u.x)};
      }
    }
    __syncthreads();
  }
  for (int e = t; e < w * w; e += 256) A[(long)(c0 + e / w) * lda + c0 + e % w] = U[(e / w) * PW + e % w];
  (void)nids; (void)row_of;
}

// L21 = A21 U11^{-1}: one thread per strip row below the top block.
template <class T>
__global__ __launch_bounds__(256) void k_tslu_l21(int n, int c0, int w, typename C2<T>::t *A, long lda, long sA) {
  typedef typename C2<T>::t cT;
  constexpr int PW = PanelW<T>::W;
  __shared__ cT U[PW * PW];
  A += (long)blockIdx.z * sA;
  for (int e = threadIdx.x; e < w * w; e += 256) U[(e / w) * PW + e % w] = A[(long)(c0 + e / w) * lda + c0 + e % w];
  __syncthreads();
  const int r = c0 + w + blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  cT *row = A + (long)r * lda + c0;
  cT x[PW];
#pragma unroll
  for (int i = 0; i < PW; ++i) {
    if (i < w) {
      cT v = row[i];
#pragma unroll
      for (int k = 0; k < i; ++k) {
        const cT u = U[k * PW + i];
        v = cT{v.x - (x[k].x * u.x - x[k].y * u.y), v.y - (x[k].x * u.y + x[k].y * u.x)};
      }
      const cT d = U[i * PW + i];
      x[i] = (d.x == 0 && d.y == 0) ? v : cdiv<T>(v, d);
      row[i] = x[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Trailing update C -= A B of the LU (M x N x K = n-k x n-k x 256): a plain
// complex GEMM, handed to rocBLAS (cgemm / zgemm on MFMA, FP32 or FP64
// accumulation) when it is large; k_gemm_sub otherwise, or everywhere with
// QPMS_HIP_GEMM=own. Row-major C = A B is column-major C^T = B^T A^T.
// ---------------------------------------------------------------------------
static bool use_rocblas() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("QPMS_HIP_GEMM");
    v = !(e && !strcmp(e, "own"));
  }
  return v == 1;
}

static rocblas_handle blas_for(hipStream_t s) {
  static std::mutex mu;
  static hipStream_t keys[16];
  static rocblas_handle vals[16];
  static int n = 0;
  std::lock_guard<std::mutex> g(mu);
  for (int i = 0; i < n; ++i) if (keys[i] == s) return vals[i];
  if (n == 16) qpms_hip_die(__FILE__, __LINE__, __func__, "too many streams for the GEMM handle table");
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success)
    qpms_hip_die(__FILE__, __LINE__, __func__, "rocblas_create_handle failed");
  if (rocblas_set_stream(h, s) != rocblas_status_success)
    qpms_hip_die(__FILE__, __LINE__, __func__, "rocblas_set_stream failed");
  keys[n] = s;
  vals[n] = h;
  return vals[n++];
}

template <class T>
static void trailing_gemm(int M, int N, int K, int batch, const typename C2<T>::t *A, long lda, long sA,
                          const typename C2<T>::t *B, long ldb, long sB, typename C2<T>::t *C, long ldc, long sC,
                          hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  if (!use_rocblas() || (long)M * N < 512L * 512) {
    hipLaunchKernelGGL(k_gemm_sub<T>, dim3((N + GT - 1) / GT, (M + GT - 1) / GT, batch), dim3(256), 0, s, M, N, K, A,
                       lda, sA, B, ldb, sB, C, ldc, sC);
    return;
  }
  rocblas_handle h = blas_for(s);
  rocblas_status st;
  if (sizeof(T) == 4) {
    const rocblas_float_complex al = {-1.f, 0.f}, be = {1.f, 0.f};
    st = rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &al,
                                       (const rocblas_float_complex *)B, (rocblas_int)ldb, sB,
                                       (const rocblas_float_complex *)A, (rocblas_int)lda, sA, &be,
                                       (rocblas_float_complex *)C, (rocblas_int)ldc, sC, batch);
  } else {
    const rocblas_double_complex al = {-1.0, 0.0}, be = {1.0, 0.0};
    st = rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &al,
                                       (const rocblas_double_complex *)B, (rocblas_int)ldb, sB,
                                       (const rocblas_double_complex *)A, (rocblas_int)lda, sA, &be,
                                       (rocblas_double_complex *)C, (rocblas_int)ldc, sC, batch);
  }
  if (st != rocblas_status_success)
    qpms_hip_die(__FILE__, __LINE__, __func__, "rocBLAS GEMM failed (status %d)", (int)st);
}

constexpr int LU_NB = 256;   // outer block: K of the trailing GEMM

// Workspace bytes for getrf_batched<T>(n, batch).
template <class T>
static size_t getrf_ws_bytes(int n, int batch) {
  const int RPB = panel_rows_per_block<T>();
  const int G = (n + RPB - 1) / RPB;
  return (size_t)batch * ((PanelWS<T>::bytes(G) + 255) / 256 * 256);
}

// Blocked right-looking LU with tournament-pivoted panels of `batch` n x n
// matrices (stride sA), in place. Outer blocks of LU_NB columns; inside a
// block the columns are factorised PW at a time (k_tslu_*), the block's
// remaining columns updated by TRSM + GEMM (K = PW); then U12 = L11^{-1} A12
// blocked by PW and the trailing A22 -= L21 U12 (K = LU_NB).
// ipiv (1-based, stride sP) lists the row interchanges as zgetrf's does;
// perm (optional) tracks P; info[z] gets the first zero pivot. ws:
// getrf_ws_bytes<T>(n, batch) bytes, or nullptr for a library-owned buffer
// (single-stream callers only).
template <class T>
static void getrf_batched(int n, int batch, typename C2<T>::t *A, long lda, long sA, int *ipiv, long sP,
                          int *perm, int *dinfo, hipStream_t s, char *ws = nullptr) {
  typedef typename C2<T>::t cT;
  constexpr int PW = PanelW<T>::W;
  const dim3 b256(256);
  if (n <= 0 || batch <= 0) return;
  const int RPB = panel_rows_per_block<T>();
  const int Gmax = (n + RPB - 1) / RPB;
  const long sW = (long)((PanelWS<T>::bytes(Gmax) + 255) / 256 * 256);
  const long setb = (long)((PanelWS<T>::set_bytes() + 255) / 256 * 256);
  if (!ws) {
    static char *own = nullptr;
    static size_t own_bytes = 0;
    const size_t need = (size_t)sW * batch;
    if (need > own_bytes) {
      if (own) { HIPCHK(hipStreamSynchronize(s)); qh_dfree(own); }
      own = (char *)qh_dmalloc(need);
      own_bytes = need;
    }
    ws = own;
  }
  char *bufA = ws, *bufB = ws + (long)Gmax * setb;   // two candidate buffers per matrix
  if (perm) hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256, 1, batch), b256, 0, s, n, perm, sP);
  for (int k0 = 0; k0 < n; k0 += LU_NB) {
    const int nb = min(LU_NB, n - k0);
    for (int j0 = 0; j0 < nb; j0 += PW) {
      const int w = min(PW, nb - j0), c0 = k0 + j0;
      // tournament-pivoted panel of the strip rows c0..n-1
      const int G = (n - c0 + RPB - 1) / RPB;
      hipLaunchKernelGGL(k_tslu_local<T>, dim3(G, 1, batch), b256, PANEL_LDS, s, n, c0, w, RPB, A, lda, sA, bufA, sW,
                         setb);
      char *cur = bufA, *nxt = bufB;
      for (int nsets = G; nsets > 1; nsets = (nsets + 1) / 2) {
        hipLaunchKernelGGL(k_tslu_merge<T>, dim3((nsets + 1) / 2, 1, batch), b256,
                           (size_t)2 * PW * PW * sizeof(cT), s, w, nsets, cur, nxt, sW, setb);
        char *tmp = cur; cur = nxt; nxt = tmp;
      }
      hipLaunchKernelGGL(k_tslu_top<T>, dim3(1, 1, batch), b256, 0, s, n, c0, w, A, lda, sA, cur, sW, ipiv, sP, perm,
                         dinfo);
      if (n - c0 - w > 0)
        hipLaunchKernelGGL(k_tslu_l21<T>, dim3((n - c0 - w + 255) / 256, 1, batch), b256, 0, s, n, c0, w, A, lda, sA);
      // the strip's interchanges on every other column
      if (n - w > 0)
        hipLaunchKernelGGL(k_laswp<T>, dim3((n - w + 255) / 256, 1, batch), b256, 0, s, n, c0, w, A, lda, sA, ipiv, sP);
      // rest of this outer block: U part by TRSM, then the rank-w update
      const int rest = k0 + nb - (c0 + w);
      if (rest > 0) {
        hipLaunchKernelGGL((k_trsm<T, true, PW>), dim3((rest + 255) / 256, 1, batch), b256, 0, s, w, rest,
                           A + (long)c0 * lda + c0, lda, sA, A + (long)c0 * lda + c0 + w, lda, sA);
        const int mr = n - c0 - w;
        if (mr > 0)
          hipLaunchKernelGGL(k_gemm_sub<T>, dim3((rest + GT - 1) / GT, (mr + GT - 1) / GT, batch), b256, 0, s, mr,
                             rest, w, A + (long)(c0 + w) * lda + c0, lda, sA, A + (long)c0 * lda + c0 + w, lda, sA,
                             A + (long)(c0 + w) * lda + c0 + w, lda, sA);
      }
    }
    const int nr = n - k0 - nb;
    if (nr > 0) {
      // U12 = L11^{-1} A12, PW rows at a time
      for (int j0 = 0; j0 < nb; j0 += PW) {
        const int w = min(PW, nb - j0), c0 = k0 + j0;
        hipLaunchKernelGGL((k_trsm<T, true, PW>), dim3((nr + 255) / 256, 1, batch), b256, 0, s, w, nr,
                           A + (long)c0 * lda + c0, lda, sA, A + (long)c0 * lda + k0 + nb, lda, sA);
        const int below = nb - j0 - w;
        if (below > 0)
          hipLaunchKernelGGL(k_gemm_sub<T>, dim3((nr + GT - 1) / GT, (below + GT - 1) / GT, batch), b256, 0, s, below,
                             nr, w, A + (long)(c0 + w) * lda + c0, lda, sA, A + (long)c0 * lda + k0 + nb, lda, sA,
                             A + (long)(c0 + w) * lda + k0 + nb, lda, sA);
      }
      // A22 -= L21 U12
      trailing_gemm<T>(nr, nr, nb, batch, A + (long)(k0 + nb) * lda + k0, lda, sA, A + (long)k0 * lda + k0 + nb, lda,
                       sA, A + (long)(k0 + nb) * lda + k0 + nb, lda, sA, s);
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
extern "C" size_t qh_getrf_ws_bytes(int prec, int n, int batch) {
  return prec == 32 ? getrf_ws_bytes<float>(n, batch) : getrf_ws_bytes<double>(n, batch);
}
extern "C" int qh_getrf_batched_ws_f64(int n, int batch, double *A, long sA, int *ipiv, int *perm, int *info, void *ws,
                                       void *stream) {
  getrf_batched<double>(n, batch, (v2d *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream), (char *)ws);
  return 0;
}
extern "C" int qh_getrf_batched_ws_f32(int n, int batch, float *A, long sA, int *ipiv, int *perm, int *info, void *ws,
                                       void *stream) {
  getrf_batched<float>(n, batch, (v2f *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream), (char *)ws);
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
    if (!(rel < 0.5 * prev)) {   // stalled: accept at the FP64 rounding floor
      ok = (rel <= 1e-14);
      ++it;
      break;
    }
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
