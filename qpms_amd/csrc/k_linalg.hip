// Dense complex linear algebra for the mode problem on gfx950: blocked
// right-looking LU with partial pivoting (zgetrf / cgetrf semantics,
// row-major, 1-based ipiv), triangular solves, and the FP32-LU + FP64
// iterative-refinement solver.
//
// Reference call sites: scatsystem.c:2214 (LAPACKE_zgetrf), 2270 (zgetrs),
// beyn.c:556-561 (zgetrf + zgetrs with l right-hand sides).
//
// LU structure (getrf_batched): outer blocks of LU_NB = 256 columns.
//   k_ts_*        factor a tall strip of PW = 32 columns with tournament
//                 pivoting: candidate rows chosen per chunk of 256 rows
//                 staged in LDS, merged a chunk's worth of candidate sets at
//                 a time, the winners moved to the top (recorded as
//                 zgetrf-style interchanges) in every column by one gather;
//                 the last level also inverts the strip's L11 and U11;
//                 no workgroup waits for another;
//   k_gemm_sub    C -= A B, or C = A B: L21 = A21 U11^{-1} and
//                 U12 = L11^{-1} A12 (products with the inverses), the
//                 rank-PW updates inside an outer block and the rank-256
//                 trailing update A22 -= L21 U12 (the O(n^3) part, rocBLAS or
//                 k_cgemm3m by size class). FP64 uses v_mfma_f64_16x16x4f64, FP32
//                 v_mfma_f32_16x16x4f32; the complex product is four real
//                 MFMA chains on de-interleaved LDS tiles.
// Every kernel takes a batch index in blockIdx.z with per-matrix strides, so
// the Beyn quadrature nodes factor in one launch sequence.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <rocblas/rocblas.h>
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

namespace {

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
// C[M x N] -= A[M x K] * B[K x N] (complex, row-major), MFMA; with ASSIGN,
// C = A B. Every workgroup reads its whole K extent before it writes its C
// tile, so C may be A (N <= 64: one column tile) or B (M <= 64: one row
// tile), which the strip updates use to apply 32 x 32 triangular inverses in
// place.
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

template <class T, bool ASSIGN = false>
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
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;  // wave quadrant
  // grid-stride over C tiles (a grid smaller than the tile count keeps its
  // workgroups on few CUs)
  for (int m0 = blockIdx.y * GT; m0 < M; m0 += gridDim.y * GT)
  for (int n0 = blockIdx.x * GT; n0 < N; n0 += gridDim.x * GT) {
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
          if (ASSIGN) {
            C[(long)row * ldc + col] = cT{pr[a][b][i], pi[a][b][i]};
          } else {
            cT c = C[(long)row * ldc + col];
            c.x -= pr[a][b][i];
            c.y -= pi[a][b][i];
            C[(long)row * ldc + col] = c;
          }
        }
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
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < (long)n * ncols; e += (long)gridDim.x * 256) {
    const int i = (int)(e / ncols), c = (int)(e % ncols);
    Y[(long)i * ldy + c] = X[(long)perm[i] * ldx + c];
  }
}

// Y[i][:] = X[i][:] for n rows of ncols (batched, per-matrix strides)
template <class T>
__global__ void k_copy_rows(int n, int ncols, const typename C2<T>::t *X, long ldx, long sX, typename C2<T>::t *Y,
                            long ldy, long sY) {
  X += (long)blockIdx.z * sX;
  Y += (long)blockIdx.z * sY;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < (long)n * ncols; e += (long)gridDim.x * 256) {
    const int i = (int)(e / ncols), c = (int)(e % ncols);
    Y[(long)i * ldy + c] = X[(long)i * ldx + c];
  }
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

// C[:, 0:NR] -= A B for few right-hand sides (NR <= 4), FP64 complex: one
// wave per row of A with its lanes 16 bytes apart along the row (each wave
// load is 1 KiB contiguous), the NR sums reduced across the wave. A is read
// once, so the residuals of the iterative refinement run at HBM speed; the
// tiled GEMM would give each 64-row tile one workgroup and leave most CUs idle.
template <int NR>
__global__ __launch_bounds__(256) void k_gemv_sub_f64(int M, int K, const v2d *__restrict__ A, long lda, long sA,
                                                      const v2d *__restrict__ B, long ldb, long sB,
                                                      v2d *__restrict__ C, long ldc, long sC) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int z = blockIdx.y;
  if (i >= M) return;
  const v2d *a = A + (long)z * sA + (long)i * lda;
  const v2d *b = B + (long)z * sB;
  double sr[NR], si[NR];
#pragma unroll
  for (int n = 0; n < NR; ++n) sr[n] = si[n] = 0.0;
#pragma unroll 4
  for (int k = lane; k < K; k += 64) {
    const v2d x = __builtin_nontemporal_load(a + k);
#pragma unroll
    for (int n = 0; n < NR; ++n) {
      const v2d y = b[(long)k * ldb + n];
      sr[n] += x.x * y.x - x.y * y.y;
      si[n] += x.x * y.y + x.y * y.x;
    }
  }
#pragma unroll
  for (int n = 0; n < NR; ++n)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sr[n] += __shfl_xor(sr[n], o);
      si[n] += __shfl_xor(si[n], o);
    }
  if (lane == 0) {
    v2d *c = C + (long)z * sC + (long)i * ldc;
#pragma unroll
    for (int n = 0; n < NR; ++n) c[n] = v2d{c[n].x - sr[n], c[n].y - si[n]};
  }
}

// Element-wise conversions and copies: grid-stride loops, so that a launch
// covers any n (a batch of Beyn nodes at order 49152 is 4.8e9 elements, more
// work-items than one dispatch can name).
__global__ void k_d2f(long n, const v2d *x, v2f *y) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    y[i] = v2f{(float)x[i].x, (float)x[i].y};
}
__global__ void k_f2d_add(long n, const v2f *d, v2d *x) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    x[i] = v2d{x[i].x + (double)d[i].x, x[i].y + (double)d[i].y};
}
__global__ void k_copy2d(long n, const v2d *x, v2d *y) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) y[i] = x[i];
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
// w <= PW) is split into chunks of R = 256 rows staged in LDS, one row per
// thread:
//   k_ts_local   each workgroup runs partial-pivoting elimination on its
//                chunk only to choose w candidate rows, and writes their
//                indices and original values;
//   k_ts_merge   up to R / PW candidate sets per workgroup are merged the
//                same way, level by level, until one set remains;
//   the last level (local when one chunk covers the strip, else the last
//                merge) also writes the w x w block L11\U11 as its
//                elimination left the pivot rows, the zgetrf-style
//                interchanges (ipiv, perm), info on a zero pivot, the
//                list of rows the interchanges move, and L11^{-1} and
//                U11^{-1};
//   k_ts_apply   moves those rows in every column of the matrix (one gather
//                per column block) and stores L11\U11;
//   then every other strip row gets L21 = A21 U11^{-1} and the next PW rows
//   of the outer block U = L11^{-1} A12, as MFMA products (k_gemm_sub).
// No workgroup ever waits for another one; each step is its own launch.
// ---------------------------------------------------------------------------
constexpr int PW_ = 32;           // strip width (FP32 and FP64)
constexpr int TS_THREADS = 256;   // threads per tournament workgroup (one row each)
constexpr int TS_WAVES = TS_THREADS / 64;
// LDS row stride (complex): conflict-free row-per-lane access with 16-byte
// lane accesses (two FP32 complex values, or one FP64)
template <class T> struct TsStride;
template <> struct TsStride<float> { static constexpr int XS = PW_ + 2; };
template <> struct TsStride<double> { static constexpr int XS = PW_ + 1; };

template <class T> struct PanelW { static constexpr int W = PW_; };
// rows per tournament workgroup, one per thread: the chunk is staged in LDS
// (70 KB FP32, 135 KB FP64)
template <class T> struct TsRows { static constexpr int R = TS_THREADS; };
template <class T> constexpr int ts_fan() { return TsRows<T>::R / PW_; }   // candidate sets merged per workgroup
template <class T> constexpr size_t ts_lds() { return (size_t)TsRows<T>::R * TsStride<T>::XS * sizeof(typename C2<T>::t); }

// Per matrix: two candidate buffers of Gmax sets (index + PW values each),
// then one output slot per strip of an outer block (the last tournament
// level's L11\U11, moved-row list and its length), kept until the block's
// interchanges have been applied to the columns outside it.
constexpr int TS_SLOTS = 8;   // strips per outer block (LU_NB / PW)
template <class T>
struct PanelWS {
  static size_t set_bytes() { return ((size_t)PW_ * 4 + (size_t)PW_ * PW_ * sizeof(typename C2<T>::t) + 255) / 256 * 256; }
  // slot: L11\U11, moved-row lists, then L11^{-1} and U11^{-1} (PW x PW each)
  __host__ __device__ static constexpr size_t fin_inv_off() { return ((size_t)PW_ * PW_ * sizeof(typename C2<T>::t) + 4 * PW_ * 4 + 16 + 255) / 256 * 256; }
  static size_t fin_bytes() { return fin_inv_off() + 2 * (size_t)PW_ * PW_ * sizeof(typename C2<T>::t); }
  static int sets(int n) { return (n + TsRows<T>::R - 1) / TsRows<T>::R; }
  static size_t bytes(int n) { return (2 * (size_t)sets(n) * set_bytes() + TS_SLOTS * fin_bytes() + 255) / 256 * 256; }
};

// Workgroup state of the elimination: each wave's packed winner and winning
// row per column (two slot sets, alternating), the pivot rows as they were
// when chosen (urow[j][j + p] = column j + p of pivot j: its U part), and
// the pivots.
template <class T>
struct TsShared {
  typedef typename C2<T>::t cT;
  unsigned long long wp[2][TS_WAVES];
  __align__(16) cT cand[2][TS_WAVES][PW_];
  __align__(16) cT urow[PW_ + 1][PW_];   // one row more: pivot j's shifted row runs j positions past its own
  int sel[PW_];    // block row of pivot j
  int skey[PW_];   // key (strip row) of pivot j
  T sval[PW_];     // 1 if the pivot j is not zero, else 0
};

// Arg-max over a wave, by DPP moves only (xor 1, xor 2, rotate 4, rotate 8
// inside each 16-lane row, then the row broadcasts; lane 63 ends with the
// wave's winner) of one packed 64-bit key per lane:
//   bits 63..32  the order bits of |x_j|_1 plus one (0: no live row) -- all 32
//                bits of an FP32 value, the upper 32 of an FP64 one (order
//                kept up to 2^-20 relative; nearer ties go to the key, as do
//                exact ties in izamax's first-maximum rule);
//   bits 31..10  0x3FFFFF - key (the smaller strip row wins a tie);
//   bits  9..0   the row within the workgroup (< 512).
// NaN magnitudes rank as zero. Non-negative floats order as their bits.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false); }
__device__ __forceinline__ unsigned int ts_order_bits(float a) { return __float_as_uint(a); }
__device__ __forceinline__ unsigned int ts_order_bits(double a) {
  return (unsigned int)((unsigned long long)__double_as_longlong(a) >> 32);
}
template <class T>
__device__ __forceinline__ unsigned long long ts_pack(bool live, T a, int key, int row) {
  if (!(a == a)) a = 0;
  const unsigned int hi = live ? ts_order_bits(a) + 1u : 0u;
  return ((unsigned long long)hi << 32) | ((unsigned long long)(0x3FFFFFu - (unsigned)key) << 10) | (unsigned)row;
}
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const unsigned lo = (unsigned)dpp_i<CTRL>((int)(unsigned)v), hi = (unsigned)dpp_i<CTRL>((int)(unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
// row broadcasts for the rows the mask enables; the others receive 0
template <int CTRL, int RM>
__device__ __forceinline__ unsigned long long dppm_u64(unsigned long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, RM, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, RM, 0xF, false);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) { return b > a ? b : a; }
__device__ __forceinline__ unsigned long long wave_amax(unsigned long long v) {
  v = umax64(v, dpp_u64<0xB1>(v));          // quad_perm [1,0,3,2]
  v = umax64(v, dpp_u64<0x4E>(v));          // quad_perm [2,3,0,1]
  v = umax64(v, dpp_u64<0x124>(v));         // row_ror 4
  v = umax64(v, dpp_u64<0x128>(v));         // row_ror 8: every lane of a row holds the row's winner
  v = umax64(v, dppm_u64<0x142, 0xA>(v));   // row_bcast 15: rows 1, 3 take rows 0, 2
  v = umax64(v, dppm_u64<0x143, 0xC>(v));   // row_bcast 31: rows 2, 3 take row 1
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

// 1 / p for the pivot as conj(p) / |p|^2 while |p|^2 is comfortably normal
// (FP32: the hardware reciprocal, one rounding more than a division; FP64: one
// IEEE division), instead of Smith's three dependent divisions, which remain
// for tiny or huge pivots.
template <class T> __device__ __forceinline__ typename C2<T>::t ts_crcp(typename C2<T>::t p);
template <> __device__ __forceinline__ v2f ts_crcp<float>(v2f p) {
  const float d = p.x * p.x + p.y * p.y;
  if (d > 1e-30f && d < 1e30f) {
    const float r = __builtin_amdgcn_rcpf(d);
    return v2f{p.x * r, -p.y * r};
  }
  return cdiv<float>(v2f{1, 0}, p);
}
template <> __device__ __forceinline__ v2d ts_crcp<double>(v2d p) {
  const double d = p.x * p.x + p.y * p.y;
  if (d > 1e-280 && d < 1e280) {
    const double r = 1.0 / d;
    return v2d{p.x * r, -p.y * r};
  }
  return cdiv<double>(v2d{1, 0}, p);
}

// Partial-pivoting elimination over the R rows of a workgroup staged in LDS
// (X[row][XS]), used only to choose pivot rows. Row i belongs to thread i
// (keys distinct, < 0 for an empty row), which holds it in registers. Per
// column j, a rolled loop (its body is a few KB of code, so the workgroup's
// single pass does not wait on instruction fetch):
//   the wave arg-max of |x_j|_1 on packed keys; each wave's winning lane
//   writes its row and lane 0 the packed key into LDS slots that alternate
//   between columns; one barrier; every thread combines the waves' keys and
//   reads the pivot row from its wave's slot (one broadcast address per
//   access); the chosen row stores itself as it is (its U part) and drops
//   out; every live row stores its multiplier l in X[row][j] and subtracts l
//   times the pivot row (four FMAs per complex entry).
// The row is kept rotated, x[p] = column j + p, shifting by one per column
// (zeros enter on the right, and stay zero: no column mask is needed).
// A wave reaches column j + 2's slots only after every wave has passed
// column j + 1's barrier, so two slot sets suffice. On return X[row][c] holds
// the multipliers of each row for the columns before its selection (all
// columns for rows not chosen), sh.urow the chosen rows' U parts and
// sh.sel/skey/sval the pivots; returns their number.
template <class T>
__device__ __forceinline__ int ts_gepp(typename C2<T>::t *X, int key, int w, TsShared<T> &sh) {
  typedef typename C2<T>::t cT;
  constexpr int XS = TsStride<T>::XS;
  static_assert(TsRows<T>::R == TS_THREADS, "one row per thread");
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  bool live = key >= 0;
  cT x[PW_];
#pragma unroll
  for (int c = 0; c < PW_; ++c) x[c] = X[t * XS + c];
  int ns = 0;
  for (int j = 0; j < w; ++j) {
    const int par = j & 1;
    const cT xj = x[0];
    const unsigned long long v = wave_amax(ts_pack<T>(live, fabs(xj.x) + fabs(xj.y), live ? key : 0, t));
    if (t == (int)(v & 0x3FFu)) {
#pragma unroll
      for (int p = 0; p < PW_; ++p) sh.cand[par][wave][p] = x[p];
    }
    if (lane == 0) sh.wp[par][wave] = v;
    __syncthreads();
    unsigned long long m = sh.wp[par][0];
#pragma unroll
    for (int q = 1; q < TS_WAVES; ++q) m = umax64(m, sh.wp[par][q]);
    const unsigned int mhi = (unsigned int)(m >> 32);
    if (mhi == 0) break;   // no row left (uniform)
    const int pr = (int)(m & 0x3FFu), pk = (int)(0x3FFFFFu - (unsigned)((m >> 10) & 0x3FFFFFu));
    const bool nonzero = mhi > 1u;   // the pivot's |x_j|_1 is not zero (uniform)
    if (t == 0) { sh.sel[j] = pr; sh.skey[j] = pk; sh.sval[j] = nonzero ? T(1) : T(0); }
    ns = j + 1;
    if (t == pr) {
      live = false;
      cT *dst = &sh.urow[j][j];
#pragma unroll
      for (int p = 0; p < PW_; ++p) dst[p] = x[p];
    }
    // Every thread shifts (straight-line code, so the shift stays in place in
    // the registers); rows that are not live, and every row at a zero pivot,
    // shift with l = 0. A live row keeps its entry at a zero pivot.
    const cT *P = sh.cand[par][pr >> 6];
    cT l = cT{0, 0};
    if (live && nonzero) l = cmul<T>(xj, ts_crcp<T>(P[0]));
    if (live) X[t * XS + j] = nonzero ? l : xj;
#pragma unroll
    for (int p = 0; p + 1 < PW_; ++p) {
      const cT u = P[p + 1], y = x[p + 1];
      x[p] = cT{fma(-l.x, u.x, fma(l.y, u.y, y.x)), fma(-l.x, u.y, fma(-l.y, u.x, y.y))};
    }
    x[PW_ - 1] = cT{0, 0};
  }
  __syncthreads();
  return ns;
}

// L11^{-1} (unit lower) and U11^{-1} (upper) of a strip's PW x PW L\U into
// out (two PW x PW blocks, row-major), for the strip's L21 and U12 updates as
// products. v: entries e = t + TS_THREADS q of L\U. PW right-looking steps
// over both inverses at once (forward for L, backward for U), each entry
// owned by one thread, one barrier per step, in the LDS of X (whose rows are
// no longer needed). A zero diagonal entry of U is taken as one: the row's
// L21 entry is then kept, as a zero pivot leaves it.
template <class T>
__device__ __forceinline__ void ts_tri_inverses(typename C2<T>::t *X, const typename C2<T>::t *v, char *out) {
  typedef typename C2<T>::t cT;
  constexpr int NE = PW_ * PW_ / TS_THREADS, PP = PW_ * PW_;
  const int t = threadIdx.x;
  cT *LU = X, *Z = X + PP, *dinv = X + 3 * PP;   // Z: L^{-1}, then U^{-1} with unscaled rows
  __syncthreads();   // every read of X's rows is done
#pragma unroll
  for (int q = 0; q < NE; ++q) {
    const int e = t + TS_THREADS * q, i = e / PW_, c = e % PW_;
    LU[e] = v[q];
    Z[e] = Z[PP + e] = (i == c) ? cT{1, 0} : cT{0, 0};
  }
  __syncthreads();
  if (t < PW_) {
    const cT d = LU[t * PW_ + t];
    dinv[t] = (d.x == 0 && d.y == 0) ? cT{1, 0} : cdiv<T>(cT{1, 0}, d);
  }
  __syncthreads();
  for (int s = 0; s < PW_; ++s) {
    const int su = PW_ - 1 - s;
#pragma unroll
    for (int q = 0; q < 2 * NE; ++q) {
      const int e = t + TS_THREADS * q;
      if (e < PP) {   // L: rows below s take row s
        const int i = e / PW_, c = e % PW_;
        if (i > s) {
          const cT l = LU[i * PW_ + s], z = Z[s * PW_ + c];
          Z[e] = cT{Z[e].x - (l.x * z.x - l.y * z.y), Z[e].y - (l.x * z.y + l.y * z.x)};
        }
      } else {        // U: rows above su take row su divided by its diagonal
        const int f = e - PP, i = f / PW_, c = f % PW_;
        if (i < su) {
          const cT u = cmul<T>(LU[i * PW_ + su], dinv[su]), z = Z[PP + su * PW_ + c];
          Z[e] = cT{Z[e].x - (u.x * z.x - u.y * z.y), Z[e].y - (u.x * z.y + u.y * z.x)};
        }
      }
    }
    __syncthreads();
  }
  cT *o = reinterpret_cast<cT *>(out);
#pragma unroll
  for (int q = 0; q < 2 * NE; ++q) {
    const int e = t + TS_THREADS * q;
    o[e] = (e < PP) ? Z[e] : cmul<T>(Z[e], dinv[(e - PP) / PW_]);
  }
}

// Output of the last tournament level: L11\U11 rows, ipiv (1-based, matrix
// rows), perm, info, and the compacted list of (from, to) strip rows the
// interchanges move. The list starts as the w top rows; each interchange
// i <-> p is applied to it with wave ballots (a row not yet listed sits at
// its own index).
template <class T>
__device__ __forceinline__ void ts_finish(typename C2<T>::t *X, int ns, int w, int c0, TsShared<T> &sh,
                                          char *fin, int *ipiv, int *perm, int *info) {
  typedef typename C2<T>::t cT;
  constexpr int XS = TsStride<T>::XS, NE = PW_ * PW_ / TS_THREADS;
  const int t = threadIdx.x;
  cT *lu = reinterpret_cast<cT *>(fin);
  int *lfrom = reinterpret_cast<int *>(fin + PW_ * PW_ * sizeof(cT));
  int *lto = lfrom + 2 * PW_;
  int *lcount = lto + 2 * PW_;
  cT v[NE];   // entries t + TS_THREADS q of L11\U11
#pragma unroll
  for (int q = 0; q < NE; ++q) {
    const int e = t + TS_THREADS * q, i = e / PW_, c = e % PW_;
    v[q] = (i < ns && c < w) ? (c >= i ? sh.urow[i][c] : X[sh.sel[i] * XS + c]) : cT{0, 0};
    lu[e] = v[q];
  }
  if (t < 64) {
    int id = (t < w) ? t : -1, pos = id, nlist = w;
    for (int i = 0; i < ns; ++i) {
      const int tg = sh.skey[i];
      const unsigned long long mt = __ballot(id == tg);
      const int lt = mt ? __ffsll((long long)mt) - 1 : -1;
      const int p = (lt >= 0) ? __shfl(pos, lt) : tg;
      const unsigned long long mi = __ballot(id >= 0 && pos == i);
      const int li = __ffsll((long long)mi) - 1;
      if (p != i) {
        if (t == li) pos = p;
        if (lt >= 0) {
          if (t == lt) pos = i;
        } else {
          if (t == nlist) { id = tg; pos = i; }
          ++nlist;
        }
      }
      if (t == 0) ipiv[c0 + i] = c0 + p + 1;
    }
    if (t == 0) {
      int z0 = 0;
      for (int i = ns - 1; i >= 0; --i)
        if (sh.sval[i] == 0) z0 = c0 + i + 1;
      if (z0 && *info == 0) *info = z0;
    }
    const bool mv = (t < nlist) && id != pos;
    const unsigned long long mm = __ballot(mv);
    const int slot = __popcll(mm & ((1ull << t) - 1ull));
    if (mv) { lfrom[slot] = id; lto[slot] = pos; }
    if (t == 0) *lcount = __popcll(mm);
    if (perm) {   // the moved rows as one gather: every lane loads before any stores
      const int pv = mv ? perm[c0 + id] : 0;
      if (mv) perm[c0 + pos] = pv;
    }
  }
  ts_tri_inverses<T>(X, v, fin + PanelWS<T>::fin_inv_off());
}

template <class T>
__global__ __launch_bounds__(TS_THREADS) void k_ts_local(int n, int c0, int w, const typename C2<T>::t *A, long lda,
                                                         long sA, char *ws, long sW, long setb, long finoff,
                                                         int final_, int *ipiv, long sP, int *perm, int *info) {
  typedef typename C2<T>::t cT;
  constexpr int R = TsRows<T>::R, XS = TsStride<T>::XS;
  extern __shared__ __align__(16) char smem_raw[];
  cT *X = reinterpret_cast<cT *>(smem_raw);
  __shared__ TsShared<T> sh;
  const int z = blockIdx.z, g = blockIdx.x, t = threadIdx.x;
  A += (long)z * sA;
  ws += (long)z * sW;
  const int m = n - c0, r0 = g * R, nr = min(R, m - r0);
  {   // coalesced along each row; every load in flight before the first LDS store
    constexpr int PER = R * PW_ / TS_THREADS;
    cT v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = t + TS_THREADS * q, r = e / PW_, c = e % PW_;
      v[q] = (r < nr && c < w) ? A[(long)(c0 + r0 + r) * lda + c0 + c] : cT{0, 0};
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = t + TS_THREADS * q;
      X[(e / PW_) * XS + e % PW_] = v[q];
    }
  }
  __syncthreads();
  const int ns = ts_gepp<T>(X, t < nr ? r0 + t : -1, w, sh);
  if (final_) {
    ts_finish<T>(X, ns, w, c0, sh, ws + finoff, ipiv + (long)z * sP, perm ? perm + (long)z * sP : nullptr, info + z);
    return;
  }
  char *set = ws + (long)g * setb;
  int *oidx = reinterpret_cast<int *>(set);
  cT *oval = reinterpret_cast<cT *>(set + PW_ * 4);
  if (t < PW_) oidx[t] = (t < ns) ? sh.skey[t] : -1;
  constexpr int PO = PW_ * PW_ / TS_THREADS;
  cT v[PO];
#pragma unroll
  for (int q = 0; q < PO; ++q) {   // the chosen rows' original values
    const int e = t + TS_THREADS * q, i = e / PW_, c = e % PW_;
    v[q] = (i < ns && c < w) ? A[(long)(c0 + sh.skey[i]) * lda + c0 + c] : cT{0, 0};
  }
#pragma unroll
  for (int q = 0; q < PO; ++q) {
    const int e = t + TS_THREADS * q;
    if (e < ns * PW_) oval[e] = v[q];
  }
}

template <class T>
__global__ __launch_bounds__(TS_THREADS) void k_ts_merge(int c0, int w, int nin, const char *wsin, char *wsout,
                                                         long sW, long setb, char *wsfin, int final_, int *ipiv,
                                                         long sP, int *perm, int *info) {
  typedef typename C2<T>::t cT;
  constexpr int R = TsRows<T>::R, F = ts_fan<T>(), XS = TsStride<T>::XS;
  extern __shared__ __align__(16) char smem_raw[];
  cT *X = reinterpret_cast<cT *>(smem_raw);
  __shared__ TsShared<T> sh;
  const int z = blockIdx.z, b = blockIdx.x, t = threadIdx.x;
  const char *in = wsin + (long)z * sW;
  const int s0 = b * F, nset = min(F, nin - s0);
  {   // every load in flight before the first LDS store
    constexpr int PER = R * PW_ / TS_THREADS;
    cT v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = t + TS_THREADS * q, r = e / PW_, c = e % PW_, si = r / PW_, i = r % PW_;
      v[q] = (si < nset) ? reinterpret_cast<const cT *>(in + (long)(s0 + si) * setb + PW_ * 4)[i * PW_ + c]
                         : cT{0, 0};
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = t + TS_THREADS * q;
      X[(e / PW_) * XS + e % PW_] = v[q];
    }
  }
  int key = -1;
  if (t < R && t / PW_ < nset) key = reinterpret_cast<const int *>(in + (long)(s0 + t / PW_) * setb)[t % PW_];
  __syncthreads();
  const int ns = ts_gepp<T>(X, key, w, sh);
  if (final_) {
    ts_finish<T>(X, ns, w, c0, sh, wsfin + (long)z * sW, ipiv + (long)z * sP, perm ? perm + (long)z * sP : nullptr,
                 info + z);
    return;
  }
  char *out = wsout + (long)z * sW + (long)b * setb;
  int *oidx = reinterpret_cast<int *>(out);
  cT *oval = reinterpret_cast<cT *>(out + PW_ * 4);
  if (t < PW_) oidx[t] = (t < ns) ? sh.skey[t] : -1;
  constexpr int PO = PW_ * PW_ / TS_THREADS;
  cT v[PO];
#pragma unroll
  for (int q = 0; q < PO; ++q) {   // the chosen rows' original values, from their input set
    const int e = t + TS_THREADS * q, i = e / PW_, c = e % PW_, r = sh.sel[i];
    v[q] = (i < ns) ? reinterpret_cast<const cT *>(in + (long)(s0 + r / PW_) * setb + PW_ * 4)[(r % PW_) * PW_ + c]
                    : cT{0, 0};
  }
#pragma unroll
  for (int q = 0; q < PO; ++q) {
    const int e = t + TS_THREADS * q;
    if (e < ns * PW_) oval[e] = v[q];
  }
}

// ---------------------------------------------------------------------------
// Partial-pivoting strips, for the entry points that return the factors
// (qpms_hip_zgetrf_dev and the ABI factorisation built on it): the pivots of
// zgetrf (zgetf2's column loop, scatsystem.c:2214): pivot j is the first row
// of largest |re| + |im| in column j at or below the diagonal, taken in full
// double (or float) precision. One launch per column (k_pp_step): every
// workgroup applies the previous column's interchange and elimination to its
// rows of the strip (the pivot row and the old top row come from copies in the
// workspace, so no workgroup reads a row another one rewrites), then reduces
// the next column's candidates; the last workgroup to finish (an atomic
// count after a device-scope release) combines the partial maxima, records
// the pivot and copies the pivot row and the top row for the next launch.
// k_pp_finish then writes the strip's slot (L11\U11, the moved-row list,
// L11^{-1}, U11^{-1}) as the tournament's last level does, so the rest of the
// LU is shared. L21 is left in place by the elimination.
// ---------------------------------------------------------------------------
constexpr int PP_ROWS = 64;   // rows per workgroup: 8 passes of 8 rows x 32 columns
// per-matrix workspace layout (in the tournament's candidate area, unused here)
constexpr long PP_CNT = 0, PP_PROW = 256, PP_TROW = 768, PP_PIV = 1280, PP_NZ = 1408, PP_PART = 1536;
__host__ __device__ constexpr long pp_part_row_off(int G) { return PP_PART + ((long)G * 8 + 255) / 256 * 256; }
__host__ __device__ constexpr long pp_ws_bytes(int n) {
  return pp_part_row_off((n + PP_ROWS - 1) / PP_ROWS) + ((long)((n + PP_ROWS - 1) / PP_ROWS) * 4 + 255) / 256 * 256;
}

__global__ void k_pp_init(char *ws, long sW) { *reinterpret_cast<int *>(ws + (long)blockIdx.x * sW + PP_CNT) = 0; }

// (magnitude, row) pairs: larger magnitude first, then the smaller row
__device__ __forceinline__ void pp_better(double &m, int &r, double m2, int r2) {
  if (m2 > m || (m2 == m && r2 < r)) { m = m2; r = r2; }
}
__device__ __forceinline__ void pp_reduce_wg(double &m, int &r, double *sm, int *sr) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double m2 = __shfl_xor(m, off);
    const int r2 = __shfl_xor(r, off);
    pp_better(m, r, m2, r2);
  }
  const int t = threadIdx.x, wave = t >> 6;
  if ((t & 63) == 0) { sm[wave] = m; sr[wave] = r; }
  __syncthreads();
  m = sm[0];
  r = sr[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) pp_better(m, r, sm[q], sr[q]);
}

template <class T>
__global__ __launch_bounds__(256) void k_pp_step(int n, int c0, int w, int j, typename C2<T>::t *A, long lda, long sA,
                                                 char *ws, long sW, int *ipiv, long sP, int *info) {
  typedef typename C2<T>::t cT;
  const int z = blockIdx.z, g = blockIdx.x, G = gridDim.x, t = threadIdx.x, k = t & 31, rr = t >> 5;
  A += (long)z * sA;
  char *wz = ws + (long)z * sW;
  ipiv += (long)z * sP;
  info += z;
  cT *prow = reinterpret_cast<cT *>(wz + PP_PROW), *trow = reinterpret_cast<cT *>(wz + PP_TROW);
  int *piv = reinterpret_cast<int *>(wz + PP_PIV), *nz = reinterpret_cast<int *>(wz + PP_NZ);
  double *part_m = reinterpret_cast<double *>(wz + PP_PART);
  int *part_r = reinterpret_cast<int *>(wz + pp_part_row_off(G));
  int *cnt = reinterpret_cast<int *>(wz + PP_CNT);
  const int s = j - 1, r = c0 + s;
  int p = -1;
  bool pnz = false;
  cT pk = cT{0, 0}, tk = cT{0, 0}, rcp = cT{0, 0};
  if (j > 0) {
    p = piv[s];
    pnz = nz[s] != 0;
    if (k < w) { pk = prow[k]; tk = trow[k]; }
    if (pnz) rcp = cdiv<T>(cT{1, 0}, prow[s]);
  }
  double best = -1.0;
  int bestrow = 0x7fffffff;
  for (int q = 0; q < PP_ROWS / 8; ++q) {
    const int i = c0 + g * PP_ROWS + q * 8 + rr;   // uniform per 32-lane group
    const bool inrow = i < n;
    cT x = (inrow && k < w) ? A[(long)i * lda + c0 + k] : cT{0, 0};
    if (j > 0) {
      if (i == r) x = pk;            // the pivot row moves to the top
      else if (i == p) x = tk;       // the old top row takes its place
      const T xs_re = __shfl(x.x, s, 32), xs_im = __shfl(x.y, s, 32);
      if (i > r && pnz) {
        const cT l = cmul<T>(cT{xs_re, xs_im}, rcp);
        if (k == s) x = l;
        else if (k > s) x = cT{x.x - (l.x * pk.x - l.y * pk.y), x.y - (l.x * pk.y + l.y * pk.x)};
      }
      if (inrow && k < w && i >= r) A[(long)i * lda + c0 + k] = x;
    }
    if (j < w && inrow && k == j && i >= c0 + j) pp_better(best, bestrow, (double)(fabs(x.x) + fabs(x.y)), i);
  }
  __shared__ double sm[4];
  __shared__ int sr[4];
  __shared__ int last;
  pp_reduce_wg(best, bestrow, sm, sr);
  if (t == 0) { part_m[g] = best; part_r[g] = bestrow; }
  __threadfence();   // release: this workgroup's rows and partial maximum
  if (t == 0) last = (atomicAdd(cnt, 1) == G - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();   // acquire: every other workgroup's rows and partial maxima
  if (j < w) {
    best = -1.0;
    bestrow = 0x7fffffff;
    for (int q = t; q < G; q += blockDim.x) pp_better(best, bestrow, part_m[q], part_r[q]);
    __syncthreads();   // sm/sr reused
    pp_reduce_wg(best, bestrow, sm, sr);
    if (bestrow < c0 + j || bestrow >= n) { bestrow = c0 + j; best = 0.0; }   // no comparable candidate (NaN column)
    if (t == 0) {
      piv[j] = bestrow;
      nz[j] = best > 0.0;
      ipiv[c0 + j] = bestrow + 1;
      if (!(best > 0.0) && *info == 0) *info = c0 + j + 1;
    }
    if (t < w) prow[t] = A[(long)bestrow * lda + c0 + t];
    else if (t >= 32 && t - 32 < w) trow[t - 32] = A[(long)(c0 + j) * lda + c0 + t - 32];
  }
  if (t == 0) *cnt = 0;
}

template <class T>
__global__ __launch_bounds__(256) void k_pp_finish(int n, int c0, int w, const typename C2<T>::t *A, long lda,
                                                   long sA, char *ws, long sW, long finoff, int *perm, long sP) {
  typedef typename C2<T>::t cT;
  constexpr int NE = PW_ * PW_ / 256;
  extern __shared__ __align__(16) char smem_raw[];
  cT *X = reinterpret_cast<cT *>(smem_raw);
  const int z = blockIdx.z, t = threadIdx.x;
  A += (long)z * sA;
  char *wz = ws + (long)z * sW;
  char *fin = wz + finoff;
  const int *piv = reinterpret_cast<const int *>(wz + PP_PIV);
  cT *lu = reinterpret_cast<cT *>(fin);
  int *lfrom = reinterpret_cast<int *>(fin + PW_ * PW_ * sizeof(cT));
  int *lto = lfrom + 2 * PW_;
  int *lcount = lto + 2 * PW_;
  cT v[NE];
#pragma unroll
  for (int q = 0; q < NE; ++q) {
    const int e = t + 256 * q, i = e / PW_, c = e % PW_;
    v[q] = (i < w && c < w) ? A[(long)(c0 + i) * lda + c0 + c] : cT{0, 0};
    lu[e] = v[q];
  }
  if (t < 64) {   // the interchanges as a list of moved rows: lane entry (original strip row, current row)
    int id = -1, pos = -1, nlist = 0;
    for (int i = 0; i < w; ++i) {
      const int pr = piv[i] - c0;
      if (pr == i) continue;
      if (!__ballot(pos == i)) { if (t == nlist) { id = i; pos = i; } ++nlist; }
      if (!__ballot(pos == pr)) { if (t == nlist) { id = pr; pos = pr; } ++nlist; }
      const bool at_i = pos == i, at_p = pos == pr;
      if (at_i) pos = pr;
      else if (at_p) pos = i;
    }
    const bool mv = (t < nlist) && id != pos;
    const unsigned long long mm = __ballot(mv);
    const int slot = __popcll(mm & ((1ull << t) - 1ull));
    if (mv) { lfrom[slot] = id; lto[slot] = pos; }
    if (t == 0) *lcount = __popcll(mm);
    if (perm) {   // every lane loads before any stores
      int *pz = perm + (long)z * sP;
      const int pv = mv ? pz[c0 + id] : 0;
      if (mv) pz[c0 + pos] = pv;
    }
  }
  ts_tri_inverses<T>(X, v, fin + PanelWS<T>::fin_inv_off());
}
constexpr size_t PP_FIN_LDS = (size_t)(3 * PW_ * PW_ + PW_) * 16;   // ts_tri_inverses' area (FP64 size)

// Row interchanges of strips as gathers A[c0 + to] = A[c0 + from] over their
// moved rows, per block of AP_COLS columns: every moved entry is loaded into
// registers, then (after a barrier) stored.
//   k_ts_apply:  one strip (its slot `fin`), columns [col_lo, col_hi), then
//                L11\U11 into the strip's top block;
//   k_ts_apply_defer: the strips of an outer block (slots 0..nslots-1, strip
//                c0 = K0 + PW slot, in order) on every column outside
//                [skip_lo, skip_hi).
constexpr int AP_COLS = 32;
template <class T>
__device__ __forceinline__ void ts_gather(int n, int c0, const char *fin, typename C2<T>::t *A, long lda, int j0,
                                          int *from, int *to, bool colok) {
  typedef typename C2<T>::t cT;
  const int t = threadIdx.x;
  const int *lfrom = reinterpret_cast<const int *>(fin + PW_ * PW_ * sizeof(cT));
  const int *lto = lfrom + 2 * PW_;
  const int k = lto[2 * PW_];
  if (t < k) { from[t] = lfrom[t]; to[t] = lto[t]; }
  __syncthreads();
  const int c = t % AP_COLS, r4 = t / AP_COLS;
  constexpr int RP = 256 / AP_COLS, NQ = 2 * PW_ / RP;   // rows per pass; passes over the longest list
  cT v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {   // every load in flight, then (after the barrier) every store
    const int e = r4 + RP * q;
    v[q] = (colok && e < k) ? A[(long)(c0 + from[e]) * lda + j0 + c] : cT{0, 0};
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int e = r4 + RP * q;
    if (colok && e < k) A[(long)(c0 + to[e]) * lda + j0 + c] = v[q];
  }
  __syncthreads();
}

template <class T>
__global__ __launch_bounds__(256) void k_ts_apply(int n, int c0, int w, typename C2<T>::t *A, long lda, long sA,
                                                  const char *fin, long sW, int col_lo, int col_hi) {
  typedef typename C2<T>::t cT;
  __shared__ int from[2 * PW_], to[2 * PW_];
  const int z = blockIdx.z, t = threadIdx.x, j0 = col_lo + blockIdx.x * AP_COLS;
  A += (long)z * sA;
  fin += (long)z * sW;
  const int c = t % AP_COLS, r4 = t / AP_COLS;
  const bool colok = j0 + c < col_hi;
  ts_gather<T>(n, c0, fin, A, lda, j0, from, to, colok);
  const cT *lu = reinterpret_cast<const cT *>(fin);
  if (colok && j0 + c >= c0 && j0 + c < c0 + w)
    for (int i = r4; i < w; i += 256 / AP_COLS) A[(long)(c0 + i) * lda + j0 + c] = lu[i * PW_ + (j0 + c - c0)];
}

template <class T>
__global__ __launch_bounds__(256) void k_ts_apply_defer(int n, int K0, int nslots, typename C2<T>::t *A, long lda,
                                                        long sA, const char *fin0, long finb, long sW, int skip_lo,
                                                        int skip_hi) {
  __shared__ int from[2 * PW_], to[2 * PW_];
  const int z = blockIdx.z, j0 = blockIdx.x * AP_COLS;
  if (j0 >= skip_lo && j0 + AP_COLS <= skip_hi) return;   // whole block inside the skipped columns (uniform)
  A += (long)z * sA;
  fin0 += (long)z * sW;
  const int c = threadIdx.x % AP_COLS;
  const bool colok = j0 + c < n && (j0 + c < skip_lo || j0 + c >= skip_hi);
  for (int q = 0; q < nslots; ++q) ts_gather<T>(n, K0 + PW_ * q, fin0 + q * finb, A, lda, j0, from, to, colok);
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

// Per-stream state of the LU driver: a rocBLAS handle bound to the stream
// and the lookahead side stream with its two events. Created on a stream's
// first use; qh_stream_release() destroys it, and callers that destroy a
// stream (the serving loop's lanes) must call it first, so that a later
// stream created at the same address never inherits a dead stream's state.
struct LaCtx {
  hipEvent_t go, done;   // ordering of the caller's stream against the device's side stream
};

constexpr int STREAM_TAB = 64;
struct StreamState {
  hipStream_t key;
  rocblas_handle blas;   // nullptr until the first library GEMM on the stream
  bool la;               // lookahead context created
  LaCtx lc;
};
static std::mutex g_tab_mu;
static StreamState g_tab[STREAM_TAB];
static int g_ntab = 0;

static StreamState *state_for(hipStream_t s) {   // caller holds g_tab_mu
  for (int i = 0; i < g_ntab; ++i) if (g_tab[i].key == s) return &g_tab[i];
  if (g_ntab == STREAM_TAB)
    qpms_hip_die(__FILE__, __LINE__, __func__, "more than %d live streams use the LU (release them)", STREAM_TAB);
  StreamState &e = g_tab[g_ntab++];
  e = StreamState{};
  e.key = s;
  return &e;
}

static rocblas_handle blas_for(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_tab_mu);
  StreamState *e = state_for(s);
  if (!e->blas) {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      qpms_hip_die(__FILE__, __LINE__, __func__, "rocblas_create_handle failed");
    if (rocblas_set_stream(h, s) != rocblas_status_success)
      qpms_hip_die(__FILE__, __LINE__, __func__, "rocblas_set_stream failed");
    e->blas = h;
  }
  return e->blas;
}

// CUs kept free of the side stream's trailing update, so that the next
// block's strips (latency-bound launches of tens of workgroups, launched as
// whole-CU workgroups) find CUs while it runs: QPMS_HIP_LU_RESERVE (default
// 32; 0: no CU mask, the strips then wait for CUs to drain).
static int lu_reserve_cus() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("QPMS_HIP_LU_RESERVE");
    v = e ? std::max(0, atoi(e)) : 32;
  }
  return v;
}

// The side stream of the lookahead: one per device for the whole process,
// masked to all CUs but the reserved ones, shared by every caller stream (the
// bulk updates of concurrent factorisations queue on it, each ordered against
// its own caller by that caller's events). Every further hardware queue
// costs: with several masked side streams alive, factorisations measured up
// to 1.6 x slower, depending on the order the queues were created in.
constexpr int LA_MAX_DEV = 64;
static hipStream_t g_la_side[LA_MAX_DEV];
static bool g_la_masked[LA_MAX_DEV];
static hipStream_t la_side(bool *masked) {   // caller holds g_tab_mu
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= LA_MAX_DEV) qpms_hip_die(__FILE__, __LINE__, __func__, "device %d outside the table", dev);
  if (!g_la_side[dev]) {
    const int reserve = lu_reserve_cus();
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (reserve > 0 && ncu > 2 * reserve && ncu <= 32 * 32) {
      uint32_t mask[32] = {0};
      for (int i = 0; i < ncu - reserve; ++i) mask[i / 32] |= 1u << (i % 32);
      HIPCHK(hipExtStreamCreateWithCUMask(&g_la_side[dev], (uint32_t)((ncu + 31) / 32), mask));
      g_la_masked[dev] = true;
    } else {
      HIPCHK(hipStreamCreateWithFlags(&g_la_side[dev], hipStreamNonBlocking));
      g_la_masked[dev] = false;
    }
  }
  *masked = g_la_masked[dev];
  return g_la_side[dev];
}

// Lookahead context of a caller stream: its two events (created on first use).
static LaCtx *la_for(hipStream_t s, hipStream_t *side, bool *masked) {
  std::lock_guard<std::mutex> g(g_tab_mu);
  StreamState *e = state_for(s);
  if (!e->la) {
    HIPCHK(hipEventCreateWithFlags(&e->lc.go, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&e->lc.done, hipEventDisableTiming));
    e->la = true;
  }
  *side = la_side(masked);
  return &e->lc;
}

// Destroys the per-stream state of `stream` (rocBLAS handle, lookahead
// events), after waiting for its work on the shared side stream.
extern "C" void qh_stream_release(void *stream) {
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> g(g_tab_mu);
  for (int i = 0; i < g_ntab; ++i) {
    if (g_tab[i].key != s) continue;
    StreamState &e = g_tab[i];
    if (e.la) {   // the caller's own work on the shared side stream ends before its events go
      HIPCHK(hipEventSynchronize(e.lc.done));
      HIPCHK(hipEventDestroy(e.lc.go));
      HIPCHK(hipEventDestroy(e.lc.done));
    }
    if (e.blas) rocblas_destroy_handle(e.blas);
    g_tab[i] = g_tab[--g_ntab];
    return;
  }
}

// Number of streams with live per-stream state (tests check that servers
// release theirs).
extern "C" int qh_stream_state_count(void) {
  std::lock_guard<std::mutex> g(g_tab_mu);
  return g_ntab;
}

static bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

template <class T>
static void trailing_gemm(int M, int N, int K, int batch, const typename C2<T>::t *A, long lda, long sA,
                          const typename C2<T>::t *B, long ldb, long sB, typename C2<T>::t *C, long ldc, long sC,
                          hipStream_t s, bool own_only = false) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  // inside a captured graph only hand-written kernels (a library call may
  // record allocations or memsets into the graph)
  const bool capturing = stream_capturing(s);
  if (capturing) own_only = true;
  if (sizeof(T) == 4 && (!own_only || capturing)) {
    // FP32: the variant the autotuner picked for this size class
    int v = qh_lu_gemm_variant(M, N, batch > 1);
    if (v == 0 && capturing) v = qh_lu_gemm_best_own(M, N, batch > 1);
    qh_lu_gemm_run(v, M, N, K, batch, (const float *)A, lda, sA, (const float *)B, ldb, sB, (float *)C, ldc, sC,
                   v == 0 ? (void *)blas_for(s) : nullptr, s);
    return;
  }
  if (own_only || !use_rocblas() || (long)M * N < 512L * 512) {
    QH_LAUNCH(k_gemm_sub<T>, dim3((N + GT - 1) / GT, (M + GT - 1) / GT, batch), dim3(256), 0, s, M, N, K, A,
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
static_assert(LU_NB / PW_ == TS_SLOTS, "one output slot per strip of an outer block");

static bool use_lookahead() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("QPMS_HIP_LU_LOOKAHEAD");
    v = !(e && !strcmp(e, "0"));
  }
  return v == 1;
}

// C -= A B (or C = A B, ASSIGN) for K <= PW: the strip's products (L21 =
// A21 U11^{-1}, U = L11^{-1} A12, the in-block rank-PW update) when they run
// as whole-CU workgroups beside the trailing update (one workgroup per CU).
// The workgroup is gcu_q quarters of 256 threads, each quarter on its own
// 64 x 64 tile of C, so that a CU has that many tiles in flight; a tile's A
// (64 x K) and B (K x 64) are staged whole (one load round trip for the
// products instead of one per GK slice), and the waves of the quarters hide
// each other's latency. The quarters step through groups of tiles together
// (every barrier is reached by all of them; a quarter past the last tile
// loads and stores nothing).
template <class T> constexpr int gcu_q() { return sizeof(T) == 4 ? 4 : 2; }
template <class T> constexpr size_t gcu_quarter_lds() {
  return (size_t)(2 * GT * (PW_ + 1) + 2 * PW_ * (GT + 1)) * sizeof(T);
}
template <class T> constexpr size_t gcu_lds() { return gcu_q<T>() * gcu_quarter_lds<T>(); }

template <class T, bool ASSIGN>
__global__ __launch_bounds__(gcu_q<T>() * 256) void k_gemm_cu(int M, int N, int K, const typename C2<T>::t *A,
                                                               long lda, const typename C2<T>::t *B, long ldb,
                                                               typename C2<T>::t *C, long ldc) {
  typedef typename C2<T>::t cT;
  typedef typename MF<T>::acc acc_t;
  constexpr int Q = gcu_q<T>(), AS = PW_ + 1, BS = GT + 1;
  extern __shared__ __align__(16) char smem_raw[];
  const int sub = threadIdx.x >> 8, tid = threadIdx.x & 255, lane = tid & 63, wv = tid >> 6;
  T *Ar = reinterpret_cast<T *>(smem_raw + sub * gcu_quarter_lds<T>());
  T *Ai = Ar + GT * AS, *Br = Ai + GT * AS, *Bi = Br + PW_ * BS;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;   // wave quadrant of the tile
  const int lr = lane & 15, lk = lane >> 4;
  const int tn = (N + GT - 1) / GT, nt = ((M + GT - 1) / GT) * tn, ngrp = (nt + Q - 1) / Q;
  for (int g = blockIdx.x; g < ngrp; g += gridDim.x) {
    const int tile = g * Q + sub;
    const bool valid = tile < nt;
    const int m0 = valid ? (tile / tn) * GT : M, n0 = valid ? (tile % tn) * GT : N;
    cT va[GT * PW_ / 256], vb[PW_ * GT / 256];
#pragma unroll
    for (int e = 0; e < GT * PW_ / 256; ++e) {
      const int idx = tid + 256 * e;
      const int ar = idx / PW_, ak = idx % PW_, bk = idx / GT, bn = idx % GT;
      va[e] = (m0 + ar < M && ak < K) ? A[(long)(m0 + ar) * lda + ak] : cT{0, 0};
      vb[e] = (bk < K && n0 + bn < N) ? B[(long)bk * ldb + n0 + bn] : cT{0, 0};
    }
    __syncthreads();   // the previous group's reads of the tiles are done
#pragma unroll
    for (int e = 0; e < GT * PW_ / 256; ++e) {
      const int idx = tid + 256 * e;
      const int ar = idx / PW_, ak = idx % PW_, bk = idx / GT, bn = idx % GT;
      Ar[ar * AS + ak] = va[e].x;
      Ai[ar * AS + ak] = va[e].y;
      Br[bk * BS + bn] = vb[e].x;
      Bi[bk * BS + bn] = vb[e].y;
    }
    __syncthreads();
    acc_t pr[2][2], pi[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) { pr[a][b] = acc_t{0, 0, 0, 0}; pi[a][b] = acc_t{0, 0, 0, 0}; }
#pragma unroll
    for (int kk = 0; kk < PW_; kk += 4) {
      if (kk >= K) break;   // uniform; entries past K are zero
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const T ar = Ar[(wm + a * 16 + lr) * AS + kk + lk], ai = Ai[(wm + a * 16 + lr) * AS + kk + lk];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const T br = Br[(kk + lk) * BS + wn + b * 16 + lr], bi = Bi[(kk + lk) * BS + wn + b * 16 + lr];
          pr[a][b] = MF<T>::mma(ar, br, pr[a][b]);
          pr[a][b] = MF<T>::mma(-ai, bi, pr[a][b]);
          pi[a][b] = MF<T>::mma(ar, bi, pi[a][b]);
          pi[a][b] = MF<T>::mma(ai, br, pi[a][b]);
        }
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = m0 + wm + a * 16 + MF<T>::row(lk, i), col = n0 + wn + b * 16 + lr;
          if (row < M && col < N) {
            if (ASSIGN) {
              C[(long)row * ldc + col] = cT{pr[a][b][i], pi[a][b][i]};
            } else {
              const cT c = C[(long)row * ldc + col];
              C[(long)row * ldc + col] = cT{c.x - pr[a][b][i], c.y - pi[a][b][i]};
            }
          }
        }
  }
}

// Strips that run beside the side stream's trailing update launch every
// kernel with a dynamic-LDS request that fills the rest of a CU's 160 KB, so
// that none of their workgroups shares a CU with a GEMM workgroup (they would
// get a small share of its issue slots); they then run on the CUs the side
// stream's mask leaves free. The request per kernel is set once with its
// dynamic-LDS limit, from the kernel's static LDS.
constexpr size_t CU_LDS = (size_t)160 << 10;
static size_t whole_cu_lds(const void *fn) {
  hipFuncAttributes at;
  HIPCHK(hipFuncGetAttributes(&at, fn));
  const size_t stat = at.sharedSizeBytes;
  const size_t pad = (CU_LDS > stat + 512 ? CU_LDS - stat - 512 : 0) / 256 * 256;
  if (pad > 0) HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pad));
  return pad;
}
template <class T>
struct WholeCu {   // dynamic-LDS requests of the strip kernels for whole-CU launches
  size_t local, merge, apply, gemm_assign, gemm_sub;
};
template <class T>
static const WholeCu<T> &whole_cu() {
  static WholeCu<T> w;
  static std::once_flag once;
  std::call_once(once, [] {
    w.local = std::max(whole_cu_lds((const void *)k_ts_local<T>), ts_lds<T>());
    w.merge = std::max(whole_cu_lds((const void *)k_ts_merge<T>), ts_lds<T>());
    w.apply = whole_cu_lds((const void *)k_ts_apply<T>);
    w.gemm_assign = std::max(whole_cu_lds((const void *)k_gemm_cu<T, true>), gcu_lds<T>());
    w.gemm_sub = std::max(whole_cu_lds((const void *)k_gemm_cu<T, false>), gcu_lds<T>());
  });
  return w;
}

// The tournament kernels stage a chunk of more than 64 KB of LDS: raise
// their dynamic-LDS limit once (before any launch sequence is captured).
template <class T>
static void ts_prepare() {
  static std::once_flag once;
  std::call_once(once, [] {
    HIPCHK(hipFuncSetAttribute((const void *)k_ts_local<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ts_lds<T>()));
    HIPCHK(hipFuncSetAttribute((const void *)k_ts_merge<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ts_lds<T>()));
  });
  (void)whole_cu<T>();   // raises the limits further for whole-CU launches
}

extern "C" void qh_cgemm3m_prepare(void);
// Every kernel attribute the LU paths set, set now (the serving loop calls
// this at creation, before it captures any launch sequence into a graph).
template <class T> static void tsb_prepare();
extern "C" void qh_lu_prepare(void) {
  ts_prepare<float>();
  ts_prepare<double>();
  tsb_prepare<float>();
  tsb_prepare<double>();
  qh_cgemm3m_prepare();
}

// Workspace bytes for getrf_batched<T>(n, batch).
template <class T>
static size_t getrf_ws_bytes(int n, int batch) {
  return (size_t)batch * PanelWS<T>::bytes(n);
}

// Blocked right-looking LU with tournament-pivoted panels of `batch` n x n
// matrices (stride sA), in place. Outer blocks of LU_NB columns; inside a
// block the columns are factorised PW at a time (k_ts_*), the block's
// remaining columns updated by products with the strip's L11^{-1} and a GEMM
// (K = PW); then U12 = L11^{-1} A12 by PW rows the same way, and the trailing
// A22 -= L21 U12 (K = LU_NB).
// ipiv (1-based, stride sP) lists the row interchanges as zgetrf's does;
// perm (optional) tracks P; info[z] gets the first zero pivot. ws:
// getrf_ws_bytes<T>(n, batch) bytes, or nullptr for a library-owned buffer
// (single-stream callers only).
template <class T>
static void getrf_batched(int n, int batch, typename C2<T>::t *A, long lda, long sA, int *ipiv, long sP,
                          int *perm, int *dinfo, hipStream_t s, char *ws = nullptr, bool lookahead = true,
                          bool own_gemm = false, bool pp = false) {
  typedef typename C2<T>::t cT;
  constexpr int PW = PanelW<T>::W;
  const dim3 b256(256);
  if (n <= 0 || batch <= 0) return;
  if (n >= (1 << 22)) qpms_hip_die(__FILE__, __LINE__, __func__, "order %d above the pivot search's 2^22 rows", n);
  const int Gmax = PanelWS<T>::sets(n);
  const long sW = (long)PanelWS<T>::bytes(n);
  const long setb = (long)PanelWS<T>::set_bytes();
  const long sWc = sW / (long)sizeof(cT);   // the workspace stride in complex entries
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
  const dim3 bts(TS_THREADS);
  ts_prepare<T>();
  char *bufA = ws, *bufB = ws + (long)Gmax * setb;   // two candidate buffers per matrix
  char *fin0 = ws + 2 * (long)Gmax * setb;           // per-strip slots of the last level's output
  const long finb = (long)PanelWS<T>::fin_bytes();
  // The strips of the outer block [k0, k0 + nb): tournament, interchanges on
  // the block's own columns, L21, then the block's remaining columns by TRSM
  // and the rank-w update. Touches only the block's columns (rows >= k0).
  if (pp && pp_ws_bytes(n) > 2 * (long)Gmax * setb)
    qpms_hip_die(__FILE__, __LINE__, __func__, "partial-pivoting workspace does not fit (order %d)", n);
  if (pp) QH_LAUNCH(k_pp_init, dim3(batch), dim3(1), 0, s, ws, sW);
  // partial pivoting: the strip in place, column by column, then its slot
  auto pp_strip = [&](int c0, int w, hipStream_t st) {
    const int G = (n - c0 + PP_ROWS - 1) / PP_ROWS;
    for (int j = 0; j <= w; ++j)
      QH_LAUNCH(k_pp_step<T>, dim3(G, 1, batch), b256, 0, st, n, c0, w, j, A, lda, sA, ws, sW, ipiv, sP, dinfo);
    QH_LAUNCH(k_pp_finish<T>, dim3(1, 1, batch), b256, PP_FIN_LDS, st, n, c0, w, A, lda, sA, ws, sW,
              (long)(fin0 - ws) + (long)((c0 % LU_NB) / PW) * finb, perm, sP);
  };
  auto block_strips = [&](int k0, int nb, hipStream_t st, bool wcu) {
    constexpr int F = ts_fan<T>();
    const WholeCu<T> &W = whole_cu<T>();
    const size_t l_local = wcu ? W.local : ts_lds<T>(), l_merge = wcu ? W.merge : ts_lds<T>();
    const size_t l_apply = wcu ? W.apply : 0, l_asg = wcu ? W.gemm_assign : 0, l_sub = wcu ? W.gemm_sub : 0;
    for (int j0 = 0; j0 < nb; j0 += PW) {
      const int w = min(PW, nb - j0), c0 = k0 + j0;
      char *fin = fin0 + (long)(j0 / PW) * finb;
      if (pp) {   // the strip factored in place (L21 included); interchanges on the block's other columns
        pp_strip(c0, w, st);
        if (c0 > k0)
          QH_LAUNCH(k_ts_apply<T>, dim3((c0 - k0 + AP_COLS - 1) / AP_COLS, 1, batch), b256, l_apply, st, n, c0, w, A,
                    lda, sA, fin, sW, k0, c0);
        if (k0 + nb > c0 + w)
          QH_LAUNCH(k_ts_apply<T>, dim3((k0 + nb - c0 - w + AP_COLS - 1) / AP_COLS, 1, batch), b256, l_apply, st, n,
                    c0, w, A, lda, sA, fin, sW, c0 + w, k0 + nb);
      } else {
      const int G = (n - c0 + TsRows<T>::R - 1) / TsRows<T>::R;
      QH_LAUNCH(k_ts_local<T>, dim3(G, 1, batch), bts, l_local, st, n, c0, w, A, lda, sA, ws, sW, setb,
                         (long)(fin - ws), G == 1 ? 1 : 0, ipiv, sP, perm, dinfo);
      char *cur = bufA, *nxt = bufB;
      for (int nsets = G; nsets > 1; nsets = (nsets + F - 1) / F) {
        const int nout = (nsets + F - 1) / F;
        QH_LAUNCH(k_ts_merge<T>, dim3(nout, 1, batch), bts, l_merge, st, c0, w, nsets, cur, nxt, sW,
                           setb, fin, nout == 1 ? 1 : 0, ipiv, sP, perm, dinfo);
        char *tmp = cur; cur = nxt; nxt = tmp;
      }
      QH_LAUNCH(k_ts_apply<T>, dim3((nb + AP_COLS - 1) / AP_COLS, 1, batch), b256, l_apply, st, n, c0, w, A,
                lda, sA, fin, sW, k0, k0 + nb);
      }
      const cT *linv = reinterpret_cast<const cT *>(fin + PanelWS<T>::fin_inv_off()), *uinv = linv + PW * PW;
      const int mr = n - c0 - w;
      const int rest = k0 + nb - (c0 + w);
      if (wcu) {   // single matrix, whole-CU launches (k_gemm_cu: several tiles per CU, K <= PW)
        const dim3 bcu(gcu_q<T>() * 256);
        auto gcu_grid = [](int M, int N) {
          const int nt = ((M + GT - 1) / GT) * ((N + GT - 1) / GT);
          return dim3(std::min((nt + gcu_q<T>() - 1) / gcu_q<T>(), 4096));
        };
        if (mr > 0 && !pp)   // L21 = A21 U11^{-1}
          QH_LAUNCH((k_gemm_cu<T, true>), gcu_grid(mr, w), bcu, l_asg, st, mr, w, w, A + (long)(c0 + w) * lda + c0,
                    lda, uinv, (long)PW, A + (long)(c0 + w) * lda + c0, lda);
        if (rest > 0) {   // the block's next columns: U = L11^{-1} A12, then A22 -= L21 U
          QH_LAUNCH((k_gemm_cu<T, true>), gcu_grid(w, rest), bcu, l_asg, st, w, rest, w, linv, (long)PW,
                    A + (long)c0 * lda + c0 + w, lda, A + (long)c0 * lda + c0 + w, lda);
          if (mr > 0)
            QH_LAUNCH((k_gemm_cu<T, false>), gcu_grid(mr, rest), bcu, l_sub, st, mr, rest, w,
                      A + (long)(c0 + w) * lda + c0, lda, A + (long)c0 * lda + c0 + w, lda,
                      A + (long)(c0 + w) * lda + c0 + w, lda);
        }
        continue;
      }
      if (mr > 0 && !pp)   // L21 = A21 U11^{-1}
        QH_LAUNCH((k_gemm_sub<T, true>), dim3(1, (mr + GT - 1) / GT, batch), b256, l_asg, st, mr, w, w,
                  A + (long)(c0 + w) * lda + c0, lda, sA, uinv, (long)PW, sWc, A + (long)(c0 + w) * lda + c0, lda,
                  sA);
      if (rest > 0) {   // the block's next columns: U = L11^{-1} A12, then A22 -= L21 U
        QH_LAUNCH((k_gemm_sub<T, true>), dim3((rest + GT - 1) / GT, 1, batch), b256, l_asg, st, w, rest, w, linv,
                  (long)PW, sWc, A + (long)c0 * lda + c0 + w, lda, sA, A + (long)c0 * lda + c0 + w, lda, sA);
        if (mr > 0)
          QH_LAUNCH(k_gemm_sub<T>, dim3((rest + GT - 1) / GT, (mr + GT - 1) / GT, batch), b256, l_sub, st, mr,
                             rest, w, A + (long)(c0 + w) * lda + c0, lda, sA, A + (long)c0 * lda + c0 + w, lda, sA,
                             A + (long)(c0 + w) * lda + c0 + w, lda, sA);
      }
    }
  };
  // Lookahead: once the trailing update has reached block k+1's columns, the
  // rest of it runs on the side stream (masked off a few CUs) while block
  // k+1's strips run here; the side stream's update is enqueued first, so it
  // starts at once, and the strips find the reserved CUs free. Block k+1's
  // interchanges are applied to the columns outside it afterwards (the strips
  // touch only their own block), after the side stream's update. Single
  // matrices only: a batch's strips already fill the device (batch x tens of
  // workgroups), and as whole-CU workgroups on the reserved CUs they would
  // queue (a batch of 8 at order 9216: 274 us per tournament launch).
  const bool la = lookahead && use_lookahead() && batch == 1 && n > 2 * LU_NB && !stream_capturing(s);
  hipStream_t side = nullptr;
  bool masked = false;
  LaCtx *cx = la ? la_for(s, &side, &masked) : nullptr;
  if (perm) QH_LAUNCH(k_iota, dim3((n + 255) / 256, 1, batch), b256, 0, s, n, perm, sP);
  block_strips(0, min(LU_NB, n), s, false);
  for (int k0 = 0; k0 < n; k0 += LU_NB) {
    const int nb = min(LU_NB, n - k0);
    if (n > nb)
      QH_LAUNCH(k_ts_apply_defer<T>, dim3((n + AP_COLS - 1) / AP_COLS, 1, batch), b256, 0, s, n, k0,
                         (nb + PW - 1) / PW, A, lda, sA, fin0, finb, sW, k0, k0 + nb);
    const int nr = n - k0 - nb;
    if (nr > 0) {
      // U12 = L11^{-1} A12, PW rows at a time (each strip's L^{-1} from its slot)
      for (int j0 = 0; j0 < nb; j0 += PW) {
        const int w = min(PW, nb - j0), c0 = k0 + j0;
        const cT *linv = reinterpret_cast<const cT *>(fin0 + (long)(j0 / PW) * finb + PanelWS<T>::fin_inv_off());
        QH_LAUNCH((k_gemm_sub<T, true>), dim3((nr + GT - 1) / GT, 1, batch), b256, 0, s, w, nr, w, linv, (long)PW,
                  sWc, A + (long)c0 * lda + k0 + nb, lda, sA, A + (long)c0 * lda + k0 + nb, lda, sA);
        const int below = nb - j0 - w;
        if (below > 0)
          QH_LAUNCH(k_gemm_sub<T>, dim3((nr + GT - 1) / GT, (below + GT - 1) / GT, batch), b256, 0, s, below,
                             nr, w, A + (long)(c0 + w) * lda + c0, lda, sA, A + (long)c0 * lda + k0 + nb, lda, sA,
                             A + (long)(c0 + w) * lda + k0 + nb, lda, sA);
      }
      // A22 -= L21 U12: the next block's columns first, then the rest and the next block's strips
      const int nb1 = min(LU_NB, nr);
      trailing_gemm<T>(nr, nb1, nb, batch, A + (long)(k0 + nb) * lda + k0, lda, sA, A + (long)k0 * lda + k0 + nb, lda,
                       sA, A + (long)(k0 + nb) * lda + k0 + nb, lda, sA, s, own_gemm);
      const bool split = la && nr > nb1;
      if (nr > nb1) {
        hipStream_t sg = split ? side : s;
        if (split) {
          HIPCHK(hipEventRecord(cx->go, s));
          HIPCHK(hipStreamWaitEvent(sg, cx->go, 0));
        }
        trailing_gemm<T>(nr, nr - nb1, nb, batch, A + (long)(k0 + nb) * lda + k0, lda, sA,
                         A + (long)k0 * lda + k0 + nb + nb1, lda, sA, A + (long)(k0 + nb) * lda + k0 + nb + nb1, lda,
                         sA, sg, own_gemm);
        if (split) HIPCHK(hipEventRecord(cx->done, sg));
      }
      block_strips(k0 + nb, nb1, s, split && masked);
      if (split) HIPCHK(hipStreamWaitEvent(s, cx->done, 0));
    }
  }
  HIPCHK(hipGetLastError());
}

// Triangular solves with few right-hand sides (nrhs <= TRV_R), in blocks of
// TRV_W rows, on the permuted right-hand side W (n x nrhs, row-major):
//   k_trsv_wave: one wave per matrix solves the w x w (w <= 64) diagonal
//                block (unit lower, forward; or upper, backward), lane i
//                owning row r0 + i, with no barrier per step: step s
//                finalises x_j (j = s forward, w - 1 - s backward, where it
//                takes the diagonal's reciprocal computed up front), every
//                lane reads it with readlane, and the rows still open
//                subtract their column entry times it; the block sits in LDS
//                column-major (lane i reads entry (i, j) at j TRV_W + i, no
//                bank conflict) and each lane loads its entries PF steps
//                ahead;
//   k_trsv_update: the rows outside the block (below it forward, above it
//                backward) subtract the block's columns times its solution,
//                one thread per row walking its row contiguously.
constexpr int TRV_R = 4, TRV_W = 64;
__device__ __forceinline__ float trv_readlane(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double trv_readlane(double v, int j) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), j), hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <class T, bool FWD>
__global__ __launch_bounds__(TRV_W) void k_trsv_wave(int r0, int w, int nrhs, const typename C2<T>::t *LU, long lda,
                                                     long sA, typename C2<T>::t *W, long sW) {
  typedef typename C2<T>::t cT;
  constexpr int PF = 4;
  __shared__ cT sL[TRV_W * TRV_W];   // sL[j TRV_W + i] = LU(r0 + i, r0 + j)
  LU += (long)blockIdx.z * sA;
  W += (long)blockIdx.z * sW;
  const int i = threadIdx.x;
  const bool mine = i < w;
  if (mine) {
    const cT *row = LU + (long)(r0 + i) * lda + r0;
    for (int j = 0; j < w; ++j) sL[j * TRV_W + i] = row[j];
  }
  cT y[TRV_R];
#pragma unroll
  for (int c = 0; c < TRV_R; ++c) y[c] = (mine && c < nrhs) ? W[(long)(r0 + i) * nrhs + c] : cT{0, 0};
  __syncthreads();
  const cT dinv = (!FWD && mine) ? cdiv<T>(cT{1, 0}, sL[i * TRV_W + i]) : cT{1, 0};
  auto col = [&](int s) { return FWD ? s : w - 1 - s; };
  cT buf[PF];
#pragma unroll
  for (int k = 0; k < PF; ++k) buf[k] = (mine && k < w) ? sL[col(k) * TRV_W + i] : cT{0, 0};
  for (int s0 = 0; s0 < w; s0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int s = s0 + k;
      if (s >= w) continue;   // uniform
      const int j = col(s);
      const cT l = buf[k];
      buf[k] = (mine && s + PF < w) ? sL[col(s + PF) * TRV_W + i] : cT{0, 0};
      if (!FWD && i == j) {
#pragma unroll
        for (int c = 0; c < TRV_R; ++c) y[c] = cmul<T>(y[c], dinv);
      }
      cT x[TRV_R];
#pragma unroll
      for (int c = 0; c < TRV_R; ++c) x[c] = cT{trv_readlane(y[c].x, j), trv_readlane(y[c].y, j)};
      if (mine && (FWD ? i > j : i < j)) {
#pragma unroll
        for (int c = 0; c < TRV_R; ++c)
          y[c] = cT{y[c].x - (l.x * x[c].x - l.y * x[c].y), y[c].y - (l.x * x[c].y + l.y * x[c].x)};
      }
    }
  }
  if (mine)
#pragma unroll
    for (int c = 0; c < TRV_R; ++c)
      if (c < nrhs) W[(long)(r0 + i) * nrhs + c] = y[c];
}

template <class T>
__global__ __launch_bounds__(64) void k_trsv_update(int i0, int nrows, int r0, int w, int nrhs,
                                                    const typename C2<T>::t *LU, long lda, long sA,
                                                    typename C2<T>::t *W, long sW) {
  typedef typename C2<T>::t cT;
  __shared__ cT sx[TRV_W * TRV_R];
  LU += (long)blockIdx.z * sA;
  W += (long)blockIdx.z * sW;
  for (int e = threadIdx.x; e < w * nrhs; e += 64) sx[(e / nrhs) * TRV_R + e % nrhs] = W[(long)r0 * nrhs + e];
  __syncthreads();
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= nrows) return;
  const int i = i0 + r;
  const cT *row = LU + (long)i * lda + r0;
  cT acc[TRV_R];
#pragma unroll
  for (int c = 0; c < TRV_R; ++c) acc[c] = cT{0, 0};
  for (int j = 0; j < w; ++j) {
    const cT l = row[j];
#pragma unroll
    for (int c = 0; c < TRV_R; ++c) {
      const cT x = sx[j * TRV_R + c];
      acc[c] = cT{acc[c].x + (l.x * x.x - l.y * x.y), acc[c].y + (l.x * x.y + l.y * x.x)};
    }
  }
#pragma unroll
  for (int c = 0; c < TRV_R; ++c)
    if (c < nrhs) {
      cT &y = W[(long)i * nrhs + c];
      y = cT{y.x - acc[c].x, y.y - acc[c].y};
    }
}

// Many right-hand sides: the triangle in diagonal blocks of TsbRows rows
// (256 FP32, 128 FP64), each solved by one launch of k_trsm_blk, and the rows
// beyond the block updated by one GEMM with K = the block height: an eighth
// (a quarter) of the dependent launches of 32-row blocks. k_trsm_blk takes a
// slab of 32 right-hand-side columns per workgroup into LDS and walks the
// block's 32-row sub-blocks: the sub-block's triangle and the panel of the
// triangle beside it (the block's other rows, those columns) are staged in
// LDS, every load issued before the substitution, which runs in registers,
// one thread per column; then all threads subtract the panel times the
// solved rows from the other rows, reading only LDS.
template <class T> struct TsbRows;
template <> struct TsbRows<float> { static constexpr int R = 256; };
template <> struct TsbRows<double> { static constexpr int R = 128; };
constexpr int TSB_SS = 33, TSB_THREADS = 512;   // LDS row stride of the triangle's columns; workgroup size
constexpr int TSB_SW = 8, TSB_BS = TSB_SW + 1;  // right-hand-side columns per workgroup; slab row stride
template <class T> constexpr size_t tsb_lds() {
  return (size_t)(TsbRows<T>::R * (TSB_BS + TSB_SS)) * sizeof(typename C2<T>::t);   // slab + (sub-block and panel)
}
template <class T, bool LOWER>
__global__ __launch_bounds__(TSB_THREADS) void k_trsm_blk(int w, int ncols, const typename C2<T>::t *L, long ldl,
                                                          long sL, typename C2<T>::t *B, long ldb, long sB) {
  typedef typename C2<T>::t cT;
  constexpr int TB = TsbRows<T>::R, NPL = TB * 32 / TSB_THREADS;   // staged triangle entries per thread
  extern __shared__ __align__(16) char smem_raw[];
  cT *sb = reinterpret_cast<cT *>(smem_raw);   // [TB][TSB_BS] the slab
  cT *sp = sb + TB * TSB_BS;                    // [TB][TSB_SS] rows of the triangle, columns of the sub-block
  L += (long)blockIdx.z * sL;
  B += (long)blockIdx.z * sB;
  const int t = threadIdx.x, c0 = blockIdx.x * TSB_SW, nc = min(TSB_SW, ncols - c0);
  for (int e = t; e < w * TSB_SW; e += TSB_THREADS) {
    const int r = e / TSB_SW, c = e % TSB_SW;
    sb[r * TSB_BS + c] = (c < nc) ? B[(long)r * ldb + c0 + c] : cT{0, 0};
  }
  const int nsub = (w + 31) / 32;
  for (int q = 0; q < nsub; ++q) {
    const int j0 = (LOWER ? q : nsub - 1 - q) * 32, jw = min(32, w - j0);
    // columns j0..j0+jw of every row of the block (the sub-block's triangle
    // and the panel), loads first
    {
      cT v[NPL];
#pragma unroll
      for (int u = 0; u < NPL; ++u) {
        const int e = t + TSB_THREADS * u, r = e >> 5, k = e & 31;
        v[u] = (r < w && k < jw) ? L[(long)r * ldl + j0 + k] : cT{0, 0};
      }
      __syncthreads();   // the previous sub-block's reads of sp are done
#pragma unroll
      for (int u = 0; u < NPL; ++u) {
        const int e = t + TSB_THREADS * u;
        sp[(e >> 5) * TSB_SS + (e & 31)] = v[u];
      }
    }
    __syncthreads();
    if (t < TSB_SW) {   // substitution on column t of the slab, in registers
      cT x[32];
#pragma unroll
      for (int r = 0; r < 32; ++r) x[r] = (r < jw) ? sb[(j0 + r) * TSB_BS + t] : cT{0, 0};
      // each row's sum in four interleaved partial sums (shorter dependent chains)
      if (LOWER) {
#pragma unroll
        for (int r = 1; r < 32; ++r) {
          if (r < jw) {
            cT a[4] = {cT{0, 0}, cT{0, 0}, cT{0, 0}, cT{0, 0}};
#pragma unroll
            for (int k = 0; k < r; ++k) {
              const cT l = sp[(j0 + r) * TSB_SS + k];
              a[k & 3] = cT{a[k & 3].x + (l.x * x[k].x - l.y * x[k].y), a[k & 3].y + (l.x * x[k].y + l.y * x[k].x)};
            }
            x[r] = cT{x[r].x - ((a[0].x + a[1].x) + (a[2].x + a[3].x)), x[r].y - ((a[0].y + a[1].y) + (a[2].y + a[3].y))};
          }
        }
      } else {
#pragma unroll
        for (int r = 31; r >= 0; --r) {
          if (r < jw) {
            cT a[4] = {cT{0, 0}, cT{0, 0}, cT{0, 0}, cT{0, 0}};
#pragma unroll
            for (int k = r + 1; k < 32; ++k) {
              const cT u = sp[(j0 + r) * TSB_SS + k];
              a[k & 3] = cT{a[k & 3].x + (u.x * x[k].x - u.y * x[k].y), a[k & 3].y + (u.x * x[k].y + u.y * x[k].x)};
            }
            const cT v = cT{x[r].x - ((a[0].x + a[1].x) + (a[2].x + a[3].x)), x[r].y - ((a[0].y + a[1].y) + (a[2].y + a[3].y))};
            x[r] = cdiv<T>(v, sp[(j0 + r) * TSB_SS + r]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 32; ++r)
        if (r < jw) sb[(j0 + r) * TSB_BS + t] = x[r];
    }
    __syncthreads();
    // the block's other rows: below the sub-block (L), above it (U)
    const int u0 = LOWER ? j0 + jw : 0, u1 = LOWER ? w : j0, c = t % TSB_SW;
    if (u1 > u0) {
      cT x[32];
#pragma unroll
      for (int k = 0; k < 32; ++k) x[k] = (k < jw) ? sb[(j0 + k) * TSB_BS + c] : cT{0, 0};
      if constexpr (sizeof(T) == 4) {
        // two rows per step, each in two partial sums
        constexpr int RS = TSB_THREADS / TSB_SW;
        for (int r = u0 + t / TSB_SW; r < u1; r += 2 * RS) {
          const int r2 = r + RS;
          const bool two = r2 < u1;
          cT a0 = cT{0, 0}, a1 = cT{0, 0}, b0 = cT{0, 0}, b1 = cT{0, 0};
#pragma unroll
          for (int k = 0; k < 32; k += 2) {
            const cT l0 = sp[r * TSB_SS + k], l1 = sp[r * TSB_SS + k + 1];
            const cT m0 = two ? sp[r2 * TSB_SS + k] : cT{0, 0}, m1 = two ? sp[r2 * TSB_SS + k + 1] : cT{0, 0};
            a0 = cT{a0.x + (l0.x * x[k].x - l0.y * x[k].y), a0.y + (l0.x * x[k].y + l0.y * x[k].x)};
            a1 = cT{a1.x + (l1.x * x[k + 1].x - l1.y * x[k + 1].y), a1.y + (l1.x * x[k + 1].y + l1.y * x[k + 1].x)};
            b0 = cT{b0.x + (m0.x * x[k].x - m0.y * x[k].y), b0.y + (m0.x * x[k].y + m0.y * x[k].x)};
            b1 = cT{b1.x + (m1.x * x[k + 1].x - m1.y * x[k + 1].y), b1.y + (m1.x * x[k + 1].y + m1.y * x[k + 1].x)};
          }
          cT &y = sb[r * TSB_BS + c];
          y = cT{y.x - (a0.x + a1.x), y.y - (a0.y + a1.y)};
          if (two) {
            cT &y2 = sb[r2 * TSB_BS + c];
            y2 = cT{y2.x - (b0.x + b1.x), y2.y - (b0.y + b1.y)};
          }
        }
      } else {   // FP64: one row per step (the two-row form exceeds the registers)
        for (int r = u0 + t / TSB_SW; r < u1; r += TSB_THREADS / TSB_SW) {
          cT acc = cT{0, 0};
#pragma unroll
          for (int k = 0; k < 32; ++k) {
            const cT l = sp[r * TSB_SS + k];
            acc = cT{acc.x + (l.x * x[k].x - l.y * x[k].y), acc.y + (l.x * x[k].y + l.y * x[k].x)};
          }
          cT &y = sb[r * TSB_BS + c];
          y = cT{y.x - acc.x, y.y - acc.y};
        }
      }
    }
  }
  __syncthreads();
  for (int e = t; e < w * TSB_SW; e += TSB_THREADS) {
    const int r = e / TSB_SW, c = e % TSB_SW;
    if (c < nc) B[(long)r * ldb + c0 + c] = sb[r * TSB_BS + c];
  }
}
template <class T>
static void tsb_prepare() {
  static std::once_flag once;
  std::call_once(once, [] {
    HIPCHK(hipFuncSetAttribute((const void *)k_trsm_blk<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)tsb_lds<T>()));
    HIPCHK(hipFuncSetAttribute((const void *)k_trsm_blk<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)tsb_lds<T>()));
  });
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
  QH_LAUNCH(k_gather_rows<T>, dim3(qh_stride_blocks(tot), 1, batch), b256, 0, s, n, nrhs, perm, sP,
                     (const cT *)B, ldb, sB, work, (long)nrhs, sW);
  if (nrhs <= TRV_R) {
    const int nblk = (n + TRV_W - 1) / TRV_W;
    for (int bi = 0; bi < nblk; ++bi) {
      const int r0 = bi * TRV_W, w = min(TRV_W, n - r0), rem = n - r0 - w;
      QH_LAUNCH((k_trsv_wave<T, true>), dim3(1, 1, batch), dim3(TRV_W), 0, s, r0, w, nrhs, LU, lda, sA, work, sW);
      if (rem > 0)
        QH_LAUNCH(k_trsv_update<T>, dim3((rem + 63) / 64, 1, batch), dim3(64), 0, s, r0 + w, rem, r0, w, nrhs,
                           LU, lda, sA, work, sW);
    }
    for (int bi = nblk - 1; bi >= 0; --bi) {
      const int r0 = bi * TRV_W, w = min(TRV_W, n - r0);
      QH_LAUNCH((k_trsv_wave<T, false>), dim3(1, 1, batch), dim3(TRV_W), 0, s, r0, w, nrhs, LU, lda, sA, work, sW);
      if (r0 > 0)
        QH_LAUNCH(k_trsv_update<T>, dim3((r0 + 63) / 64, 1, batch), dim3(64), 0, s, 0, r0, r0, w, nrhs, LU, lda,
                           sA, work, sW);
    }
    QH_LAUNCH(k_copy_rows<T>, dim3(qh_stride_blocks(tot), 1, batch), b256, 0, s, n, nrhs,
                       (const cT *)work, (long)nrhs, sW, B, ldb, sB);
    HIPCHK(hipGetLastError());
    return;
  }
  tsb_prepare<T>();
  const dim3 gs((nrhs + TSB_SW - 1) / TSB_SW, 1, batch);
  for (int r0 = 0; r0 < n; r0 += TsbRows<T>::R) {
    const int w = min(TsbRows<T>::R, n - r0);
    QH_LAUNCH((k_trsm_blk<T, true>), gs, dim3(TSB_THREADS), tsb_lds<T>(), s, w, nrhs, LU + (long)r0 * lda + r0, lda, sA,
              work + (long)r0 * nrhs, (long)nrhs, sW);
    const int rem = n - r0 - w;
    if (rem > 0)
      QH_LAUNCH(k_gemm_sub<T>, dim3((nrhs + GT - 1) / GT, (rem + GT - 1) / GT, batch), b256, 0, s, rem, nrhs, w,
                         LU + (long)(r0 + w) * lda + r0, lda, sA, (const cT *)(work + (long)r0 * nrhs), (long)nrhs, sW,
                         work + (long)(r0 + w) * nrhs, (long)nrhs, sW);
  }
  const int nblk = (n + TsbRows<T>::R - 1) / TsbRows<T>::R;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int r0 = bi * TsbRows<T>::R, w = min(TsbRows<T>::R, n - r0);
    QH_LAUNCH((k_trsm_blk<T, false>), gs, dim3(TSB_THREADS), tsb_lds<T>(), s, w, nrhs, LU + (long)r0 * lda + r0, lda, sA,
              work + (long)r0 * nrhs, (long)nrhs, sW);
    if (r0 > 0)
      QH_LAUNCH(k_gemm_sub<T>, dim3((nrhs + GT - 1) / GT, (r0 + GT - 1) / GT, batch), b256, 0, s, r0, nrhs, w,
                         LU + r0, lda, sA, (const cT *)(work + (long)r0 * nrhs), (long)nrhs, sW, work, (long)nrhs, sW);
  }
  // copy back (B may have a different leading dimension), one launch for the batch
  QH_LAUNCH(k_copy_rows<T>, dim3(qh_stride_blocks(tot), 1, batch), b256, 0, s, n, nrhs,
                     (const cT *)work, (long)nrhs, sW, B, ldb, sB);
  HIPCHK(hipGetLastError());
}

extern "C" int qpms_hip_zgetrf_dev(size_t n, _Complex double *A_dev, int *ipiv_dev, void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  int *dinfo = (int *)qh_dmalloc(sizeof(int));
  HIPCHK(hipMemsetAsync(dinfo, 0, sizeof(int), s));
  getrf_batched<double>((int)n, 1, (v2d *)A_dev, (long)n, 0, ipiv_dev, 0, nullptr, dinfo, s, nullptr, true, false,
                        /*pp=*/true);
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  qh_dfree(dinfo);
  return info;
}

// ---------------------------------------------------------------------------
// Whole-factorisation timing for the GEMM autotuner: FP32 LUs of a
// well-conditioned pseudo-random matrix (entries in [-1/2, 1/2]^2 from a hash
// of the indices, 3 sqrt(n) added on the diagonal), with lookahead, as the
// serving loop and the LU entry points run them, on a stream of their own.
// ---------------------------------------------------------------------------
__global__ void k_tune_fill(long n, v2f *A) {
  const long tot = n * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    unsigned long long h = (unsigned long long)e * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    const float re = (float)(h & 0xFFFFFF) * (1.f / 16777216.f) - 0.5f;
    const float im = (float)((h >> 24) & 0xFFFFFF) * (1.f / 16777216.f) - 0.5f;
    A[e] = v2f{re + ((e / n == e % n) ? 3.f * sqrtf((float)n) : 0.f), im};
  }
}

namespace {
struct TuneCtx {
  v2f *A;
  char *ws;
  int *ipiv, *info;
  hipStream_t s;
  hipEvent_t e0, e1;
  int nmax;
};
float tune_lu(int n, void *p) {
  TuneCtx &c = *static_cast<TuneCtx *>(p);
  if (n > c.nmax) return -1.f;
  float best = -1.f;
  for (int rep = 0; rep < 2; ++rep) {   // the first run warms the shapes' kernels up
    QH_LAUNCH(k_tune_fill, dim3(qh_stride_blocks((long)n * n)), dim3(256), 0, c.s, (long)n, c.A);
    HIPCHK(hipEventRecord(c.e0, c.s));
    getrf_batched<float>(n, 1, c.A, n, 0, c.ipiv, n, nullptr, c.info, c.s, c.ws, true);
    HIPCHK(hipEventRecord(c.e1, c.s));
    HIPCHK(hipEventSynchronize(c.e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c.e0, c.e1));
    if (rep) best = ms;
  }
  return best;
}
}  // namespace

// Tunes the single-matrix classes by whole factorisations (largest order
// that fits the free memory, at most 16384; the class above keeps its GEMM
// timing). Returns the classes tuned.
extern "C" int qh_autotune_lu(float *report) {
  TuneCtx c{};
  HIPCHK(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));   // released with its LU state at the end
  for (int n = 16384; n >= 2048 && !c.A; n /= 2) {
    void *a = nullptr, *w = nullptr;
    if (hipMalloc(&a, (size_t)n * n * 8) != hipSuccess ||
        hipMalloc(&w, getrf_ws_bytes<float>(n, 1) + 8 * (size_t)n + 256) != hipSuccess) {
      (void)hipGetLastError();
      if (a) HIPCHK(hipFree(a));
      if (w) HIPCHK(hipFree(w));
      continue;
    }
    c.A = (v2f *)a;
    c.ws = (char *)w;
    c.nmax = n;
  }
  if (!c.A) {
    HIPCHK(hipStreamDestroy(c.s));
    return 0;
  }
  c.ipiv = (int *)(c.ws + getrf_ws_bytes<float>(c.nmax, 1));
  c.info = c.ipiv + c.nmax;
  HIPCHK(hipEventCreate(&c.e0));
  HIPCHK(hipEventCreate(&c.e1));
  const int tuned = qh_autotune_lu_classes(tune_lu, &c, report);
  HIPCHK(hipStreamSynchronize(c.s));
  HIPCHK(hipEventDestroy(c.e0));
  HIPCHK(hipEventDestroy(c.e1));
  HIPCHK(hipFree(c.A));
  HIPCHK(hipFree(c.ws));
  qh_stream_release(c.s);
  HIPCHK(hipStreamDestroy(c.s));
  return tuned;
}

// ---------------------------------------------------------------------------
// Allocation-free variant of qpms_hip_zgetrf_dev for callers that cannot
// allocate device memory (code under graph capture, serving loops that
// forbid hipMalloc). Its scratch -- the info word and the strip workspace
// -- is the file-scope array below; the GEMMs are the
// hand-written kernel (the library GEMM may allocate) and there is no
// lookahead side stream. The host never names the array in a kernel
// argument or copy: it uses the device address hipGetSymbolAddress returned
// for the current device at initialisation.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEV = 64;
static char *g_scratch_ptr[NOALLOC_MAX_DEV];   // device address of `scratch`, per device

// Called by the library's initialisation with the device current (once per
// device); records the device address of the scratch array.
extern "C" void qh_noalloc_scratch_init(void) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEV)
    qpms_hip_die(__FILE__, __LINE__, __func__, "device %d outside the scratch table", dev);
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[dev] = (char *)p;
}

// Largest order the scratch array takes (FP64 tournament workspace after a
// 256-byte info slot).
extern "C" size_t qpms_hip_zgetrf_noalloc_max_n(void) {
  size_t n = 256;
  while (256 + PanelWS<double>::bytes((int)(n + 256)) <= NOALLOC_SCRATCH_BYTES) n += 256;
  return n;
}

extern "C" int qpms_hip_zgetrf_dev_noalloc(size_t n, _Complex double *A_dev, int *ipiv_dev, void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  if (n == 0) return 0;
  const size_t need = 256 + PanelWS<double>::bytes((int)(n < (1u << 22) ? n : (1u << 22)));
  if (n >= (1u << 22) || need > NOALLOC_SCRATCH_BYTES) {
    fprintf(stderr, "qpms_hip_zgetrf_dev_noalloc: order %zu needs %zu bytes of scratch, above the %zu-byte array "
            "(largest order %zu)\n", n, need, NOALLOC_SCRATCH_BYTES, qpms_hip_zgetrf_noalloc_max_n());
    return -1;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  char *base = (dev >= 0 && dev < NOALLOC_MAX_DEV) ? g_scratch_ptr[dev] : nullptr;
  if (!base) qpms_hip_die(__FILE__, __LINE__, __func__, "scratch of device %d not initialised", dev);
  int *dinfo = (int *)base;
  qh_zero(dinfo, 256, s);   // a kernel: the info word must start at zero on every call
  getrf_batched<double>((int)n, 1, (v2d *)A_dev, (long)n, 0, ipiv_dev, 0, nullptr, dinfo, s, base + 256,
                        /*lookahead=*/false, /*own_gemm=*/true, /*pp=*/true);
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return info;
}

extern "C" int qpms_hip_zgetrs_dev(size_t n, size_t nrhs, const _Complex double *LU_dev, const int *ipiv_dev,
                                   _Complex double *B_dev, void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  int *perm = (int *)qh_dmalloc(n * sizeof(int));
  v2d *work = (v2d *)qh_dmalloc(n * nrhs * sizeof(v2d));
  QH_LAUNCH(k_perm_from_ipiv, dim3(1), dim3(64), 0, s, (int)n, ipiv_dev, perm);
  getrs_batched<double>((int)n, (int)nrhs, 1, (const v2d *)LU_dev, (long)n, 0, perm, 0, (v2d *)B_dev, (long)nrhs, 0,
                        work, 0, s);
  HIPCHK(hipStreamSynchronize(s));
  qh_dfree(perm);
  qh_dfree(work);
  return 0;
}

extern "C" int qh_launch_block_tw(size_t nr, size_t L, const double *T, const double *W, double *M, size_t col0,
                                  void *stream) {
  QH_LAUNCH(k_block_tw, dim3((unsigned)((L + 255) / 256), (unsigned)nr), dim3(256), 0,
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
// Caller-workspace entry points: their callers (the serving loop's lanes)
// already overlap work across several streams, so no lookahead side stream.
extern "C" int qh_getrf_batched_ws_f64(int n, int batch, double *A, long sA, int *ipiv, int *perm, int *info, void *ws,
                                       void *stream) {
  getrf_batched<double>(n, batch, (v2d *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream), (char *)ws,
                        false);
  return 0;
}
extern "C" int qh_getrf_batched_ws_f32(int n, int batch, float *A, long sA, int *ipiv, int *perm, int *info, void *ws,
                                       void *stream) {
  getrf_batched<float>(n, batch, (v2f *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream), (char *)ws,
                       false);
  return 0;
}
// The same with lookahead when `lookahead` (large orders): the bulk of each
// trailing GEMM runs on a side stream of `stream` while the next block's
// strips run on `stream` (the side stream belongs to `stream` until
// released).
extern "C" int qh_getrf_batched_wsla_f32(int n, int batch, float *A, long sA, int *ipiv, int *perm, int *info,
                                         void *ws, int lookahead, void *stream) {
  getrf_batched<float>(n, batch, (v2f *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream), (char *)ws,
                       lookahead != 0);
  return 0;
}
extern "C" int qh_getrf_batched_wsla_f64(int n, int batch, double *A, long sA, int *ipiv, int *perm, int *info,
                                         void *ws, int lookahead, void *stream) {
  getrf_batched<double>(n, batch, (v2d *)A, n, sA, ipiv, n, perm, info, (hipStream_t)qh_stream(stream), (char *)ws,
                        lookahead != 0);
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
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  if (N <= 4) {   // few right-hand sides: stream A once, a wave per row
    const dim3 g((unsigned)((M + 3) / 4), (unsigned)batch);
    hipStream_t s = (hipStream_t)qh_stream(stream);
#define QH_GV(NR) QH_LAUNCH(k_gemv_sub_f64<NR>, g, dim3(256), 0, s, M, K, (const v2d *)A, lda, sA, \
                                     (const v2d *)B, ldb, sB, (v2d *)C, ldc, sC)
    switch (N) { case 1: QH_GV(1); break; case 2: QH_GV(2); break; case 3: QH_GV(3); break; default: QH_GV(4); }
#undef QH_GV
    HIPCHK(hipGetLastError());
    return 0;
  }
  // more right-hand sides (the refinement residuals of Beyn nodes, N = l):
  // the library GEMM outside graph capture (it streams A near the HBM rate;
  // the tiled kernel reached a fifth of it at N = 32), the tiled kernel inside
  trailing_gemm<double>(M, N, K, batch, (const v2d *)A, lda, sA, (const v2d *)B, ldb, sB, (v2d *)C, ldc, sC,
                        (hipStream_t)qh_stream(stream));
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int qh_gemm_sub_f32(int M, int N, int K, int batch, const float *A, long lda, long sA, const float *B,
                               long ldb, long sB, float *C, long ldc, long sC, void *stream) {
  QH_LAUNCH(k_gemm_sub<float>, dim3((N + GT - 1) / GT, (M + GT - 1) / GT, batch), dim3(256), 0,
                     (hipStream_t)qh_stream(stream), M, N, K, (const v2f *)A, lda, sA, (const v2f *)B, ldb, sB,
                     (v2f *)C, ldc, sC);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int qh_d2f(long n, const double *x, float *y, void *stream) {
  QH_LAUNCH(k_d2f, dim3(qh_stride_blocks(n)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (const v2d *)x, (v2f *)y);
  return 0;
}
extern "C" int qh_f2d_add(long n, const float *d, double *x, void *stream) {
  QH_LAUNCH(k_f2d_add, dim3(qh_stride_blocks(n)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (const v2f *)d, (v2d *)x);
  return 0;
}
extern "C" int qh_copy_c64(long n, const double *x, double *y, void *stream) {
  QH_LAUNCH(k_copy2d, dim3(qh_stride_blocks(n)), dim3(256), 0, (hipStream_t)qh_stream(stream), n,
                     (const v2d *)x, (v2d *)y);
  return 0;
}
// Frobenius norm of n complex values (synchronises).
extern "C" double qh_norm2_c64(long n, const double *x, void *stream) {
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const int nb = 512;
  double *part = (double *)qh_dmalloc(nb * sizeof(double));
  QH_LAUNCH(k_sumsq, dim3(nb), dim3(256), 0, s, n, (const v2d *)x, part);
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
