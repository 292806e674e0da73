// Complex FP32 GEMM update C -= A B on MFMA for the LU's trailing update
// (M x N x K = (n - k) x (n - k) x 256 per outer block: the O(n^3) part of the
// mixed-precision factorisation).
//
// Row-major complex64 operands (interleaved re, im), batched through
// blockIdx.z with per-matrix strides. The complex product uses three real
// products instead of four (the "3M" method; N. J. Higham, Accuracy and
// Stability of Numerical Algorithms, 2nd ed., sec. 23.2.4):
//   P1 = Ar Br,  P2 = Ai Bi,  P3 = (Ar + Ai)(Br + Bi),
//   Re(A B) = P1 - P2,  Im(A B) = P3 - P1 - P2,
// so each complex multiply-add costs three MFMA products: 3/4 of the matrix
// work of the four-product form. The sums Ar + Ai and Br + Bi are formed
// once per tile when it is staged in LDS.
//
// Tiling: a workgroup of 256 threads (4 waves in 2 x 2) owns a BM x BN tile
// of C; each wave a (BM/2) x (BN/2) sub-tile made of 32 x 32 MFMA blocks
// (v_mfma_f32_32x32x2f32). K advances BK = 16 complex per LDS stage; the
// operands are de-interleaved into two planes each (A: Ar, Ai as [k][m]; B:
// Br, Bi as [k][n], rows padded by one float; the sums Ar + Ai, Br + Bi are
// formed in registers as the operands are read), double buffered, with one
// barrier per stage. When K is a multiple of BK (the LU's updates) the global
// loads of a stage are issued two stages before it is multiplied, into one of
// two register sets (unconditional loads, so each set's wait covers that set
// only); otherwise one stage before.
//
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2; the tile a workgroup takes is remapped so that every XCD works through
// a contiguous band of C tiles (which share rows of A and columns of B)
// instead of every 8th tile.
//
// MFMA 32x32x2 lane layout (f32): the A operand of lane l is A[l % 32][l / 32],
// the B operand B[l / 32][l % 32]; accumulator element i of lane l is
// D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
#include <hip/hip_runtime.h>
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v16f __attribute__((ext_vector_type(16)));

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

namespace {

constexpr int BK = 16;       // complex K per stage
constexpr int THREADS = 256;

template <int BM, int BN>
struct Tile {
  static constexpr int WM = BM / 2, WN = BN / 2;       // wave sub-tile
  static constexpr int TM = WM / 32, TN = WN / 32;     // MFMA blocks per wave
  static constexpr int SA = BM + 1, SB = BN + 1;       // padded plane row strides (floats)
  static constexpr int PA = BK * SA, PB = BK * SB;     // plane sizes
  static constexpr int STAGE = 2 * (PA + PB);          // floats per stage
  static constexpr size_t LDS = 2 * (size_t)STAGE * sizeof(float);
  // global staging per thread
  static constexpr int A_TPR = THREADS / BM;           // threads per A row
  static constexpr int A_PER = BK / A_TPR;             // complex per thread
  static constexpr int B_TPR = THREADS / BK;           // threads per B row (16)
  static constexpr int B_PER = BN / B_TPR;             // complex per thread
  static_assert(BM % 64 == 0 && BN % 64 == 0 && THREADS % BM == 0 && BK % A_TPR == 0, "tile shape");
};

template <int BM, int BN, bool KF>
__global__ __launch_bounds__(THREADS) void k_cgemm3m(int M, int N, int K, const v2f *__restrict__ A, long lda,
                                                     long sA, const v2f *__restrict__ B, long ldb, long sB,
                                                     v2f *__restrict__ C, long ldc, long sC) {
  using T = Tile<BM, BN>;
  extern __shared__ __align__(16) float lds[];
  A += (long)blockIdx.z * sA;
  B += (long)blockIdx.z * sB;
  C += (long)blockIdx.z * sC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware tile order: hardware block p runs on XCD p % 8
  const int gx = gridDim.x, ntile = gx * gridDim.y, nfull = ntile & ~7;
  int p = blockIdx.x + blockIdx.y * gx;
  if (p < nfull) p = (p & 7) * (nfull >> 3) + (p >> 3);
  const int m0 = (p / gx) * BM, n0 = (p % gx) * BN;
  const int wm = (wave >> 1) * T::WM, wn = (wave & 1) * T::WN;
  const int lr = lane & 31, lk = lane >> 5;

  // this thread's share of the global tile loads
  const int ar = tid / T::A_TPR, ac = (tid % T::A_TPR) * T::A_PER;   // A: row, first k
  const int bk = tid / T::B_TPR, bc = (tid % T::B_TPR) * T::B_PER;   // B: k, first column
  const bool arow_ok = m0 + ar < M;
  // KF (K a multiple of BK): the operand loads are unconditional (a row of A
  // or a column of B past the matrix reads the last one and reaches only
  // entries of C that are not stored), so the wait before a register set's
  // LDS store covers that set only, and two sets alternate: the loads of a
  // stage are issued two stages before it is multiplied. Otherwise entries
  // past the matrix are zero-filled and one set is used.
  v2f ga[T::A_PER], gb[T::B_PER], ha[T::A_PER], hb[T::B_PER];
  const long arow = (long)min(m0 + ar, M - 1) * lda;
  auto load = [&](int k0, v2f *xa, v2f *xb) {
    if (KF) {
#pragma unroll
      for (int c = 0; c < T::A_PER; ++c) xa[c] = A[arow + k0 + ac + c];
      const long brow = (long)(k0 + bk) * ldb;
#pragma unroll
      for (int c = 0; c < T::B_PER; ++c) xb[c] = B[brow + min(n0 + bc + c, N - 1)];
      return;
    }
#pragma unroll
    for (int c = 0; c < T::A_PER; ++c) {
      const int k = k0 + ac + c;
      xa[c] = (arow_ok && k < K) ? A[(long)(m0 + ar) * lda + k] : v2f{0.f, 0.f};
    }
    const bool bk_ok = k0 + bk < K;
#pragma unroll
    for (int c = 0; c < T::B_PER; ++c) {
      const int n = n0 + bc + c;
      xb[c] = (bk_ok && n < N) ? B[(long)(k0 + bk) * ldb + n] : v2f{0.f, 0.f};
    }
  };
  auto store = [&](int buf, const v2f *xa, const v2f *xb) {
    float *s = lds + buf * T::STAGE;
    float *Ar = s, *Ai = s + T::PA;
    float *Br = s + 2 * T::PA, *Bi = Br + T::PB;
#pragma unroll
    for (int c = 0; c < T::A_PER; ++c) {
      const int o = (ac + c) * T::SA + ar;
      Ar[o] = xa[c].x;
      Ai[o] = xa[c].y;
    }
#pragma unroll
    for (int c = 0; c < T::B_PER; ++c) {
      const int o = bk * T::SB + bc + c;
      Br[o] = xb[c].x;
      Bi[o] = xb[c].y;
    }
  };

  v16f acc[3][T::TM][T::TN];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int i = 0; i < T::TM; ++i)
#pragma unroll
      for (int j = 0; j < T::TN; ++j) acc[p][i][j] = v16f{0};

  auto multiply = [&](int buf) {
    const float *s = lds + buf * T::STAGE;
    const float *Ar = s, *Ai = s + T::PA;
    const float *Br = s + 2 * T::PA, *Bi = Br + T::PB;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      const int ka = (kk + lk) * T::SA + wm + lr, kb = (kk + lk) * T::SB + wn + lr;
      float a[3][T::TM], b[3][T::TN];
#pragma unroll
      for (int i = 0; i < T::TM; ++i) {
        a[0][i] = Ar[ka + 32 * i];
        a[1][i] = Ai[ka + 32 * i];
        a[2][i] = a[0][i] + a[1][i];
      }
#pragma unroll
      for (int j = 0; j < T::TN; ++j) {
        b[0][j] = Br[kb + 32 * j];
        b[1][j] = Bi[kb + 32 * j];
        b[2][j] = b[0][j] + b[1][j];
      }
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int i = 0; i < T::TM; ++i)
#pragma unroll
          for (int j = 0; j < T::TN; ++j)
            acc[p][i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[p][i], b[p][j], acc[p][i][j], 0, 0, 0);
    }
  };

  if (KF) {
    // stages in pairs, so that the register sets alternate without copies:
    // at stage st the set g holds stage st + 1 and h receives st + 2. Loads
    // past the last stage re-read it (straight-line code, no waits on
    // conditions); what they store is never multiplied.
    const int ns = K / BK;
    load(0, ga, gb);
    store(0, ga, gb);
    load(min(1, ns - 1) * BK, ga, gb);
    __syncthreads();
    int st = 0;
    for (; st + 1 < ns; st += 2) {
      load(min(st + 2, ns - 1) * BK, ha, hb);
      __builtin_amdgcn_sched_barrier(0);   // the loads stay ahead of the products
      multiply(0);
      store(1, ga, gb);
      __syncthreads();
      load(min(st + 3, ns - 1) * BK, ga, gb);
      __builtin_amdgcn_sched_barrier(0);
      multiply(1);
      store(0, ha, hb);
      __syncthreads();
    }
    if (st < ns) multiply(0);   // odd stage count: the last stage, stored by the loop
  } else {
    load(0, ga, gb);
    store(0, ga, gb);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += BK) {
      const bool next = k0 + BK < K;
      if (next) load(k0 + BK, ga, gb);
      multiply(cur);
      if (next) store(cur ^ 1, ga, gb);
      __syncthreads();
      cur ^= 1;
    }
  }

#pragma unroll
  for (int i = 0; i < T::TM; ++i)
#pragma unroll
    for (int j = 0; j < T::TN; ++j) {
      const int col = n0 + wn + 32 * j + lr;
      if (col >= N) continue;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = m0 + wm + 32 * i + 8 * (e >> 2) + 4 * lk + (e & 3);
        if (row < M) {
          const float p1 = acc[0][i][j][e], p2 = acc[1][i][j][e], p3 = acc[2][i][j][e];
          v2f &c = C[(long)row * ldc + col];
          c = v2f{c.x - (p1 - p2), c.y - (p3 - p1 - p2)};
        }
      }
    }
}

template <int BM, int BN>
void gemm_prepare() {
  static bool attr = [] {
    HIPCHK(hipFuncSetAttribute((const void *)k_cgemm3m<BM, BN, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)Tile<BM, BN>::LDS));
    HIPCHK(hipFuncSetAttribute((const void *)k_cgemm3m<BM, BN, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)Tile<BM, BN>::LDS));
    return true;
  }();
  (void)attr;
}

template <int BM, int BN>
void gemm_launch(int M, int N, int K, int batch, const v2f *A, long lda, long sA, const v2f *B, long ldb, long sB,
            v2f *C, long ldc, long sC, hipStream_t s) {
  gemm_prepare<BM, BN>();
  if (K <= 0) return;
  const dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, batch);
  if (K % BK == 0)
    QH_LAUNCH((k_cgemm3m<BM, BN, true>), grid, dim3(THREADS), (Tile<BM, BN>::LDS), s, M, N, K, A, lda, sA, B, ldb, sB,
              C, ldc, sC);
  else
    QH_LAUNCH((k_cgemm3m<BM, BN, false>), grid, dim3(THREADS), (Tile<BM, BN>::LDS), s, M, N, K, A, lda, sA, B, ldb,
              sB, C, ldc, sC);
}

}  // namespace

extern "C" void qh_cgemm3m_prepare(void) {
  gemm_prepare<128, 128>();
  gemm_prepare<128, 64>();
  gemm_prepare<64, 64>();
}

// C[M x N] -= A[M x K] B[K x N], complex FP32 (3M on MFMA), batched. `tile`
// selects the workgroup tile: 0 = 128 x 128, 1 = 128 x 64, 2 = 64 x 64.
extern "C" int qh_cgemm3m_sub(int M, int N, int K, int batch, const float *A, long lda, long sA, const float *B,
                              long ldb, long sB, float *C, long ldc, long sC, int tile, void *stream) {
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return 0;
  hipStream_t s = (hipStream_t)qh_stream(stream);
  const v2f *a = (const v2f *)A, *b = (const v2f *)B;
  v2f *c = (v2f *)C;
  switch (tile) {
    case 0: gemm_launch<128, 128>(M, N, K, batch, a, lda, sA, b, ldb, sB, c, ldc, sC, s); break;
    case 1: gemm_launch<128, 64>(M, N, K, batch, a, lda, sA, b, ldb, sB, c, ldc, sC, s); break;
    case 2: gemm_launch<64, 64>(M, N, K, batch, a, lda, sA, b, ldb, sB, c, ldc, sC, s); break;
    default: qpms_hip_die(__FILE__, __LINE__, __func__, "unknown GEMM tile %d", tile);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Autotuner of the FP32 LU trailing update (C -= A B with K = 256, the
// factorisation's O(n^3) part). Candidates: the library GEMM (rocBLAS
// cgemm), the 3M kernel above with its three tile shapes, and the
// four-product kernel of k_linalg.hip. A size class is the larger of M, N
// (buckets below) and whether the call is batched; each class is timed once
// on scratch operands of its representative shape (HIP events, best of three
// after a warm-up), and the fastest candidate is recorded. Tuning allocates
// scratch memory, so it runs only when asked (qpms_hip_autotune, called by
// the serving loop at creation); untuned classes use the built-in defaults,
// and QPMS_HIP_GEMM=rocblas|own|3m forces one variant everywhere.
// ---------------------------------------------------------------------------
#include <mutex>
#include <cstring>
#include <cstdlib>
#include <rocblas/rocblas.h>

namespace {
constexpr int NCLASS = 6;
constexpr int CLASS_DIM[NCLASS] = {768, 2048, 4096, 8192, 16384, 32768};   // representative (upper) dims
constexpr int NVAR = 5;
constexpr float OWN_MARGIN = 1.03f;   // preference for a hand-written variant over the library
const char *const VAR_NAME[NVAR] = {"rocblas", "3m_128x128", "3m_128x64", "3m_64x64", "own_4m"};
std::mutex g_tune_mu;
// [batched][class] -> variant; -1: not tuned
int g_tuned[2][NCLASS] = {{-1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}};
float g_tuned_ms[2][NCLASS][NVAR];
// while whole factorisations are timed: single-matrix updates of this class use this variant
int g_override_class = -1, g_override_var = -1;

int size_class(int M, int N) {
  const int d = M > N ? M : N;
  for (int c = 0; c < NCLASS - 1; ++c)
    if (d <= CLASS_DIM[c]) return c;
  return NCLASS - 1;
}

int forced_variant() {
  static int v = [] {
    const char *e = getenv("QPMS_HIP_GEMM");
    if (!e) return -1;
    if (!strcmp(e, "rocblas")) return 0;
    if (!strcmp(e, "3m")) return 1;
    if (!strcmp(e, "3m128x64")) return 2;
    if (!strcmp(e, "3m64")) return 3;
    if (!strcmp(e, "own")) return 4;
    return -1;
  }();
  return v;
}
}  // namespace

// Variant of the FP32 trailing update for an M x N update (batched or not):
// 0 library, 1-3 3M tiles (128x128, 128x64, 64x64), 4 four-product kernel.
extern "C" int qh_lu_gemm_variant(int M, int N, int batched) {
  const int f = forced_variant();
  if (f >= 0) return f;
  std::lock_guard<std::mutex> g(g_tune_mu);
  const int c = size_class(M, N), t = g_tuned[batched ? 1 : 0][c];
  if (!batched && c == g_override_class) return g_override_var;
  if (t >= 0) return t;
  return (M >= 512 && N >= 512) ? 0 : 4;   // untuned: the library GEMM for large updates
}

// The fastest hand-written variant of a size class (tuned timings; the 3M
// 128 x 128 kernel when untuned).
extern "C" int qh_lu_gemm_best_own(int M, int N, int batched) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  const int c = size_class(M, N), b = batched ? 1 : 0;
  if (g_tuned[b][c] < 0) return 1;
  int best = 1;
  for (int v = 2; v < NVAR; ++v)
    if (g_tuned_ms[b][c][v] < g_tuned_ms[b][c][best]) best = v;
  return best;
}

// Variant v of C -= A B (FP32 complex) on stream s; the library variant
// through the caller's rocBLAS handle.
extern "C" int qh_lu_gemm_run(int v, int M, int N, int K, int batch, const float *A, long lda, long sA,
                              const float *B, long ldb, long sB, float *C, long ldc, long sC, void *blas,
                              void *stream) {
  switch (v) {
    case 0: {
      const rocblas_float_complex al = {-1.f, 0.f}, be = {1.f, 0.f};
      const rocblas_status st = rocblas_cgemm_strided_batched(
          (rocblas_handle)blas, rocblas_operation_none, rocblas_operation_none, N, M, K, &al,
          (const rocblas_float_complex *)B, (rocblas_int)ldb, sB, (const rocblas_float_complex *)A, (rocblas_int)lda,
          sA, &be, (rocblas_float_complex *)C, (rocblas_int)ldc, sC, batch);
      if (st != rocblas_status_success)
        qpms_hip_die(__FILE__, __LINE__, __func__, "rocBLAS cgemm failed (status %d)", (int)st);
      return 0;
    }
    case 1: case 2: case 3:
      return qh_cgemm3m_sub(M, N, K, batch, A, lda, sA, B, ldb, sB, C, ldc, sC, v - 1, stream);
    case 4:
      return qh_gemm_sub_f32(M, N, K, batch, A, lda, sA, B, ldb, sB, C, ldc, sC, stream);
  }
  qpms_hip_die(__FILE__, __LINE__, __func__, "unknown GEMM variant %d", v);
}

// Times every candidate for every size class (batched: 4 matrices of the
// class's size / 2) on scratch operands and records the fastest; the
// single-matrix classes of order 2048 and above are then re-decided by whole
// factorisations (qh_autotune_lu). Returns the number of GEMM-timed classes. `report` (optional, >= 2 * 6 * 5 floats) receives
// the measured milliseconds [batched][class][variant].
extern "C" int qpms_hip_autotune(float *report) {
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success)
    qpms_hip_die(__FILE__, __LINE__, __func__, "rocblas_create_handle failed");
  rocblas_set_stream(h, s);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int K = 256;
  int tuned = 0;
  for (int bt = 0; bt < 2; ++bt)
    for (int c = 0; c < NCLASS; ++c) {
      const int d = bt ? CLASS_DIM[c] / 2 : CLASS_DIM[c], batch = bt ? 4 : 1;
      const size_t na = (size_t)d * K, nc = (size_t)d * d;
      float *A = nullptr, *B = nullptr, *C = nullptr;
      if (hipMalloc(&A, batch * na * 8) != hipSuccess || hipMalloc(&B, batch * na * 8) != hipSuccess ||
          hipMalloc(&C, batch * nc * 8) != hipSuccess) {
        (void)hipGetLastError();
        if (A) HIPCHK(hipFree(A));
        if (B) HIPCHK(hipFree(B));
        if (C) HIPCHK(hipFree(C));
        continue;   // class too large for the free memory: keep its default
      }
      HIPCHK(hipMemsetAsync(A, 0, batch * na * 8, s));
      HIPCHK(hipMemsetAsync(B, 0, batch * na * 8, s));
      HIPCHK(hipMemsetAsync(C, 0, batch * nc * 8, s));
      int best = -1;
      float bestms = 1e30f;
      for (int v = 0; v < NVAR; ++v) {
        float ms = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {
          HIPCHK(hipEventRecord(e0, s));
          qh_lu_gemm_run(v, d, d, K, batch, A, K, (long)na, B, d, (long)na, C, d, (long)nc, h, s);
          HIPCHK(hipEventRecord(e1, s));
          HIPCHK(hipEventSynchronize(e1));
          float t = 0;
          HIPCHK(hipEventElapsedTime(&t, e0, e1));
          if (rep > 0 && t < ms) ms = t;   // rep 0 warms up
        }
        if (report) report[(bt * NCLASS + c) * NVAR + v] = ms;
        g_tuned_ms[bt][c][v] = ms;
        if (ms < bestms) { bestms = ms; best = v; }
      }
      // a hand-written variant within OWN_MARGIN of the library is preferred
      // (it also runs inside captured graphs and needs no library workspace)
      if (best == 0) {
        int own = -1;
        for (int v = 1; v < NVAR; ++v)
          if (own < 0 || g_tuned_ms[bt][c][v] < g_tuned_ms[bt][c][own]) own = v;
        if (own > 0 && g_tuned_ms[bt][c][own] <= OWN_MARGIN * bestms) best = own;
      }
      {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tuned[bt][c] = best;
      }
      ++tuned;
      HIPCHK(hipFree(A));
      HIPCHK(hipFree(B));
      HIPCHK(hipFree(C));
    }
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  rocblas_destroy_handle(h);
  HIPCHK(hipStreamDestroy(s));
  qh_autotune_lu(nullptr);   // the large single-matrix classes, by whole factorisations
  return tuned;
}

// Whole-factorisation tuning of the single-matrix classes from LU_TUNE_MIN on:
// the trailing update of a blocked LU runs beside the panel chain (on a
// CU-masked stream, next to whole-CU strip kernels), where the ranking of
// the candidates can differ from that of an isolated GEMM. For each such
// class the FP32 LU of a matrix of the class's order is timed once per
// candidate (rocBLAS, 3M 128 x 64, 3M 64 x 64), with only that class's
// updates switched, and the fastest is recorded. lu(n, ms) factors a fresh
// matrix of order n and returns the milliseconds. As for the GEMM timings, a
// hand-written candidate within OWN_MARGIN of the library is preferred.
constexpr int LU_TUNE_MIN = 2048;
extern "C" int qh_autotune_lu_classes(float (*lu)(int n, void *ctx), void *ctx, float *report) {
  const int cand[3] = {0, 2, 3};
  int tuned = 0;
  for (int c = 0; c < NCLASS; ++c) {
    const int d = CLASS_DIM[c];
    if (d < LU_TUNE_MIN) continue;
    float ms[3];
    for (int k = 0; k < 3; ++k) {
      {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_override_class = c;
        g_override_var = cand[k];
      }
      ms[k] = lu(d, ctx);
      if (report) report[c * 3 + k] = ms[k];
    }
    int best = -1;
    for (int k = 0; k < 3; ++k)
      if (ms[k] > 0 && (best < 0 || ms[k] < ms[best])) best = k;
    if (best == 0) {   // the faster hand-written candidate, if within OWN_MARGIN of the library
      const int own = (ms[2] > 0 && (ms[1] <= 0 || ms[2] < ms[1])) ? 2 : 1;
      if (ms[own] > 0 && ms[own] <= OWN_MARGIN * ms[0]) best = own;
    }
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_override_class = g_override_var = -1;
    if (best >= 0) { g_tuned[0][c] = cand[best]; ++tuned; }
  }
  return tuned;
}

// The tuned table as text: one line per class, "batched dim variant-name".
extern "C" int qpms_hip_autotune_table(char *buf, size_t len) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  size_t o = 0;
  for (int bt = 0; bt < 2; ++bt)
    for (int c = 0; c < NCLASS; ++c) {
      const int v = g_tuned[bt][c];
      const int n = snprintf(buf + o, o < len ? len - o : 0, "%d %d %s\n", bt, CLASS_DIM[c], v >= 0 ? VAR_NAME[v] : "default");
      if (n > 0) o += (size_t)n;
    }
  return (int)o;
}
