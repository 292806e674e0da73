// Measured peaks that the roofline fractions are read against, next to the
// datasheet values (8 TB/s HBM, 78.6 TFLOP/s FP64 vector):
//
//  k_triad  STREAM triad a = b + s c over doubles: 16-byte non-temporal
//           loads and stores, one contiguous chunk per workgroup, 24
//           algorithmic bytes per element.
//  k_fill   write-only stream a = s over doubles, same layout: the ceiling
//           of a kernel that only writes, such as the translation-matrix
//           build.
//  k_fma64  eight independent FP64 FMA chains per lane; the chains stay
//           bounded (x -> 0.999 x + 1e-3) and their sum is stored, so no
//           work is dead. 16 flops per lane per iteration.
#include <hip/hip_runtime.h>
#include "qpms_internal.h"
#include "dev_launch.hpp"

typedef double v2d __attribute__((ext_vector_type(2)));

namespace {

// Each workgroup streams one contiguous chunk of PROBE_CHUNK 16-byte elements
// (256 KiB per array), its lanes 16 bytes apart.
constexpr long PROBE_CHUNK = 16384;
__global__ __launch_bounds__(256) void k_triad(long n2, v2d *__restrict__ a, const v2d *__restrict__ b,
                                               const v2d *__restrict__ c, double s) {
  const long b0 = (long)blockIdx.x * PROBE_CHUNK, b1 = min(b0 + PROBE_CHUNK, n2);
  for (long i = b0 + threadIdx.x; i < b1; i += 256) {
    const v2d x = __builtin_nontemporal_load(b + i), y = __builtin_nontemporal_load(c + i);
    __builtin_nontemporal_store(x + s * y, a + i);
  }
}

__global__ __launch_bounds__(256) void k_fill(long n2, v2d *__restrict__ a, double s) {
  const long b0 = (long)blockIdx.x * PROBE_CHUNK, b1 = min(b0 + PROBE_CHUNK, n2);
  const v2d v = v2d{s, s};
  for (long i = b0 + threadIdx.x; i < b1; i += 256) __builtin_nontemporal_store(v, a + i);
}

__global__ __launch_bounds__(256) void k_fma64(int iters, double *__restrict__ out) {
  const double m = 0.999, q = 1e-3;
  double r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = 1e-3 * (threadIdx.x + j);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = fma(r[j], m, q);
  }
  double sum = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) sum += r[j];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

}  // namespace

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

// n doubles per array (even); a, b, c 16-byte aligned device buffers.
extern "C" int qpms_hip_probe_triad(size_t n, double *a, const double *b, const double *c, double s, void *stream) {
  if (n == 0) return 0;
  if (n & 1) qpms_hip_die(__FILE__, __LINE__, __func__, "triad length %zu must be even", n);
  hipStream_t st = (hipStream_t)qh_stream(stream);
  const long n2 = (long)(n / 2);
  QH_LAUNCH(k_triad, dim3((unsigned)((n2 + PROBE_CHUNK - 1) / PROBE_CHUNK)), dim3(256), 0, st, n2, (v2d *)a,
                     (const v2d *)b, (const v2d *)c, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// n doubles (even), 8 n bytes written.
extern "C" int qpms_hip_probe_fill(size_t n, double *a, double s, void *stream) {
  if (n == 0) return 0;
  if (n & 1) qpms_hip_die(__FILE__, __LINE__, __func__, "fill length %zu must be even", n);
  hipStream_t st = (hipStream_t)qh_stream(stream);
  const long n2 = (long)(n / 2);
  QH_LAUNCH(k_fill, dim3((unsigned)((n2 + PROBE_CHUNK - 1) / PROBE_CHUNK)), dim3(256), 0, st, n2, (v2d *)a, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// blocks x 256 lanes, iters iterations: 16 * iters * 256 * blocks flops;
// out holds blocks * 256 doubles.
extern "C" int qpms_hip_probe_fma64(int blocks, int iters, double *out, void *stream) {
  if (blocks <= 0 || iters <= 0) return 0;
  hipStream_t st = (hipStream_t)qh_stream(stream);
  QH_LAUNCH(k_fma64, dim3(blocks), dim3(256), 0, st, iters, out);
  HIPCHK(hipGetLastError());
  return 0;
}
