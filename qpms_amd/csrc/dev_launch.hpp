// Host-side validation of every kernel launch of the library.
//
// An AQL dispatch packet holds the grid size in work-items as a 32-bit value
// per dimension and the workgroup size as 16-bit values; a grid whose
// work-item count does not fit, a workgroup above 1024 work-items or a
// dynamic LDS request above the CU's 160 KiB would give a malformed packet
// or a silently wrapped grid. QH_LAUNCH checks the shape before
// hipLaunchKernelGGL and aborts with the kernel's name and the shape
// (the library's abort convention); a grid with a zero dimension is an empty
// problem and launches nothing.
#pragma once
#include <hip/hip_runtime.h>
#include "qpms_internal.h"

constexpr size_t QH_MAX_LDS_BYTES = 160 * 1024;

// Returns false for an empty grid (nothing to launch); aborts on a shape the
// hardware cannot take.
inline bool qh_check_launch(const char *kname, const char *file, int line, dim3 g, dim3 b, size_t lds) {
  if (g.x == 0 || g.y == 0 || g.z == 0) return false;
  const unsigned long long wg = (unsigned long long)b.x * b.y * b.z;
  const unsigned long long gx = (unsigned long long)g.x * b.x, gy = (unsigned long long)g.y * b.y,
                           gz = (unsigned long long)g.z * b.z;
  if (b.x == 0 || b.y == 0 || b.z == 0 || wg > 1024 || gx > 0xFFFFFFFFull || gy > 0xFFFFFFFFull ||
      gz > 0xFFFFFFFFull || lds > QH_MAX_LDS_BYTES)
    qpms_hip_die(file, line, kname,
                 "invalid launch shape: grid (%u, %u, %u) x block (%u, %u, %u), %llu x %llu x %llu work-items, "
                 "dynamic LDS %zu bytes",
                 g.x, g.y, g.z, b.x, b.y, b.z, gx, gy, gz, lds);
  return true;
}

#define QH_LAUNCH(K, G, B, LDS, ST, ...)                                                      \
  do {                                                                                        \
    const dim3 qh_g_ = (G), qh_b_ = (B);                                                      \
    if (qh_check_launch(#K, __FILE__, __LINE__, qh_g_, qh_b_, (size_t)(LDS)))                 \
      hipLaunchKernelGGL(K, qh_g_, qh_b_, LDS, ST, __VA_ARGS__);                              \
  } while (0)

// Blocks of 256 work-items for a grid-stride loop over n elements: enough
// to fill the device, never more than 2^20 (so the work-item count stays far
// below 2^32 whatever n is).
inline unsigned qh_stride_blocks(long n) {
  long b = (n + 255) / 256;
  if (b > (1L << 20)) b = 1L << 20;
  return (unsigned)(b < 1 ? 1 : b);
}
