// Device/stream plumbing shared by the host C sources (C ABI, no torch).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include "qpms_internal.h"

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  qpms_hip_die(__FILE__, __LINE__, __func__, "HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

static std::once_flag g_once;
static hipStream_t g_stream = nullptr;
static int g_init_status = -1;

extern "C" int qh_dev_init(void) {
  std::call_once(g_once, [] {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
      fprintf(stderr, "qpms_hip: no HIP device available (%s)\n", hipGetErrorString(e));
      g_init_status = (e != hipSuccess) ? (int)e : -1;
      return;
    }
    HIPCHK(hipSetDevice(0));
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    qh_noalloc_scratch_init();
    g_init_status = 0;
  });
  return g_init_status;
}

static void ensure_dev() {
  if (qh_dev_init() != 0)
    qpms_hip_die(__FILE__, __LINE__, __func__, "no usable HIP device: the GPU path has no CPU fallback");
}

extern "C" void *qh_stream(void *s) { ensure_dev(); return s ? s : (void *)g_stream; }

extern "C" void *qh_dmalloc(size_t bytes) {
  ensure_dev();
  void *p = nullptr;
  if (bytes == 0) bytes = 256;
  HIPCHK(hipMalloc(&p, bytes));
  return p;
}
extern "C" void qh_dfree(void *p) { if (p) HIPCHK(hipFree(p)); }
extern "C" void qh_h2d(void *dst, const void *src, size_t bytes, void *s) {
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)qh_stream(s)));
}
extern "C" void qh_d2h(void *dst, const void *src, size_t bytes, void *s) {
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)qh_stream(s)));
}
extern "C" void qh_sync(void *s) { HIPCHK(hipStreamSynchronize((hipStream_t)qh_stream(s))); }
extern "C" void qh_memset(void *dst, int v, size_t bytes, void *s) {
  if (bytes) HIPCHK(hipMemsetAsync(dst, v, bytes, (hipStream_t)qh_stream(s)));
}

// ---- public extensions -------------------------------------------------
extern "C" int qpms_hip_init(void) { return qh_dev_init(); }
extern "C" const char *qpms_hip_version(void) { return "qpms_hip 0.1 (gfx950)"; }
extern "C" int qpms_hip_sync(void) { qh_sync(nullptr); return 0; }
extern "C" void *qpms_hip_stream(void) { return qh_stream(nullptr); }
extern "C" void *qpms_hip_malloc(size_t bytes) { return qh_dmalloc(bytes); }
extern "C" void qpms_hip_free(void *p) { qh_dfree(p); }
extern "C" int qpms_hip_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  qh_h2d(dst, src, bytes, nullptr); qh_sync(nullptr); return 0;
}
extern "C" int qpms_hip_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  qh_d2h(dst, src, bytes, nullptr); qh_sync(nullptr); return 0;
}
