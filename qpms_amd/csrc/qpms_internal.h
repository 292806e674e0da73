/* Internal declarations shared by the host C sources and the HIP launchers. */
#ifndef QPMS_INTERNAL_H
#define QPMS_INTERNAL_H

#include "../../include/qpms_hip.h"
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Abort with `file:line, func: message` on stderr — the reference's error
 * convention (qpms_error.h:166-244, error.c:43-55). */
void qpms_hip_die(const char *file, int line, const char *func, const char *fmt, ...)
	__attribute__((noreturn, format(printf, 4, 5)));
#define QH_DIE(...) qpms_hip_die(__FILE__, __LINE__, __func__, __VA_ARGS__)
#define QH_ENSURE(cond, ...) do { if (!(cond)) QH_DIE(__VA_ARGS__); } while (0)
#define QH_MALLOC(ptr, bytes) do { (ptr) = malloc(bytes); \
	if (!(ptr) && (size_t)(bytes) != 0) QH_DIE("malloc of %zu bytes failed", (size_t)(bytes)); } while (0)
#define QH_CALLOC(ptr, n, sz) do { (ptr) = calloc((n), (sz)); \
	if (!(ptr) && (size_t)(n) != 0 && (size_t)(sz) != 0) QH_DIE("calloc of %zu x %zu bytes failed", (size_t)(n), (size_t)(sz)); } while (0)

/* Device-resident tables of one translation calculator. Built lazily on the
 * first GPU call and kept until qpms_trans_calculator_free. The calculator
 * struct has no spare field, so the tables live in a registry keyed by the
 * calculator address (see calculator.c). */
typedef struct qh_trans_tables {
	int lMax;
	int nelem;
	int qslots;              /* lMax + 1 multiplier slots per (y, yu) */
	/* Multipliers, zero-padded: A at [(y*nelem + yu)*qslots + q],
	 * B the same (the B slot q corresponds to the reference's q). */
	double *dA;              /* complex as (re, im) pairs */
	double *dB;
	/* Per (y, yu) the number of A terms (qmax_A + 1) and B terms (Qmax_B + 1). */
	unsigned char *dqA, *dqB;
} qh_trans_tables;

const qh_trans_tables *qh_tables_get(const qpms_trans_calculator *c);
void qh_tables_release(const qpms_trans_calculator *c);

/* Host-side multiplier evaluation (calculator.c). */
int qh_gaunt_q_max(int m, int n, int mu, int nu);
int qh_gauntB_Q_max(int M, int n, int mu, int nu);

/* ---- HIP launchers (k_*.hip). All take a hipStream_t as void*. ---------- */
int qh_dev_init(void);
void qh_noalloc_scratch_init(void);     /* k_linalg.hip: device address of the no-alloc scratch */
int qh_zero(void *dst, size_t bytes, void *stream);   /* k_serve.hip: clear by a kernel */
void *qh_stream(void *s);               /* NULL → library stream */
void *qh_dmalloc(size_t bytes);
void qh_dfree(void *p);
void qh_h2d(void *dst, const void *src, size_t bytes, void *stream);
void qh_d2h(void *dst, const void *src, size_t bytes, void *stream);
void qh_sync(void *stream);
void qh_memset(void *dst, int v, size_t bytes, void *stream);

int qh_launch_ab_batch(const qh_trans_tables *t, size_t npts, const double *kr_dev,
		const double *theta_dev, const double *phi_dev, int kind,
		double *A_dev, double *B_dev, void *stream);
int qh_launch_w_build(const qh_trans_tables *t, size_t npart, const double *pos_dev,
		double k_re, double k_im, int kind, size_t row_p0, size_t row_np,
		double *W_dev, size_t ld, void *stream);
int qh_launch_modeproblem(const qh_trans_tables *t, size_t npart, const double *pos_dev,
		const int *ttype_dev, const double *tdiag_dev, double k_re, double k_im,
		double *M_dev, size_t ld, void *stream);
int qh_launch_mie(int lMax, int nsph, const double *params_dev, double *out_dev, void *stream);
int qh_launch_block_tw(size_t nr, size_t L, const double *T, const double *W, double *M, size_t col0, void *stream);

/* k_linalg.hip: batched complex LU / solves. Complex arrays are interleaved
 * (re, im); `s*` arguments are per-matrix strides in complex elements,
 * leading dimensions are the matrix order (packed). */
int qh_getrf_batched_f64(int n, int batch, double *A, long sA, int *ipiv, int *perm, int *info, void *stream);
int qh_getrs_batched_f64(int n, int nrhs, int batch, const double *LU, long sA, const int *perm,
		double *B, long sB, double *work, void *stream);
int qh_getrf_batched_f32(int n, int batch, float *A, long sA, int *ipiv, int *perm, int *info, void *stream);
int qh_getrs_batched_f32(int n, int nrhs, int batch, const float *LU, long sA, const int *perm,
		float *B, long sB, float *work, void *stream);
int qh_gemm_sub_f64(int M, int N, int K, int batch, const double *A, long lda, long sA, const double *B,
		long ldb, long sB, double *C, long ldc, long sC, void *stream);
int qh_gemm_sub_f32(int M, int N, int K, int batch, const float *A, long lda, long sA, const float *B,
		long ldb, long sB, float *C, long ldc, long sC, void *stream);
/* k_gemm.hip: the 3M FP32 complex GEMM and the tuned LU trailing update */
int qh_cgemm3m_sub(int M, int N, int K, int batch, const float *A, long lda, long sA, const float *B,
		long ldb, long sB, float *C, long ldc, long sC, int tile, void *stream);
int qh_lu_gemm_variant(int M, int N, int batched);
int qh_lu_gemm_best_own(int M, int N, int batched);
int qh_lu_gemm_run(int v, int M, int N, int K, int batch, const float *A, long lda, long sA,
		const float *B, long ldb, long sB, float *C, long ldc, long sC, void *blas, void *stream);
/* Whole-factorisation tuning of the trailing-update variant per size class
 * (k_gemm.hip drives, k_linalg.hip times the LUs). */
int qh_autotune_lu_classes(float (*lu)(int n, void *ctx), void *ctx, float *report);
int qh_autotune_lu(float *report);
int qh_d2f(long n, const double *x, float *y, void *stream);
int qh_f2d_add(long n, const float *d, double *x, void *stream);
int qh_copy_c64(long n, const double *x, double *y, void *stream);
double qh_norm2_c64(long n, const double *x, void *stream);

#ifdef __cplusplus
}
#endif
#endif
