/* Beyn's contour-integral eigensolver (beyn.c:69-86, 345-646) and the finite
 * scattering-system wrapper (scatsystem.c:2281-2325), GPU-resident.
 *
 * Per quadrature node: M(z) from the caller's callback (host, as in the
 * reference), then X = M(z)^{-1} V on the GPU (FP64 LU, or with
 * QPMS_HIP_LU_MODE=mixed the FP32-MFMA LU with FP64 refinement), and the
 * moment sums A0, A1 (and their even-node "coarse" versions) accumulate in
 * HBM. The SVD of A0 (one-sided Jacobi), V0^H A1, the reduced K x K
 * eigenproblem (Hessenberg + shifted QR in one workgroup) and V0 S all run on
 * the GPU in FP64 (k_beyn.hip).
 *
 * Differences from the reference, by design:
 *  - the probe matrix V is drawn from a seeded generator (environment variable
 *    QPMS_HIP_BEYN_SEED, default 1) instead of time(NULL) (beyn.c:290), and
 *    beyn_solve_v takes it from the caller;
 *  - nothing is printed (the reference prints unconditionally, beyn.c:354-600);
 *  - the scattering-system wrapper uses l = min(m, QPMS_HIP_BEYN_L) columns
 *    (default 64) instead of l = m (scatsystem.c:2318).
 */
#include "qpms_internal.h"
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

int qh_beyn_accum(long n, const double *X, double dzr, double dzi, double wr, double wi, int coarse,
		double *A0, double *A1, double *A0c, double *A1c, void *stream);
int qh_jacobi_svd(int m, int l, const double *A, double *Ut_out, double *sigma_host, double *Vt_out, void *stream,
		void *ws);
size_t qh_jacobi_svd_ws_bytes(int m, int l);
int qh_ctrans_gemm(int m, int l, int K, const double *Ut, const double *A, double *C, void *stream);
int qh_v0s(int m, int K, const double *S, const double *Ut, double *Y, void *stream);
int qh_small_eig(int K, const double *B, double *lam, double *S, void *stream, void *ws);
int qh_bmat(int K, int l, const double *T, const double *Vt, const double *sigma, double *B, void *stream);

/* (Re(z - c)/rRe)^2 + (Im(z - c)/rIm)^2 < 1 with the half-axes stored after
 * the nodes (beyn.c:62-67). */
static bool qh_ellipse_inside(struct beyn_contour_t *c, complex double z) {
	const double rRe = creal(c->z_dz[c->n][0]), rIm = creal(c->z_dz[c->n][1]);
	const double a = creal(z - c->centre) / rRe, b = cimag(z - c->centre) / rIm;
	return a * a + b * b < 1;
}

beyn_contour_t *beyn_contour_ellipse(complex double centre, double rRe, double rIm, size_t n) {
	beyn_contour_t *c;
	QH_MALLOC(c, sizeof(beyn_contour_t) + (n + 1) * sizeof(c->z_dz[0]));
	c->centre = centre;
	c->n = n;
	for (size_t i = 0; i < n; ++i) {
		const double t = i * 2 * M_PI / n, st = sin(t), ct = cos(t);
		c->z_dz[i][0] = centre + ct * rRe + I * st * rIm;
		c->z_dz[i][1] = (-rRe * st + I * rIm * ct) * (2 * M_PI / n);
	}
	c->z_dz[n][0] = rRe;   /* half-axes kept after the nodes, as the reference does */
	c->z_dz[n][1] = rIm;
	c->inside_test = qh_ellipse_inside;
	return c;
}

/* Rotation taking the "upper half" (IM_PLUS) shape to the requested
 * orientation, and which half-axis runs along the straight cut
 * (beyn.c:127-147). */
static complex double qh_half_orient(beyn_contour_halfellipse_orientation o, double rRe, double rIm,
		double *l, double *h) {
	*l = rRe; *h = rIm;
	switch (o) {
		case BEYN_CONTOUR_HALFELLIPSE_RE_PLUS: *l = rIm; *h = rRe; return -I;
		case BEYN_CONTOUR_HALFELLIPSE_RE_MINUS: *l = rIm; *h = rRe; return I;
		case BEYN_CONTOUR_HALFELLIPSE_IM_PLUS: return 1;
		case BEYN_CONTOUR_HALFELLIPSE_IM_MINUS: return -1;
	}
	QH_DIE("invalid half-ellipse orientation %d", (int)o);
	return 0;
}

/* Nodes on the straight cut get the signed zero of the side the contour
 * encloses, so that functions with a branch cut there see the right side
 * (beyn.c:91-117). */
static complex double qh_align_zero(complex double z, beyn_contour_halfellipse_orientation o) {
	double re = creal(z), im = cimag(z);
	switch (o) {
		case BEYN_CONTOUR_HALFELLIPSE_RE_PLUS: if (re == 0) re = +0.0; break;
		case BEYN_CONTOUR_HALFELLIPSE_RE_MINUS: if (re == 0) re = -0.0; break;
		case BEYN_CONTOUR_HALFELLIPSE_IM_PLUS: if (im == 0) im = +0.0; break;
		case BEYN_CONTOUR_HALFELLIPSE_IM_MINUS: if (im == 0) im = -0.0; break;
	}
	return CMPLX(re, im);
}

beyn_contour_t *beyn_contour_halfellipse(complex double centre, double rRe, double rIm, size_t n,
		beyn_contour_halfellipse_orientation o) {
	beyn_contour_t *c;
	QH_MALLOC(c, sizeof(beyn_contour_t) + (n + 1) * sizeof(c->z_dz[0]) + sizeof(beyn_contour_halfellipse_orientation));
	c->centre = centre;
	c->n = n;
	double l, h;
	const complex double f = qh_half_orient(o, rRe, rIm, &l, &h);
	const size_t nline = n / 2, narc = n - nline;
	/* arc: midpoint rule over t in (0, pi) (beyn.c:149-154) */
	for (size_t i = 0; i < narc; ++i) {
		const double t = (i + 0.5) * M_PI / narc;
		c->z_dz[i][0] = centre + f * (cos(t) * l + I * sin(t) * h);
		c->z_dz[i][1] = f * (-l * sin(t) + I * h * cos(t)) * (M_PI / narc);
	}
	/* cut: nline equal steps from +l back to -l (beyn.c:155-159) */
	for (size_t i = 0; i < nline; ++i) {
		const double t = 0.5 * (1 - (double)nline) + i;
		c->z_dz[narc + i][0] = qh_align_zero(centre + f * t * 2 * l / nline, o);
		c->z_dz[narc + i][1] = f * 2 * l / nline;
	}
	c->z_dz[n][0] = rRe;
	c->z_dz[n][1] = rIm;
	memcpy(&c->z_dz[n + 1][0], &o, sizeof o);
	c->inside_test = NULL;
	return c;
}

beyn_contour_t *beyn_contour_kidney(complex double centre, double rRe, double rIm, const double rounding,
		const size_t n, beyn_contour_halfellipse_orientation o) {
	QH_ENSURE(rounding >= 0 && rounding < .5, "rounding must lie in the interval [0, 0.5)");
	beyn_contour_t *c;
	QH_MALLOC(c, sizeof(beyn_contour_t) + (n + 1) * sizeof(c->z_dz[0]) + sizeof(beyn_contour_halfellipse_orientation));
	c->centre = centre;
	c->n = n;
	double l, h;
	const complex double f = qh_half_orient(o, rRe, rIm, &l, &h);
	/* A half-disc of radius h (stretched by l/h along the cut) whose two
	 * corners are replaced by arcs of radius y = rounding*h tangent to both
	 * the cut and the big arc; the parameter is arc length in the unstretched
	 * shape, sampled uniformly (beyn.c:199-247). */
	const double y = rounding * h;
	const double x = sqrt((h - y) * (h - y) - y * y);
	const double alpha = asin(y / (h - y));
	const double ar = l / h;
	const double tmax = 2 * (x + (M_PI_2 + alpha) * y + (M_PI_2 - alpha) * h);
	const double dt = tmax / n;
	size_t i = 0;
	double t, t0 = 0, t1 = x;
	for (; t = i * dt, t <= t1; ++i) {                      /* cut, centre -> right corner */
		c->z_dz[i][0] = qh_align_zero(centre + (t - t0) * ar * f, o);
		c->z_dz[i][1] = dt * ar * f;
	}
	t0 = t1; t1 = t0 + (M_PI_2 + alpha) * y;
	for (; t = i * dt, t < t1; ++i) {                       /* right rounding arc */
		const double phi = (t - t0) / y - M_PI_2;
		c->z_dz[i][0] = centre + (ar * (x + y * cos(phi)) + y * (1 + sin(phi)) * I) * f;
		c->z_dz[i][1] = dt * (-ar * sin(phi) + cos(phi) * I) * f;
	}
	t0 = t1; t1 = t0 + (M_PI - 2 * alpha) * h;
	for (; t = i * dt, t < t1; ++i) {                       /* big arc */
		const double phi = (t - t0) / h + alpha;
		c->z_dz[i][0] = centre + (ar * h * cos(phi) + h * sin(phi) * I) * f;
		c->z_dz[i][1] = dt * (-ar * sin(phi) + cos(phi) * I) * f;
	}
	t0 = t1; t1 = t0 + (M_PI_2 + alpha) * y;
	for (; t = i * dt, t < t1; ++i) {                       /* left rounding arc */
		const double phi = (t - t0) / y + M_PI - alpha;
		c->z_dz[i][0] = centre + (ar * (-x + y * cos(phi)) + y * (1 + sin(phi)) * I) * f;
		c->z_dz[i][1] = dt * (-ar * sin(phi) + cos(phi) * I) * f;
	}
	t0 = t1;
	for (; t = i * dt, i < n; ++i) {                        /* cut, left corner -> centre */
		c->z_dz[i][0] = qh_align_zero(centre + (t - t0 - x) * ar * f, o);
		c->z_dz[i][1] = dt * ar * f;
	}
	c->inside_test = NULL;
	return c;
}

void beyn_result_free(beyn_result_t *r) {
	if (r) {
		free(r->eigval);
		free(r->eigval_err);
		free(r->residuals);
		free(r->eigvec);
		free(r->ranktest_SV);
	}
	free(r);
}

/* xorshift64* + Box–Muller (beyn.c:47-58 draws N(0,1) for Re and Im) */
static uint64_t qh_rng_next(uint64_t *s) {
	*s ^= *s >> 12; *s ^= *s << 25; *s ^= *s >> 27;
	return *s * 2685821657736338717ULL;
}
static double qh_rng_u01(uint64_t *s) { return ((qh_rng_next(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
static double qh_randn(uint64_t *s) {
	const double u1 = qh_rng_u01(s), u2 = qh_rng_u01(s);
	return sqrt(-2 * log(u1)) * cos(2 * M_PI * u2);
}

/* Steps 3-6 of Beyn's algorithm on device A0, A1 (m x l), on `stream`.
 * Returns K (retained) and fills the result arrays (host). With
 * res_tol > 0 the residual ||M(z) v|| of every candidate is computed from the
 * M callback and candidates above res_tol are dropped. */
static int qh_beyn_process(size_t m, size_t l, beyn_function_M_t M, void *params,
		const complex double *dA0, const complex double *dA1, complex double z0,
		double rank_tol, size_t rank_sel_min, double res_tol,
		complex double *eigval, complex double *eigvec /* l x m, nullable */, double *residuals, double *sigma,
		void *stream, void *ws) {
	/* device scratch: from `ws` (qh_beyn_ws_bytes) when given, else allocated */
	char *wp = ws;
	#define QH_TAKE(b) (ws ? (void *)((wp += (((b) + 255) & ~(size_t)255)) - (((b) + 255) & ~(size_t)255)) : qh_dmalloc(b))
	#define QH_GIVE(p) do { if (!ws) qh_dfree(p); } while (0)
	complex double *dUt = QH_TAKE(m * l * sizeof(complex double));
	complex double *dVt = QH_TAKE(l * l * sizeof(complex double));
	void *svd_ws = ws ? QH_TAKE(qh_jacobi_svd_ws_bytes((int)m, (int)l)) : NULL;
	qh_jacobi_svd((int)m, (int)l, (const double *)dA0, (double *)dUt, sigma, (double *)dVt, stream, svd_ws);
	size_t K = 0;
	for (size_t k = 0; k < l; ++k)
		if (k < rank_sel_min || sigma[k] > rank_tol) ++K;
	if (K == 0) { QH_GIVE(dUt); QH_GIVE(dVt); return 0; }
	QH_ENSURE(K <= 64, "Beyn: %zu candidate eigenvalues exceed the reduced-problem limit of 64", K);
	/* TM2 = V0^H A1 (K x l) */
	complex double *dT = QH_TAKE(K * l * sizeof(complex double));
	qh_ctrans_gemm((int)m, (int)l, (int)K, (const double *)dUt, (const double *)dA1, (double *)dT, stream);
	/* B = TM2 W0 Sigma^{-1}, W0 column j = row j of Vt */
	double *dsig = QH_TAKE(l * sizeof(double));
	qh_h2d(dsig, sigma, l * sizeof(double), stream);
	complex double *dB = QH_TAKE((K * K + K + K * K) * sizeof(complex double));
	complex double *dlam = dB + K * K, *dS = dlam + K;
	qh_bmat((int)K, (int)l, (const double *)dT, (const double *)dVt, dsig, (double *)dB, stream);
	void *eig_ws = ws ? QH_TAKE(256) : NULL;
	int fail = qh_small_eig((int)K, (const double *)dB, (double *)dlam, (double *)dS, stream, eig_ws);
	QH_ENSURE(!fail, "Beyn: the reduced eigenproblem did not converge");
	complex double *lam;
	QH_MALLOC(lam, K * sizeof(complex double));
	qh_d2h(lam, dlam, K * sizeof(complex double), stream);
	complex double *dY = NULL, *Y = NULL;
	if (eigvec || res_tol > 0) {
		dY = QH_TAKE(K * m * sizeof(complex double));
		qh_v0s((int)m, (int)K, (const double *)dS, (const double *)dUt, (double *)dY, stream);
		QH_MALLOC(Y, K * m * sizeof(complex double));
		qh_d2h(Y, dY, K * m * sizeof(complex double), stream);
	}
	qh_sync(stream);
	size_t kept = 0;
	complex double *Mm = NULL, *dMm = NULL, *dr = NULL;
	if (res_tol > 0) {
		QH_MALLOC(Mm, m * m * sizeof(complex double));
		dMm = qh_dmalloc(m * m * sizeof(complex double));
		dr = qh_dmalloc(m * sizeof(complex double));
	}
	for (size_t k = 0; k < K; ++k) {
		const complex double z = z0 + lam[k];
		double residual = 0;
		if (res_tol > 0) {
			/* ||M(z) v|| on the GPU: r = 0 - M v */
			QH_ENSURE(M(Mm, m, z, params) == 0, "Beyn: M callback failed");
			qh_h2d(dMm, Mm, m * m * sizeof(complex double), stream);
			qh_memset(dr, 0, m * sizeof(complex double), stream);
			qh_gemm_sub_f64((int)m, 1, (int)m, 1, (const double *)dMm, (long)m, 0,
				(const double *)(dY + k * m), 1, 0, (double *)dr, 1, 0, stream);
			residual = qh_norm2_c64((long)m, (const double *)dr, stream);
			if (residual > res_tol) continue;
		}
		eigval[kept] = z;
		if (eigvec) {
			memcpy(eigvec + kept * m, Y + k * m, m * sizeof(complex double));
			residuals[kept] = residual;
		}
		++kept;
	}
	qh_dfree(dMm); qh_dfree(dr);
	QH_GIVE(dsig);
	free(Mm); free(Y); free(lam);
	if (dY) QH_GIVE(dY);
	QH_GIVE(dB); QH_GIVE(dT); QH_GIVE(dUt); QH_GIVE(dVt);
	#undef QH_TAKE
	#undef QH_GIVE
	return (int)kept;
}

/* Device scratch of the tail without residuals, for callers that cannot
 * allocate while other work runs (the serving loop). */
size_t qh_beyn_ws_bytes(size_t m, size_t l) {
	const size_t a = 255;
	#define R256(b) ((((size_t)(b)) + a) & ~a)
	return R256(m * l * 16) + R256(l * l * 16) + R256(qh_jacobi_svd_ws_bytes((int)m, (int)l)) + R256(l * l * 16)
		+ R256(l * 8) + R256((2 * l * l + l) * 16) + 256 + R256(l * m * 16) + 1024;
	#undef R256
}

/* Serving-loop entry: eigenvalues only, no residual filter. */
int qh_beyn_process_s(size_t m, size_t l, const complex double *dA0, const complex double *dA1, complex double z0,
		double rank_tol, size_t rank_sel_min, complex double *eigval, double *sigma, void *stream, void *ws) {
	return qh_beyn_process(m, l, NULL, NULL, dA0, dA1, z0, rank_tol, rank_sel_min, 0.0, eigval, NULL, NULL, sigma,
		stream, ws);
}

beyn_result_t *beyn_solve_v(size_t m, size_t l, beyn_function_M_t M, beyn_function_M_inv_Vhat_t M_inv_Vhat,
		void *params, const beyn_contour_t *contour, double rank_tol, size_t rank_sel_min,
		double res_tol, const complex double *Vhat) {
	QH_ENSURE(l <= m, "We expect L <= M, but we got L = %zu, M = %zu ", l, m);
	QH_ENSURE(l <= 64 || l == 0, "this implementation supports at most 64 probe columns (l = %zu)", l);
	const size_t N = contour->n;
	const complex double z0 = contour->centre;
	const int mixed = getenv("QPMS_HIP_LU_MODE") && !strcmp(getenv("QPMS_HIP_LU_MODE"), "mixed");
	const size_t ml = m * l;
	complex double *dAll = qh_dmalloc(6 * ml * sizeof(complex double));
	complex double *dV = dAll, *dX = dV + ml, *dA0 = dX + ml, *dA1 = dA0 + ml, *dA0c = dA1 + ml, *dA1c = dA0c + ml;
	qh_memset(dA0, 0, 4 * ml * sizeof(complex double), NULL);
	qh_h2d(dV, Vhat, ml * sizeof(complex double), NULL);
	complex double *Mh = NULL, *dM = NULL, *Xh = NULL;
	int *dpiv = NULL, *dperm = NULL, *dinfo = NULL;
	complex double *dwork = NULL;
	if (!M_inv_Vhat) {
		QH_MALLOC(Mh, m * m * sizeof(complex double));
		dM = qh_dmalloc(m * m * sizeof(complex double));
		dpiv = qh_dmalloc(m * sizeof(int));
		dperm = qh_dmalloc(m * sizeof(int));
		dinfo = qh_dmalloc(sizeof(int));
		dwork = qh_dmalloc(ml * sizeof(complex double));
	} else {
		QH_MALLOC(Xh, ml * sizeof(complex double));
	}
	for (size_t n = 0; n < N; ++n) {
		const complex double z = contour->z_dz[n][0], dz = contour->z_dz[n][1];
		if (M_inv_Vhat) {
			QH_ENSURE(M_inv_Vhat(Xh, m, l, Vhat, z, params) == 0, "Beyn: M_inv_Vhat callback failed");
			qh_h2d(dX, Xh, ml * sizeof(complex double), NULL);
		} else {
			QH_ENSURE(M(Mh, m, z, params) == 0, "Beyn: M callback failed");
			qh_h2d(dM, Mh, m * m * sizeof(complex double), NULL);
			if (mixed) {
				qpms_hip_solve_mixed_dev(m, l, dM, dV, dX, 1e-17, 30, NULL);   /* refine to the FP64 floor */
			} else {
				qh_memset(dinfo, 0, sizeof(int), NULL);
				qh_getrf_batched_f64((int)m, 1, (double *)dM, 0, dpiv, dperm, dinfo, NULL);
				qh_copy_c64((long)ml, (const double *)dV, (double *)dX, NULL);
				qh_getrs_batched_f64((int)m, (int)l, 1, (const double *)dM, 0, dperm, (double *)dX, 0,
					(double *)dwork, NULL);
			}
		}
		const complex double w = z - z0;
		qh_beyn_accum((long)ml, (const double *)dX, creal(dz), cimag(dz), creal(w), cimag(w), (n % 2) == 0,
			(double *)dA0, (double *)dA1, (double *)dA0c, (double *)dA1c, NULL);
		qh_sync(NULL);   /* Mh / Xh are reused by the next callback */
	}
	beyn_result_t *r;
	QH_CALLOC(r, 1, sizeof(*r));
	r->vlen = m;
	QH_CALLOC(r->eigval, l ? l : 1, sizeof(complex double));
	QH_CALLOC(r->eigval_err, l ? l : 1, sizeof(complex double));
	QH_CALLOC(r->residuals, l ? l : 1, sizeof(double));
	QH_CALLOC(r->eigvec, (l ? l : 1) * m, sizeof(complex double));
	QH_CALLOC(r->ranktest_SV, l ? l : 1, sizeof(double));
	double *sv_coarse;
	QH_MALLOC(sv_coarse, (l ? l : 1) * sizeof(double));
	/* coarse pass first (its values become the error estimate), then fine */
	qh_beyn_process(m, l, M, params, dA0c, dA1c, z0, rank_tol, rank_sel_min, res_tol,
		r->eigval_err, NULL, NULL, sv_coarse, NULL, NULL);
	const int K = qh_beyn_process(m, l, M, params, dA0, dA1, z0, rank_tol, rank_sel_min, res_tol,
		r->eigval, r->eigvec, r->residuals, r->ranktest_SV, NULL, NULL);
	for (size_t i = 0; i < l; ++i) r->eigval_err[i] -= r->eigval[i];   /* beyn.c:592 */
	r->neig = (size_t)K;
	free(sv_coarse);
	free(Mh); free(Xh);
	qh_dfree(dM); qh_dfree(dpiv); qh_dfree(dperm); qh_dfree(dinfo); qh_dfree(dwork);
	qh_dfree(dAll);
	return r;
}

beyn_result_t *beyn_solve(size_t m, size_t l, beyn_function_M_t M, beyn_function_M_inv_Vhat_t M_inv_Vhat,
		void *params, const beyn_contour_t *contour, double rank_tol, size_t rank_sel_min, double res_tol) {
	const char *s = getenv("QPMS_HIP_BEYN_SEED");
	uint64_t st = s ? strtoull(s, NULL, 10) : 1;
	if (st == 0) st = 1;
	st = st * 0x9E3779B97F4A7C15ULL + 0x2545F4914F6CDD1DULL;
	complex double *V;
	QH_MALLOC(V, m * l * sizeof(complex double));
	for (size_t i = 0; i < m * l; ++i) {
		const double re = qh_randn(&st), im = qh_randn(&st);
		V[i] = re + I * im;
	}
	beyn_result_t *r = beyn_solve_v(m, l, M, M_inv_Vhat, params, contour, rank_tol, rank_sel_min, res_tol, V);
	free(V);
	return r;
}

/* ---------------------------------------------------------------------- */
/* Scattering-system wrapper (scatsystem.c:2281-2325)                        */
/* ---------------------------------------------------------------------- */
struct qh_finite_beyn_param { const qpms_scatsys_t *ss; };

static int qh_finite_eval_ImTS(complex double *target, size_t m, complex double omega, void *params) {
	const struct qh_finite_beyn_param *p = params;
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_at_omega(p->ss, omega);
	QH_ENSURE(m == p->ss->fecv_size, "matrix size mismatch");
	qpms_scatsysw_build_modeproblem_matrix_full(target, ssw);
	qpms_scatsys_at_omega_free(ssw);
	return 0;
}

beyn_result_t *qpms_scatsys_finite_find_eigenmodes(const qpms_scatsys_t *ss, qpms_iri_t iri,
		complex double omega_centre, double omega_rr, double omega_ri, size_t contour_npoints,
		double rank_tol, size_t rank_sel_min, double res_tol) {
	QH_ENSURE(ss->lattice_dimension == 0,
		"This method is applicable only to nonperiodic systems. Use qpms_scatsys_periodic_find_eigenmodes() instead.");
	QH_ENSURE(iri == QPMS_NO_IRREP, "irrep-packed eigenmode search is not supported (full system only)");
	const size_t n = ss->fecv_size;
	const char *sl = getenv("QPMS_HIP_BEYN_L");
	size_t l = sl ? (size_t)strtoul(sl, NULL, 10) : 64;
	if (l > n) l = n;
	if (l > 64) l = 64;
	beyn_contour_t *contour = beyn_contour_ellipse(omega_centre, omega_rr, omega_ri, contour_npoints);
	struct qh_finite_beyn_param p = {ss};
	beyn_result_t *r = beyn_solve(n, l, qh_finite_eval_ImTS, NULL, &p, contour, rank_tol, rank_sel_min, res_tol);
	free(contour);
	return r;
}

/* Periodic wrapper (scatsystem.c:2327-2370): M(omega) = I - T(omega) W(omega, k)
 * through the Ewald builder at the adjusted eta. */
struct qh_periodic_beyn_param { const qpms_scatsys_t *ss; const double *k; };

static int qh_periodic_eval_ImTW(complex double *target, size_t m, complex double omega, void *params) {
	const struct qh_periodic_beyn_param *p = params;
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_at_omega(p->ss, omega);
	QH_ENSURE(m == p->ss->fecv_size, "matrix size mismatch");
	qpms_scatsys_at_omega_k_t sswk;
	sswk.ssw = ssw;
	sswk.k[0] = p->k[0]; sswk.k[1] = p->k[1]; sswk.k[2] = p->k[2];
	sswk.eta = qpms_ss_adjusted_eta(p->ss, ssw->wavenumber, p->k);
	qpms_scatsyswk_build_modeproblem_matrix_full(target, &sswk);
	qpms_scatsys_at_omega_free(ssw);
	return 0;
}

beyn_result_t *qpms_scatsys_periodic_find_eigenmodes(const qpms_scatsys_t *ss, const double k[3],
		complex double omega_centre, double omega_rr, double omega_ri, size_t contour_npoints,
		double rank_tol, size_t rank_sel_min, double res_tol) {
	QH_ENSURE(ss->lattice_dimension > 0, "qpms_scatsys_finite_find_eigenmodes(): this method is applicable only to periodic systems.");
	const size_t n = ss->fecv_size;
	const char *sl = getenv("QPMS_HIP_BEYN_L");
	size_t l = sl ? (size_t)strtoul(sl, NULL, 10) : 64;
	if (l > n) l = n;
	if (l > 64) l = 64;
	beyn_contour_t *contour = beyn_contour_ellipse(omega_centre, omega_rr, omega_ri, contour_npoints);
	struct qh_periodic_beyn_param p = {ss, k};
	beyn_result_t *r = beyn_solve(n, l, qh_periodic_eval_ImTW, NULL, &p, contour, rank_tol, rank_sel_min, res_tol);
	free(contour);
	return r;
}
