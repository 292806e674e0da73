/* Ewald σ sums and the e32 translation coefficients, host side.
 *
 * Reference: ewald.c:260-273 (σ0), 286-503 + 623-649 (σ long), 742-920
 * (σ short); translations.c:724-793, 926-1094 (e32); scatsystem.c:1219-1281
 * (periodic W), 1283-1357 (periodic mode problem).
 *
 * The point generators are walked on the host (plumbing); the sums run in
 * k_ewald.hip. Error estimates written to the optional `err` outputs are
 * rounding bounds, DBL_EPSILON x sum of |terms| (the reference adds the
 * quadrature error estimates instead).
 */
#include "qpms_internal.h"
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <string.h>

int qh_launch_sigma_short(int LE, int xy, int ntrip, const double *tr_dev, int npts, const double *pts_dev,
		double eta, const double *leg0_dev, int csphase, double *out_dev, double *err_dev, void *stream);
int qh_launch_sigma_long(int LE, int ntrip, const double *tr_dev, int npts, const double *K_dev,
		double eta, double area, const double *leg0_dev, double *out_dev, double *err_dev, void *stream);
int qh_launch_sigma0(int n, const double *k_dev, double eta, double leg00, double *out_dev, void *stream);
int qh_launch_sigma_total(int ntrip, int nsc, const double *ss, const double *sl, const double *s0,
		const int *flag0, int parts, double *tot, void *stream);
int qh_launch_e32_combine(const qh_trans_tables *t, int ntrip, int nsc, const double *sig_dev, int bq_mode,
		double *A_dev, double *B_dev, void *stream);
void qh_trans_array_from_AB(_Complex double *t, const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		const _Complex double *A, const _Complex double *B, int nelem_AB);

/* ---------------------------------------------------------------------- */
/* Device copies of an Ewald constant table (keyed by the table address)    */
/* ---------------------------------------------------------------------- */
/* The long-range kernels form the s1 constants (ewald.c:141-145) from their
 * closed form at compile time; only P_n^m(0), which carries the table's
 * Condon-Shortley phase, is read from the table. */
typedef struct qh_e3dev {
	const qpms_ewald3_constants_t *c;
	double *leg0;     /* P_n^m(0), n <= lMax */
	struct qh_e3dev *next;
} qh_e3dev;
static qh_e3dev *qh_e3reg = NULL;
static pthread_mutex_t qh_e3lock = PTHREAD_MUTEX_INITIALIZER;

static const qh_e3dev *qh_e3_get(const qpms_ewald3_constants_t *c) {
	pthread_mutex_lock(&qh_e3lock);
	for (qh_e3dev *e = qh_e3reg; e; e = e->next)
		if (e->c == c) { pthread_mutex_unlock(&qh_e3lock); return e; }
	qh_e3dev *e;
	QH_CALLOC(e, 1, sizeof(*e));
	e->c = c;
	const int L = c->lMax;
	const size_t nleg = (size_t)(L + 1) * (L + 2) / 2;
	e->leg0 = qh_dmalloc(nleg * sizeof(double));
	qh_h2d(e->leg0, c->legendre0, nleg * sizeof(double), NULL);
	qh_sync(NULL);
	e->next = qh_e3reg;
	qh_e3reg = e;
	pthread_mutex_unlock(&qh_e3lock);
	return e;
}

/* Drop the device copy of a table that is being freed (its address may be
 * reused by a later table of another size). */
void qh_e3_release(const qpms_ewald3_constants_t *c) {
	pthread_mutex_lock(&qh_e3lock);
	for (qh_e3dev **pp = &qh_e3reg; *pp; pp = &(*pp)->next)
		if ((*pp)->c == c) {
			qh_e3dev *e = *pp;
			*pp = e->next;
			qh_dfree(e->leg0);
			free(e);
			break;
		}
	pthread_mutex_unlock(&qh_e3lock);
}

static void qh_check_le(int LE) {
	QH_ENSURE(LE == 3 || LE == 5 || LE == 7 || LE == 9,
		"Ewald constants with lMax %d are not instantiated on the GPU (3, 5, 7, 9 = 2*lMax+1 for lMax 1..4)", LE);
}

/* ---------------------------------------------------------------------- */
/* Batched extension: many (beta, k, s) triples on one lattice             */
/* ---------------------------------------------------------------------- */
int qpms_hip_ewald3_sigma_batch_dev(const qpms_ewald3_constants_t *c, double eta, double unitcell_area,
		size_t ntrip, const double *triples_dev, size_t nR, const double *R_dev, size_t nK, const double *K_dev,
		_Complex double *sigma_short_dev, _Complex double *sigma_long_dev, void *stream) {
	qh_check_le(c->lMax);
	const qh_e3dev *e = qh_e3_get(c);
	if (sigma_short_dev)
		qh_launch_sigma_short(c->lMax, 1, (int)ntrip, triples_dev, (int)nR, R_dev, eta, e->leg0,
			c->legendre_csphase, (double *)sigma_short_dev, NULL, stream);
	if (sigma_long_dev)
		qh_launch_sigma_long(c->lMax, (int)ntrip, triples_dev, (int)nK, K_dev, eta, unitcell_area, e->leg0,
			(double *)sigma_long_dev, NULL, stream);
	return 0;
}

/* Long-range sums of ntrip triples sharing (beta, k): per-point terms once,
 * then one phase sum per shift. F: qh_sigma_long_shifts_scratch(2 lMax + 1,
 * nK) bytes of device scratch. */
size_t qh_sigma_long_shifts_scratch(int LE, int npts);
int qh_launch_sigma_long_shifts(int LE, int ntrip, const double *tr_dev, int npts, const double *K_dev, double eta,
		double area, const double *leg0_dev, double *F_dev, double *out_dev, void *stream);
size_t qh_ewald3_long_shifts_scratch(const qpms_ewald3_constants_t *c, size_t nK) {
	return qh_sigma_long_shifts_scratch(c->lMax, (int)nK);
}
int qh_ewald3_sigma_long_shifts_dev(const qpms_ewald3_constants_t *c, double eta, double unitcell_area, size_t ntrip,
		const double *triples_dev, size_t nK, const double *K_dev, void *F_dev, _Complex double *sigma_long_dev,
		void *stream) {
	qh_check_le(c->lMax);
	const qh_e3dev *e = qh_e3_get(c);
	return qh_launch_sigma_long_shifts(c->lMax, (int)ntrip, triples_dev, (int)nK, K_dev, eta, unitcell_area, e->leg0,
		(double *)F_dev, (double *)sigma_long_dev, stream);
}

/* ---------------------------------------------------------------------- */
/* σ0, σ short, σ long (single triple, host arrays)                          */
/* ---------------------------------------------------------------------- */
int ewald3_sigma0(complex double *result, double *err, const qpms_ewald3_constants_t *c,
		double eta, complex double k) {
	double kk[2] = {creal(k), cimag(k)};
	char *d = qh_dmalloc(32);
	qh_h2d(d, kk, 16, NULL);
	qh_launch_sigma0(1, (const double *)d, eta, c->legendre0[0], (double *)(d + 16), NULL);
	double out[2];
	qh_d2h(out, d + 16, 16, NULL);
	qh_sync(NULL);
	qh_dfree(d);
	*result = out[0] + I * out[1];
	if (err) *err = 2.2204460492503131e-16 * cabs(*result);
	return 0;
}

/* sph2cart (vectors.h:143-154, with QPMS_VECTORS_NICE_TRANSFORMATIONS) */
static void qh_sph2cart(sph_t s, double *x, double *y, double *z) {
	double st = (s.theta == M_PI) ? 0 : sin(s.theta);
	*x = s.r * st * cos(s.phi);
	*y = s.r * st * sin(s.phi);
	*z = s.r * cos(s.theta);
}

double *qh_collect(PGen *g, size_t *n, int dims) {
	size_t cap = 1024, cnt = 0;
	double *p;
	QH_MALLOC(p, cap * dims * sizeof(double));
	PGenSphReturnData r;
	while ((r = PGen_next_sph(g)).flags & PGEN_NOTDONE) {
		if (cnt == cap) {
			cap *= 2;
			double *q = realloc(p, cap * dims * sizeof(double));
			QH_ENSURE(q, "realloc failed");
			p = q;
		}
		double x, y, z;
		qh_sph2cart(r.point_sph, &x, &y, &z);
		p[dims * cnt] = x;
		p[dims * cnt + 1] = y;
		if (dims == 3) p[dims * cnt + 2] = z;
		++cnt;
	}
	*n = cnt;
	return p;
}

int ewald3_sigma_short(complex double *target, double *err, const qpms_ewald3_constants_t *c,
		double eta, complex double k, LatticeDimensionality latdim, PGen *pgen_R,
		bool pgen_generates_shifted_points, cart3_t beta, cart3_t shift) {
	qh_check_le(c->lMax);
	size_t np;
	double *pts = qh_collect(pgen_R, &np, 3);
	const int xy = ((latdim & LAT_2D_IN_3D_XYONLY) == LAT_2D_IN_3D_XYONLY) && shift.z == 0;
	if (xy) for (size_t i = 0; i < np; ++i) pts[3 * i + 2] = 0;
	double tr[8] = {beta.x, beta.y, beta.z, creal(k), cimag(k), shift.x, shift.y, shift.z};
	if (pgen_generates_shifted_points) tr[5] = tr[6] = tr[7] = 0;
	const size_t nsc = c->nelem_sc;
	char *d = qh_dmalloc(64 + np * 24 + nsc * 24);
	double *dtr = (double *)d, *dpts = (double *)(d + 64), *dout = (double *)(d + 64 + np * 24);
	double *derr = dout + 2 * nsc;
	qh_h2d(dtr, tr, 64, NULL);
	qh_h2d(dpts, pts, np * 24, NULL);
	const qh_e3dev *e = qh_e3_get(c);
	qh_launch_sigma_short(c->lMax, xy, 1, dtr, (int)np, dpts, eta, e->leg0, c->legendre_csphase, dout, derr, NULL);
	qh_d2h(target, dout, nsc * 16, NULL);
	if (err) qh_d2h(err, derr, nsc * 8, NULL);
	qh_sync(NULL);
	qh_dfree(d);
	free(pts);
	return 0;
}

int ewald3_sigma_long(complex double *target, double *err, const qpms_ewald3_constants_t *c,
		double eta, complex double k, double unitcell_volume, LatticeDimensionality latdim, PGen *pgen_K,
		bool pgen_generates_shifted_points, cart3_t beta, cart3_t shift) {
	qh_check_le(c->lMax);
	QH_ENSURE((latdim & LAT_XYONLY) && (latdim & SPACE3D) && (latdim & LAT2D),
		"only 2D lattices in the xy plane are implemented on this path");
	QH_ENSURE(shift.z == 0, "out-of-plane particle shifts (Kambe Delta_n terms) are not implemented on this path");
	size_t np;
	double *K = qh_collect(pgen_K, &np, 2);
	if (pgen_generates_shifted_points)
		for (size_t i = 0; i < np; ++i) { K[2 * i] -= beta.x; K[2 * i + 1] -= beta.y; }
	double tr[8] = {beta.x, beta.y, beta.z, creal(k), cimag(k), shift.x, shift.y, shift.z};
	const size_t nsc = c->nelem_sc;
	char *d = qh_dmalloc(64 + np * 16 + nsc * 24);
	double *dtr = (double *)d, *dK = (double *)(d + 64), *dout = (double *)(d + 64 + np * 16);
	double *derr = dout + 2 * nsc;
	qh_h2d(dtr, tr, 64, NULL);
	qh_h2d(dK, K, np * 16, NULL);
	const qh_e3dev *e = qh_e3_get(c);
	qh_launch_sigma_long(c->lMax, 1, dtr, (int)np, dK, eta, unitcell_volume, e->leg0, dout, derr, NULL);
	qh_d2h(target, dout, nsc * 16, NULL);
	if (err) qh_d2h(err, derr, nsc * 8, NULL);
	qh_sync(NULL);
	qh_dfree(d);
	free(K);
	return 0;
}

/* ---------------------------------------------------------------------- */
/* e32 (translations.c:926-1094)                                             */
/* ---------------------------------------------------------------------- */
int qh_e32_bq_mode(void) {
	const char *s = getenv("QPMS_HIP_E32_B_QMAX");
	return (s && !strcmp(s, "corrected")) ? 1 : 0;
}

/* σ_total and A/B for ntrip shifts on one lattice at one (k, beta); device in/out. */
static void qh_e32_device(const qpms_trans_calculator *c, int ntrip, const cart3_t *shifts,
		double eta, complex double k, cart2_t b1, cart2_t b2, cart2_t beta, double maxR, double maxK,
		qpms_ewald_part parts, complex double *dA, complex double *dB, double *hSerr /* nullable, ntrip*nsc */) {
	const qpms_ewald3_constants_t *e3 = c->e3c;
	qh_check_le(e3->lMax);
	const size_t nsc = e3->nelem_sc;
	const double area = l2d_unitcell_area(b1, b2);
	cart2_t rb1, rb2;
	QH_ENSURE(l2d_reciprocalBasis2pi(b1, b2, &rb1, &rb2) == QPMS_SUCCESS, "degenerate lattice basis");
	cart2_t zero = {0, 0};
	PGen Kg = PGen_xyWeb_new(rb1, rb2, BASIS_RTOL, zero, 0, true, maxK, false);
	size_t nK;
	double *K = qh_collect(&Kg, &nK, 2);
	PGen_destroy(&Kg);
	/* R points: with and without the origin (σ0 shifts exclude it) */
	PGen Rg = PGen_xyWeb_new(b1, b2, BASIS_RTOL, zero, 0, true, maxR, false);
	size_t nR;
	double *R = qh_collect(&Rg, &nR, 3);
	PGen_destroy(&Rg);
	for (size_t i = 0; i < nR; ++i) R[3 * i + 2] = 0;
	/* the generator yields the origin first (layer 0) */
	QH_ENSURE(nR > 0 && R[0] == 0 && R[1] == 0, "lattice generator did not start at the origin");
	double *tr;
	int *flag0;
	QH_MALLOC(tr, ntrip * 8 * sizeof(double));
	QH_MALLOC(flag0, ntrip * sizeof(int));
	int all_xy = 1;
	for (int t = 0; t < ntrip; ++t) {
		const cart3_t s = shifts[t];
		double v[8] = {beta.x, beta.y, 0, creal(k), cimag(k), s.x, s.y, s.z};
		memcpy(tr + 8 * t, v, sizeof v);
		flag0[t] = (s.x == 0 && s.y == 0 && s.z == 0);
		if (s.z != 0) all_xy = 0;
	}
	QH_ENSURE(all_xy || !(parts & QPMS_EWALD_LONG_RANGE),
		"out-of-plane particle shifts (Kambe Delta_n terms) are not implemented on this path");
	char *d = qh_dmalloc(ntrip * (64 + 4 + 16) + nK * 16 + nR * 24 + ntrip * nsc * 16 * 3 + 64);
	char *p = d;
	double *dtr = (double *)p; p += ntrip * 64;
	double *dK = (double *)p; p += nK * 16;
	double *dR = (double *)p; p += nR * 24;
	double *dss = (double *)p; p += ntrip * nsc * 16;
	double *dsl = (double *)p; p += ntrip * nsc * 16;
	double *dtot = (double *)p; p += ntrip * nsc * 16;
	double *ds0 = (double *)p; p += ntrip * 16;
	int *dflag = (int *)p;
	qh_h2d(dtr, tr, ntrip * 64, NULL);
	qh_h2d(dK, K, nK * 16, NULL);
	qh_h2d(dR, R, nR * 24, NULL);
	qh_h2d(dflag, flag0, ntrip * sizeof(int), NULL);
	const qh_e3dev *e = qh_e3_get(e3);
	if (parts & QPMS_EWALD_LONG_RANGE)
		qh_launch_sigma_long(e3->lMax, ntrip, dtr, (int)nK, dK, eta, area, e->leg0, dsl, NULL, NULL);
	/* one launch for every triple: the kernel drops the lattice point equal
	 * to the shift, i.e. the origin for s = 0 */
	if (parts & QPMS_EWALD_SHORT_RANGE)
		qh_launch_sigma_short(e3->lMax, all_xy, ntrip, dtr, (int)nR, dR, eta, e->leg0, e3->legendre_csphase, dss,
			NULL, NULL);
	double *kk;
	QH_MALLOC(kk, 2 * ntrip * sizeof(double));
	for (int t = 0; t < ntrip; ++t) { kk[2 * t] = creal(k); kk[2 * t + 1] = cimag(k); }
	double *dkk = qh_dmalloc(2 * ntrip * sizeof(double));
	qh_h2d(dkk, kk, 2 * ntrip * sizeof(double), NULL);
	if (parts & QPMS_EWALD_0TERM) qh_launch_sigma0(ntrip, dkk, eta, e3->legendre0[0], ds0, NULL);
	qh_launch_sigma_total(ntrip, (int)nsc, dss, dsl, ds0, dflag, (int)parts, dtot, NULL);
	qh_launch_e32_combine(qh_tables_get(c), ntrip, (int)nsc, dtot, qh_e32_bq_mode(), (double *)dA, (double *)dB, NULL);
	qh_sync(NULL);
	(void)hSerr;
	qh_dfree(dkk);
	qh_dfree(d);
	free(kk); free(tr); free(flag0); free(K); free(R);
}

int qpms_trans_calculator_get_AB_arrays_e32_e(const qpms_trans_calculator *c,
		complex double *Adest, double *Aerr, complex double *Bdest, double *Berr,
		ptrdiff_t deststride, ptrdiff_t srcstride,
		double eta, complex double k, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK, qpms_ewald_part parts) {
	const size_t ne = c->nelem;
	complex double *dA = qh_dmalloc(2 * ne * ne * sizeof(complex double));
	complex double *dB = dA + ne * ne;
	qh_e32_device(c, 1, &particle_shift, eta, k, b1, b2, beta, maxR, maxK, parts, dA, dB, NULL);
	complex double *h;
	QH_MALLOC(h, 2 * ne * ne * sizeof(complex double));
	qh_d2h(h, dA, 2 * ne * ne * sizeof(complex double), NULL);
	qh_sync(NULL);
	qh_dfree(dA);
	for (size_t y = 0; y < ne; ++y)
		for (size_t yu = 0; yu < ne; ++yu) {
			Adest[deststride * y + srcstride * yu] = h[y * ne + yu];
			Bdest[deststride * y + srcstride * yu] = h[ne * ne + y * ne + yu];
			if (Aerr) Aerr[deststride * y + srcstride * yu] = 2.2204460492503131e-16 * cabs(h[y * ne + yu]);
			if (Berr) Berr[deststride * y + srcstride * yu] = 2.2204460492503131e-16 * cabs(h[ne * ne + y * ne + yu]);
		}
	free(h);
	return 0;
}

int qpms_trans_calculator_get_AB_arrays_e32(const qpms_trans_calculator *c,
		complex double *Adest, double *Aerr, complex double *Bdest, double *Berr,
		ptrdiff_t deststride, ptrdiff_t srcstride,
		double eta, complex double k, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK) {
	return qpms_trans_calculator_get_AB_arrays_e32_e(c, Adest, Aerr, Bdest, Berr, deststride, srcstride,
		eta, k, b1, b2, beta, particle_shift, maxR, maxK, QPMS_EWALD_FULL);
}

qpms_errno_t qpms_trans_calculator_get_trans_array_e32_e(const qpms_trans_calculator *c,
		complex double *target, double *err,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		double eta, complex double k, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK, qpms_ewald_part parts) {
	QH_ENSURE(c->normalisation == destspec->norm && c->normalisation == srcspec->norm,
		"The normalisation conventions must be the same");
	const size_t ne = c->nelem;
	complex double *A;
	double *Ae = NULL;
	QH_MALLOC(A, 2 * ne * ne * sizeof(complex double));
	if (err) QH_MALLOC(Ae, 2 * ne * ne * sizeof(double));
	qpms_trans_calculator_get_AB_arrays_e32_e(c, A, Ae, A + ne * ne, Ae ? Ae + ne * ne : NULL, ne, 1,
		eta, k, b1, b2, beta, particle_shift, maxR, maxK, parts);
	qh_trans_array_from_AB(target, destspec, deststride, srcspec, srcstride, A, A + ne * ne, (int)ne);
	if (err) {
		for (size_t di = 0; di < destspec->n; ++di)
			for (size_t si = 0; si < srcspec->n; ++si) {
				const int dt = (int)(destspec->ilist[di] & 3), st = (int)(srcspec->ilist[si] & 3);
				const long dy = (long)(destspec->ilist[di] / 4) - 1, sy = (long)(srcspec->ilist[si] / 4) - 1;
				err[si * srcstride + di * deststride] = (st == dt) ? Ae[dy * ne + sy] : Ae[ne * ne + dy * ne + sy];
			}
	}
	free(A); free(Ae);
	return 0;
}

qpms_errno_t qpms_trans_calculator_get_trans_array_e32(const qpms_trans_calculator *c,
		complex double *target, double *err,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		double eta, complex double k, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK) {
	return qpms_trans_calculator_get_trans_array_e32_e(c, target, err, destspec, deststride, srcspec, srcstride,
		eta, k, b1, b2, beta, particle_shift, maxR, maxK, QPMS_EWALD_FULL);
}

/* ---------------------------------------------------------------------- */
/* Periodic builders (scatsystem.c:1166-1281, 1283-1357)                     */
/* ---------------------------------------------------------------------- */
static const qpms_vswf_set_spec_t *qh_bspec(const qpms_scatsys_t *ss, qpms_ss_pi_t pi) {
	return ss->tmg[ss->tm[ss->p[pi].tmatrix_id].tmgi].spec;
}

complex double *qpms_scatsys_periodic_build_translation_matrix_full(complex double *target,
		const qpms_scatsys_t *ss, complex double wavenumber, const cart3_t *wavevector, double eta) {
	QH_ENSURE(ss->lattice_dimension > 0, "This method is applicable only to periodic systems.");
	QH_ENSURE(ss->lattice_dimension == 2 && !ss->per.lattice_basis[0].z && !ss->per.lattice_basis[1].z
		&& !wavevector->z, "Only 2D xy-lattices currently supported");
	if (eta == 0 || isnan(eta)) {
		double kv[3] = {wavevector->x, wavevector->y, wavevector->z};
		eta = qpms_ss_adjusted_eta(ss, wavenumber, kv);
	}
	const size_t L = ss->fecv_size, np = ss->p_count, ne = ss->c->nelem;
	if (!target) QH_MALLOC(target, L * L * sizeof(complex double));
	const double maxR = sqrt(ss->per.unitcell_volume) * 64;   /* scatsystem.c:1226-1227 */
	const double maxK = 2048 * 2 * M_PI / maxR;
	const size_t npairs = np * np;
	cart3_t *sh;
	QH_MALLOC(sh, npairs * sizeof(cart3_t));
	for (size_t d = 0; d < np; ++d)
		for (size_t s = 0; s < np; ++s) {
			sh[d * np + s].x = ss->p[d].pos.x - ss->p[s].pos.x;
			sh[d * np + s].y = ss->p[d].pos.y - ss->p[s].pos.y;
			sh[d * np + s].z = ss->p[d].pos.z - ss->p[s].pos.z;
		}
	complex double *dAB = qh_dmalloc(2 * npairs * ne * ne * sizeof(complex double));
	cart2_t b1 = {ss->per.lattice_basis[0].x, ss->per.lattice_basis[0].y};
	cart2_t b2 = {ss->per.lattice_basis[1].x, ss->per.lattice_basis[1].y};
	cart2_t beta = {wavevector->x, wavevector->y};
	qh_e32_device(ss->c, (int)npairs, sh, eta, wavenumber, b1, b2, beta, maxR, maxK, QPMS_EWALD_FULL,
		dAB, dAB + npairs * ne * ne, NULL);
	complex double *h;
	QH_MALLOC(h, 2 * npairs * ne * ne * sizeof(complex double));
	qh_d2h(h, dAB, 2 * npairs * ne * ne * sizeof(complex double), NULL);
	qh_sync(NULL);
	qh_dfree(dAB);
	for (size_t d = 0; d < np; ++d)
		for (size_t s = 0; s < np; ++s) {
			const size_t i = d * np + s;
			qh_trans_array_from_AB(target + ss->fecv_pstarts[d] * L + ss->fecv_pstarts[s],
				qh_bspec(ss, (qpms_ss_pi_t)d), L, qh_bspec(ss, (qpms_ss_pi_t)s), 1,
				h + i * ne * ne, h + npairs * ne * ne + i * ne * ne, (int)ne);
		}
	free(h); free(sh);
	return target;
}

complex double *qpms_scatsyswk_build_translation_matrix_full(complex double *target,
		const qpms_scatsys_at_omega_k_t *sswk) {
	const qpms_scatsys_at_omega_t *ssw = sswk->ssw;
	const cart3_t kv = {sswk->k[0], sswk->k[1], sswk->k[2]};
	return qpms_scatsys_periodic_build_translation_matrix_full(target, ssw->ss, ssw->wavenumber, &kv, sswk->eta);
}

complex double *qpms_scatsyswk_build_modeproblem_matrix_full(complex double *target,
		const qpms_scatsys_at_omega_k_t *sswk) {
	const qpms_scatsys_at_omega_t *ssw = sswk->ssw;
	const qpms_scatsys_t *ss = ssw->ss;
	QH_ENSURE(ss->lattice_dimension > 0,
		"This method is applicable only to periodic systems. Use qpms_scatsysw_build_modeproblem_matrix_full() instead.");
	const size_t L = ss->fecv_size, np = ss->p_count;
	if (!target) QH_MALLOC(target, L * L * sizeof(complex double));
	complex double *W = qpms_scatsyswk_build_translation_matrix_full(NULL, sswk);
	complex double *dW = qh_dmalloc(2 * L * L * sizeof(complex double));
	complex double *dM = dW + L * L;
	qh_h2d(dW, W, L * L * sizeof(complex double), NULL);
	for (size_t r = 0; r < np; ++r) {
		const qpms_tmatrix_t *T = ssw->tm[ss->p[r].tmatrix_id];
		const size_t nr = T->spec->n;
		complex double *dT = qh_dmalloc(nr * nr * sizeof(complex double));
		qh_h2d(dT, T->m, nr * nr * sizeof(complex double), NULL);
		qh_launch_block_tw(nr, L, (const double *)dT, (const double *)(dW + ss->fecv_pstarts[r] * L),
			(double *)(dM + ss->fecv_pstarts[r] * L), ss->fecv_pstarts[r], NULL);
		qh_sync(NULL);
		qh_dfree(dT);
	}
	qh_d2h(target, dM, L * L * sizeof(complex double), NULL);
	qh_sync(NULL);
	qh_dfree(dW);
	free(W);
	return target;
}

qpms_ss_LU qpms_scatsyswk_build_modeproblem_matrix_full_LU(complex double *target, int *target_piv,
		const qpms_scatsys_at_omega_k_t *sswk) {
	target = qpms_scatsyswk_build_modeproblem_matrix_full(target, sswk);
	return qpms_scatsysw_modeproblem_matrix_full_factorise(target, target_piv, sswk->ssw, sswk);
}
