/* Translation calculator: init-time coefficient tables (host, FP64) and the
 * AB-array entry points, which run the k_ab_batch kernel on the GPU.
 *
 * Reference: translations.c:117-133 (normalisation factors), 313-368 (A
 * multipliers), 372-437 (B multipliers, Cruzan form), 439-508 (init),
 * 610-698 (AB arrays), 701-812 (trans arrays).
 *
 * The Gaunt coefficients a_q(m,n,mu,nu) (Xu's recursions in the reference,
 * gaunt.c:396) and the Cruzan 3j factors (gsl_sf_coupling_3j) are evaluated
 * here from Racah's closed form for the Wigner 3j symbol in long double,
 * using exact integer factorials; only the ratios a_q / a_0 enter the
 * multipliers.
 */
#include "qpms_internal.h"
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <string.h>

void qpms_hip_die(const char *file, int line, const char *func, const char *fmt, ...) {
	fprintf(stderr, "%s:%d, %s: ", file, line, func);
	va_list ap;
	va_start(ap, fmt);
	vfprintf(stderr, fmt, ap);
	va_end(ap);
	fputc('\n', stderr);
	fflush(stderr);
	abort();
}

/* ---------------------------------------------------------------------- */
/* Exact factorials and the Wigner 3j symbol                               */
/* ---------------------------------------------------------------------- */
static long double qh_fact(int n) {
	if (n < 0) return NAN;
	long double f = 1.0L;
	for (int i = 2; i <= n; ++i) f *= (long double)i;
	return f;
}

static inline int imin2(int a, int b) { return a < b ? a : b; }
static inline int imax2(int a, int b) { return a > b ? a : b; }
static inline int min1pow(int k) { return (k % 2) ? -1 : 1; }
static inline complex double ipow(int k) {
	switch (((k % 4) + 4) % 4) {
		case 0: return 1;
		case 1: return I;
		case 2: return -1;
		default: return -I;
	}
}

/* Racah's formula (e.g. Edmonds, Angular Momentum in QM, eq. 3.7.3). */
static long double qh_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
	if (m1 + m2 + m3 != 0) return 0;
	if (j3 < abs(j1 - j2) || j3 > j1 + j2) return 0;
	if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0;
	long double tri = qh_fact(j1 + j2 - j3) * qh_fact(j1 - j2 + j3) * qh_fact(-j1 + j2 + j3)
		/ qh_fact(j1 + j2 + j3 + 1);
	long double pre = sqrtl(tri * qh_fact(j1 + m1) * qh_fact(j1 - m1) * qh_fact(j2 + m2)
		* qh_fact(j2 - m2) * qh_fact(j3 + m3) * qh_fact(j3 - m3));
	int kmin = imax2(0, imax2(j2 - j3 - m1, j1 - j3 + m2));
	int kmax = imin2(j1 + j2 - j3, imin2(j1 - m1, j2 + m2));
	long double sum = 0;
	for (int k = kmin; k <= kmax; ++k) {
		long double den = qh_fact(k) * qh_fact(j3 - j2 + k + m1) * qh_fact(j3 - j1 + k - m2)
			* qh_fact(j1 + j2 - j3 - k) * qh_fact(j1 - k - m1) * qh_fact(j2 - k + m2);
		sum += (k % 2 ? -1.0L : 1.0L) / den;
	}
	return min1pow(j1 - j2 - m3) * pre * sum;
}

/* Gaunt coefficient a_q(m, n, mu, nu), p = n + nu - 2q (Xu 1996, eq. 4):
 *   a_q = (-1)^{m+mu} (2p+1) sqrt[(n+m)!(nu+mu)!(p-m-mu)! / ((n-m)!(nu-mu)!(p+m+mu)!)]
 *         * (n nu p; m mu -m-mu) (n nu p; 0 0 0). */
static long double qh_gaunt(int m, int n, int mu, int nu, int q) {
	int p = n + nu - 2 * q;
	long double f = sqrtl(qh_fact(n + m) * qh_fact(nu + mu) * qh_fact(p - m - mu)
		/ (qh_fact(n - m) * qh_fact(nu - mu) * qh_fact(p + m + mu)));
	return min1pow(m + mu) * (2 * p + 1) * f * qh_wigner3j(n, nu, p, m, mu, -m - mu)
		* qh_wigner3j(n, nu, p, 0, 0, 0);
}

int qh_gaunt_q_max(int m, int n, int mu, int nu) {  /* gaunt.h:6-8 */
	return imin2(n, imin2(nu, (n + nu - abs(m + mu)) / 2));
}
int qh_gauntB_Q_max(int M, int n, int mu, int nu) { /* translations.c:113-115 */
	return imin2(n, imin2(nu, (n + nu + 1 - abs(M + mu)) / 2));
}

/* ---------------------------------------------------------------------- */
/* Normalisation factors (normalisation.h:18-76, translations.c:117-133)   */
/* ---------------------------------------------------------------------- */
static double qh_normfactor(qpms_normalisation_t norm, int l, int m) {
	switch (norm & QPMS_NORMALISATION_NORM_BITS) {
		case QPMS_NORMALISATION_NORM_POWER: return 1;
		case QPMS_NORMALISATION_NORM_SPHARM: return sqrt(l * (l + 1.0));
		case QPMS_NORMALISATION_NORM_NONE:
			return sqrt(l * (l + 1.0) * 4 * M_PI / (2 * l + 1)) * sqrt((double)(qh_fact(l + m) / qh_fact(l - m)));
		default: QH_DIE("invalid normalisation bits %d", (int)norm);
	}
}
static complex double qh_factor_N_noCS(qpms_normalisation_t norm, int l, int m) {
	complex double f = qh_normfactor(norm, l, m);
	if (norm & QPMS_NORMALISATION_N_MINUS) f *= -1;
	if (norm & QPMS_NORMALISATION_N_I) f *= I;
	if (norm & QPMS_NORMALISATION_INVERSE) f = 1 / f;
	return f;
}
static complex double qh_factor_M_noCS(qpms_normalisation_t norm, int l, int m) {
	complex double f = qh_normfactor(norm, l, m);
	if (norm & QPMS_NORMALISATION_M_MINUS) f *= -1;
	if (norm & QPMS_NORMALISATION_M_I) f *= I;
	if (norm & QPMS_NORMALISATION_INVERSE) f = 1 / f;
	return f;
}
/* exp(normlogfac) = sqrt[(n-m)!(nu+mu)! / ((n+m)!(nu-mu)!)] */
static long double qh_normlogfac_exp(int m, int n, int mu, int nu) {
	return sqrtl(qh_fact(n - m) * qh_fact(nu + mu) / (qh_fact(n + m) * qh_fact(nu - mu)));
}
static double qh_trans_normfac(qpms_normalisation_t norm, int m, int n, int mu, int nu) {
	int csphase = (norm & QPMS_NORMALISATION_CSPHASE) ? -1 : 1;
	double f = (csphase == 1) ? min1pow(m - mu) : 1.;
	f *= sqrt((n * (n + 1.)) / (nu * (nu + 1.)));
	f *= sqrt((2. * n + 1) / (2. * nu + 1));
	return f;
}

/* translations.c:313-368 */
static void qh_multipliers_A(qpms_normalisation_t norm, complex double *dest,
		int m, int n, int mu, int nu, int qmax) {
	long double a0 = qh_gaunt(-m, n, mu, nu, 0);
	complex double normfac = qh_trans_normfac(norm, m, n, mu, nu)
		* qh_factor_N_noCS(norm, nu, mu) / qh_factor_N_noCS(norm, n, m);
	normfac *= min1pow(m);
	long double ex = qh_fact(2 * n) / qh_fact(n + 1) * qh_fact(2 * nu + 2) / qh_fact(nu + 1)
		* qh_fact(n + nu + m - mu) / (qh_fact(n - m) * qh_fact(nu + mu))
		* qh_fact(n + nu) / qh_fact(2 * (n + nu)) * qh_normlogfac_exp(m, n, mu, nu);
	complex double presum = (double)ex * normfac / (4. * n) * ipow(n + nu);
	for (int q = 0; q <= qmax; ++q) {
		int p = n + nu - 2 * q;
		int d = mu - m;
		double ppfac = (d >= 0) ? 1.0
			: min1pow(d) * (double)(qh_fact(p + d) / qh_fact(p - d));
		long double aq = qh_gaunt(-m, n, mu, nu, q);
		double summandfac = (n * (n + 1) + nu * (nu + 1) - p * (p + 1)) * min1pow(q) * (double)(aq / a0);
		dest[q] = presum * summandfac * ppfac * ipow(n - nu);
	}
}

/* translations.c:372-379 */
static double qh_cruzan_bfactor(int M, int n, int mu, int nu, int p) {
	long double pre = sqrtl(qh_fact(n + M) / qh_fact(n - M) * qh_fact(nu + mu) / qh_fact(nu - mu)
		* qh_fact(p - M - mu + 1) / qh_fact(p + M + mu + 1));
	return (double)(min1pow(mu + M) * (2 * p + 3) * pre
		* qh_wigner3j(n, nu, p + 1, M, mu, -M - mu) * qh_wigner3j(n, nu, p, 0, 0, 0));
}

/* translations.c:382-437 */
static void qh_multipliers_B(qpms_normalisation_t norm, complex double *dest,
		int m, int n, int mu, int nu, int Qmax) {
	complex double normfac = qh_trans_normfac(norm, m, n, mu, nu)
		* qh_factor_M_noCS(norm, nu, mu) / qh_factor_N_noCS(norm, n, m);
	double presum = min1pow(1 - m) * (2 * nu + 1) / (2. * (n * (n + 1)))
		* (double)(qh_fact(n + m) / qh_fact(n - m) * qh_fact(nu - mu) / qh_fact(nu + mu)
			* qh_normlogfac_exp(m, n, mu, nu));
	for (int q = 0; q <= Qmax; ++q) {
		int p = n + nu - 2 * q;
		int d = mu - m;
		double ppfac = (d >= 0) ? 1.0
			: min1pow(d) * (double)(qh_fact(p + 1 + d) / qh_fact(p + 1 - d));
		double t = sqrt((double)((p + 1) * (p + 1) - (n - nu) * (n - nu))
			* (double)((n + nu + 1) * (n + nu + 1) - (p + 1) * (p + 1)));
		dest[q] = presum * normfac * t * ppfac * qh_cruzan_bfactor(-m, n, mu, nu, p)
			* ipow(p + 1) * ipow(n - nu);
	}
}

/* Ferrers P_l^m(0) with the Condon–Shortley phase: (-1)^{(l+m)/2} (l+m-1)!!/(l-m)!!
 * for l+m even, 0 otherwise (DLMF 14.5.1). */
static double qh_legendre_at_zero(int l, int m) {
	if ((l + m) % 2) return 0;
	long double num = 1, den = 1;
	for (int k = l + m - 1; k > 1; k -= 2) num *= k;
	for (int k = l - m; k > 1; k -= 2) den *= k;
	return (double)(min1pow((l + m) / 2) * num / den);
}

/* ---------------------------------------------------------------------- */
/* Public: init / free                                                      */
/* ---------------------------------------------------------------------- */
qpms_trans_calculator *qpms_trans_calculator_init(qpms_l_t lMax, qpms_normalisation_t norm) {
	if (norm & (QPMS_NORMALISATION_SPHARM_REAL | QPMS_NORMALISATION_REVERSE_AZIMUTHAL_PHASE))
		QH_DIE("Translation operators for real or inverse complex spherical harmonics based waves are not implemented.");
	QH_ENSURE(lMax > 0, "lMax must be positive, got %d", lMax);
	qpms_trans_calculator *c;
	QH_MALLOC(c, sizeof(*c));
	c->lMax = lMax;
	c->nelem = qpms_lMax2nelem(lMax);
	c->normalisation = norm;
	const size_t ne = c->nelem;
	QH_MALLOC(c->A_multipliers, (1 + ne * ne) * sizeof(complex double *));
	QH_MALLOC(c->B_multipliers, (1 + ne * ne) * sizeof(complex double *));
	size_t sumA = 0, sumB = 0;
	for (size_t y = 0; y < ne; ++y)
		for (size_t yu = 0; yu < ne; ++yu) {
			int n = (int)sqrt((double)y + 1), m = (int)y - n * (n + 1) + 1;
			int nu = (int)sqrt((double)yu + 1), mu = (int)yu - nu * (nu + 1) + 1;
			sumA += 1 + qh_gaunt_q_max(-m, n, mu, nu);
			sumB += 1 + qh_gauntB_Q_max(-m, n, mu, nu);
		}
	QH_MALLOC(c->A_multipliers[0], sumA * sizeof(complex double));
	QH_MALLOC(c->B_multipliers[0], sumB * sizeof(complex double));
	for (size_t y = 0; y < ne; ++y)
		for (size_t yu = 0; yu < ne; ++yu) {
			size_t i = y * ne + yu;
			int n = (int)sqrt((double)y + 1), m = (int)y - n * (n + 1) + 1;
			int nu = (int)sqrt((double)yu + 1), mu = (int)yu - nu * (nu + 1) + 1;
			int qA = qh_gaunt_q_max(-m, n, mu, nu), qB = qh_gauntB_Q_max(-m, n, mu, nu);
			c->A_multipliers[i + 1] = c->A_multipliers[i] + qA + 1;
			c->B_multipliers[i + 1] = c->B_multipliers[i] + qB + 1;
			qh_multipliers_A(norm, c->A_multipliers[i], m, n, mu, nu, qA);
			qh_multipliers_B(norm, c->B_multipliers[i], m, n, mu, nu, qB);
		}
	int csphase = (norm & QPMS_NORMALISATION_CSPHASE) ? -1 : 1;
	c->e3c = qpms_ewald3_constants_init(2 * lMax + 1, csphase);
	const int L = 2 * lMax + 1;
	QH_MALLOC(c->legendre0, (size_t)(L + 1) * (L + 2) / 2 * sizeof(double));
	for (int l = 0; l <= L; ++l)
		for (int m = 0; m <= l; ++m)
			c->legendre0[l * (l + 1) / 2 + m] = qh_legendre_at_zero(l, m);
	return c;
}

void qpms_trans_calculator_free(qpms_trans_calculator *c) {
	if (!c) return;
	qh_tables_release(c);
	free(c->A_multipliers[0]);
	free(c->A_multipliers);
	free(c->B_multipliers[0]);
	free(c->B_multipliers);
	qpms_ewald3_constants_free(c->e3c);
	free(c->legendre0);
	free(c);
}

/* ---------------------------------------------------------------------- */
/* Device table registry                                                    */
/* ---------------------------------------------------------------------- */
typedef struct qh_reg_entry {
	const qpms_trans_calculator *c;
	qh_trans_tables t;
	struct qh_reg_entry *next;
} qh_reg_entry;
static qh_reg_entry *qh_registry = NULL;
static pthread_mutex_t qh_registry_lock = PTHREAD_MUTEX_INITIALIZER;

const qh_trans_tables *qh_tables_get(const qpms_trans_calculator *c) {
	pthread_mutex_lock(&qh_registry_lock);
	for (qh_reg_entry *e = qh_registry; e; e = e->next)
		if (e->c == c) { pthread_mutex_unlock(&qh_registry_lock); return &e->t; }
	qh_reg_entry *e;
	QH_CALLOC(e, 1, sizeof(*e));
	e->c = c;
	const int ne = (int)c->nelem, qs = c->lMax + 1;
	e->t.lMax = c->lMax; e->t.nelem = ne; e->t.qslots = qs;
	size_t nslot = (size_t)ne * ne * qs;
	double *hA, *hB; unsigned char *hqA, *hqB;
	QH_CALLOC(hA, 2 * nslot, sizeof(double));
	QH_CALLOC(hB, 2 * nslot, sizeof(double));
	QH_MALLOC(hqA, (size_t)ne * ne);
	QH_MALLOC(hqB, (size_t)ne * ne);
	for (int i = 0; i < ne * ne; ++i) {
		int nA = (int)(c->A_multipliers[i + 1] - c->A_multipliers[i]);
		int nB = (int)(c->B_multipliers[i + 1] - c->B_multipliers[i]);
		QH_ENSURE(nA <= qs && nB <= qs, "multiplier count exceeds slot count");
		hqA[i] = (unsigned char)nA; hqB[i] = (unsigned char)nB;
		for (int q = 0; q < nA; ++q) {
			hA[2 * ((size_t)i * qs + q)] = creal(c->A_multipliers[i][q]);
			hA[2 * ((size_t)i * qs + q) + 1] = cimag(c->A_multipliers[i][q]);
		}
		for (int q = 0; q < nB; ++q) {
			hB[2 * ((size_t)i * qs + q)] = creal(c->B_multipliers[i][q]);
			hB[2 * ((size_t)i * qs + q) + 1] = cimag(c->B_multipliers[i][q]);
		}
	}
	e->t.dA = qh_dmalloc(2 * nslot * sizeof(double));
	e->t.dB = qh_dmalloc(2 * nslot * sizeof(double));
	e->t.dqA = qh_dmalloc((size_t)ne * ne);
	e->t.dqB = qh_dmalloc((size_t)ne * ne);
	qh_h2d(e->t.dA, hA, 2 * nslot * sizeof(double), NULL);
	qh_h2d(e->t.dB, hB, 2 * nslot * sizeof(double), NULL);
	qh_h2d(e->t.dqA, hqA, (size_t)ne * ne, NULL);
	qh_h2d(e->t.dqB, hqB, (size_t)ne * ne, NULL);
	qh_sync(NULL);
	free(hA); free(hB); free(hqA); free(hqB);
	e->next = qh_registry;
	qh_registry = e;
	pthread_mutex_unlock(&qh_registry_lock);
	return &e->t;
}

void qh_tables_release(const qpms_trans_calculator *c) {
	pthread_mutex_lock(&qh_registry_lock);
	qh_reg_entry **pp = &qh_registry;
	while (*pp) {
		if ((*pp)->c == c) {
			qh_reg_entry *e = *pp;
			*pp = e->next;
			qh_dfree(e->t.dA); qh_dfree(e->t.dB); qh_dfree(e->t.dqA); qh_dfree(e->t.dqB);
			free(e);
			break;
		}
		pp = &(*pp)->next;
	}
	pthread_mutex_unlock(&qh_registry_lock);
}

/* ---------------------------------------------------------------------- */
/* AB arrays (translations.c:610-698)                                       */
/* ---------------------------------------------------------------------- */
int qpms_trans_calculator_get_AB_arrays_batch_dev(const qpms_trans_calculator *c,
		size_t npts, const complex double *kr_dev, const double *theta_dev, const double *phi_dev,
		int r_ge_d, qpms_bessel_t J, complex double *A_dev, complex double *B_dev, void *stream) {
	if (r_ge_d) J = QPMS_BESSEL_REGULAR;
	QH_ENSURE(J >= 1 && J <= 4, "invalid Bessel kind %d", (int)J);
	const qh_trans_tables *t = qh_tables_get(c);
	return qh_launch_ab_batch(t, npts, (const double *)kr_dev, theta_dev, phi_dev, (int)J,
		(double *)A_dev, (double *)B_dev, stream);
}

int qpms_trans_calculator_get_AB_arrays(const qpms_trans_calculator *c,
		complex double *Adest, complex double *Bdest,
		size_t deststride, size_t srcstride,
		csph_t kdlj, bool r_ge_d, qpms_bessel_t J) {
	const size_t ne = c->nelem;
	const size_t bytes = ne * ne * sizeof(complex double);
	/* one staging allocation: kr (16 B), theta, phi (8 B each), A, B */
	char *d = qh_dmalloc(32 + 2 * bytes);
	double hin[4] = {creal(kdlj.r), cimag(kdlj.r), kdlj.theta, kdlj.phi};
	qh_h2d(d, hin, sizeof hin, NULL);
	complex double *dA = (complex double *)(d + 32), *dB = (complex double *)(d + 32 + bytes);
	qpms_trans_calculator_get_AB_arrays_batch_dev(c, 1, (const complex double *)d,
		(const double *)(d + 16), (const double *)(d + 24), r_ge_d, J, dA, dB, NULL);
	complex double *hA, *hB;
	QH_MALLOC(hA, 2 * bytes);
	hB = hA + ne * ne;
	qh_d2h(hA, dA, 2 * bytes, NULL);
	qh_sync(NULL);
	qh_dfree(d);
	for (size_t y = 0; y < ne; ++y)
		for (size_t yu = 0; yu < ne; ++yu) {
			Adest[deststride * y + srcstride * yu] = hA[y * ne + yu];
			Bdest[deststride * y + srcstride * yu] = hB[y * ne + yu];
		}
	free(hA);
	return 0;
}

int qpms_trans_calculator_get_AB_arrays_ext(const qpms_trans_calculator *c,
		complex double *Adest, complex double *Bdest,
		size_t deststride, size_t srcstride,
		complex double kdlj_r, double kdlj_theta, double kdlj_phi,
		int r_ge_d, int J) {
	csph_t kdlj = {kdlj_r, kdlj_theta, kdlj_phi};
	return qpms_trans_calculator_get_AB_arrays(c, Adest, Bdest, deststride, srcstride,
		kdlj, r_ge_d, (qpms_bessel_t)J);
}

/* uvswfi → (type, y) (indexing.h:87-94) */
static int qh_uvswfi2ty(qpms_uvswfi_t u, int *t, long *y) {
	*t = (int)(u & 3);
	*y = (long)(u / 4) - 1;
	if (*t == 0 || *t == 3 || *y < 0) return QPMS_ERROR;
	return QPMS_SUCCESS;
}

/* Scatter A/B (nelem × nelem of lMax_AB) into a bspec-ordered block
 * (translations_inlines.h:8-35). */
void qh_trans_array_from_AB(complex double *t, const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		const complex double *A, const complex double *B, int nelem_AB) {
	for (size_t di = 0; di < destspec->n; ++di) {
		int dt; long dy;
		QH_ENSURE(qh_uvswfi2ty(destspec->ilist[di], &dt, &dy) == QPMS_SUCCESS,
			"Invalid u. vswf index %llx.", destspec->ilist[di]);
		for (size_t si = 0; si < srcspec->n; ++si) {
			int st; long sy;
			QH_ENSURE(qh_uvswfi2ty(srcspec->ilist[si], &st, &sy) == QPMS_SUCCESS,
				"Invalid u. vswf index %llx.", srcspec->ilist[si]);
			t[si * srcstride + di * deststride] = (st == dt) ? A[dy * nelem_AB + sy] : B[dy * nelem_AB + sy];
		}
	}
}

static void qh_check_symmetric_norm(qpms_normalisation_t norm) {
	int nb = norm & QPMS_NORMALISATION_NORM_BITS;
	if (nb != QPMS_NORMALISATION_NORM_SPHARM && nb != QPMS_NORMALISATION_NORM_POWER)
		QH_DIE("Only spherical harmonic and power normalisation supported.");
	if ((!(norm & QPMS_NORMALISATION_N_I) != !(norm & QPMS_NORMALISATION_M_I))
			|| (!(norm & QPMS_NORMALISATION_N_MINUS) != !(norm & QPMS_NORMALISATION_M_MINUS)))
		QH_DIE("Only normalisations without a phase factors between M and N waves are supported.");
}

qpms_errno_t qpms_trans_calculator_get_trans_array(const qpms_trans_calculator *c,
		complex double *target,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		csph_t kdlj, bool r_ge_d, qpms_bessel_t J) {
	qh_check_symmetric_norm(c->normalisation);
	QH_ENSURE(c->normalisation == destspec->norm && c->normalisation == srcspec->norm,
		"The normalisation conventions must be the same");
	QH_ENSURE(c->lMax >= destspec->lMax && c->lMax >= srcspec->lMax, "bspec lMax exceeds calculator lMax");
	const size_t ne = c->nelem;
	complex double *A;
	QH_MALLOC(A, 2 * ne * ne * sizeof(complex double));
	complex double *B = A + ne * ne;
	int r = qpms_trans_calculator_get_AB_arrays(c, A, B, ne, 1, kdlj, r_ge_d, J);
	qh_trans_array_from_AB(target, destspec, deststride, srcspec, srcstride, A, B, (int)ne);
	free(A);
	return r;
}

qpms_errno_t qpms_trans_calculator_get_trans_array_lc3p(const qpms_trans_calculator *c,
		complex double *target,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		complex double k, cart3_t destpos, cart3_t srcpos, qpms_bessel_t J) {
	/* vectors.h:359-365 */
	double dx = destpos.x - srcpos.x, dy = destpos.y - srcpos.y, dz = destpos.z - srcpos.z;
	double r = sqrt(dx * dx + dy * dy + dz * dz);
	csph_t kd;
	kd.r = r * k;
	kd.theta = r ? acos(dz / r) : M_PI_2;
	kd.phi = atan2(dy, dx);
	return qpms_trans_calculator_get_trans_array(c, target, destspec, deststride,
		srcspec, srcstride, kd, false, J);
}
