/* Ewald-sum constant tables (ewald.c:104-257), host, init-time.
 *
 * s1_constfacs[y(m,n)][j] = -2 i^{n+1} sqrt(pi) ((n-m)/2)! ((n+m)/2)! (-1)^j
 *                           / (j! ((n-m)/2 - j)! ((n+m)/2 - j)!)      (n+m even)
 * S1_constfacs (out-of-plane shifts) and s1_constfacs_1Dz as documented in
 * ewald.h:63-130; legendre0/±1 are Ferrers P_l^m at 0 and ±1 with the given
 * Condon–Shortley sign.
 */
#include "qpms_internal.h"

void qh_e3_release(const qpms_ewald3_constants_t *c);
#include <complex.h>
#include <math.h>

static double qh_factorial(int n) {
	double f = 1;
	for (int i = 2; i <= n; ++i) f *= i;
	return f;
}
/* (n2/2)! = Γ(n2/2 + 1) */
static double qh_factorial_of_half(int n2) {
	if (n2 % 2 == 0) return qh_factorial(n2 / 2);
	const int k = n2 / 2 + 1;
	double f2 = 1;
	for (int j = 2 * k - 1; j > 0; j -= 2) f2 *= j;
	return f2 * pow(2, -k) * sqrt(M_PI);
}
static complex double qh_ipow(int k) {
	switch (((k % 4) + 4) % 4) { case 0: return 1; case 1: return I; case 2: return -1; default: return -I; }
}
static int qh_min1pow(int k) { return (k % 2) ? -1 : 1; }

static double qh_ferrers(int l, int m, double x, int csphase) {
	/* unnormalised Ferrers P_l^m(x), m >= 0, CS phase (-1)^m iff csphase == -1 */
	double s = sqrt((1 - x) * (1 + x));
	double pmm = 1;
	for (int i = 1; i <= m; ++i) pmm *= (2 * i - 1) * s;
	if (csphase == -1 && (m % 2)) pmm = -pmm;
	if (l == m) return pmm;
	double p1 = x * (2 * m + 1) * pmm, p0 = pmm;
	for (int ll = m + 2; ll <= l; ++ll) {
		double pl = ((2 * ll - 1) * x * p1 - (ll + m - 1) * p0) / (ll - m);
		p0 = p1; p1 = pl;
	}
	return p1;
}

qpms_ewald3_constants_t *qpms_ewald3_constants_init(qpms_l_t lMax, int csphase) {
	qpms_ewald3_constants_t *c;
	QH_MALLOC(c, sizeof(*c));
	c->lMax = lMax;
	c->nelem_sc = qpms_lMax2nelem_sc(lMax);
	const size_t ne = c->nelem_sc;
	QH_MALLOC(c->s1_jMaxes, ne * sizeof(qpms_l_t));
	QH_MALLOC(c->s1_constfacs, ne * sizeof(complex double *));
	size_t sz = 0;
	for (size_t y = 0; y < ne; ++y) {
		int n = (int)sqrt((double)y), m = (int)y - n * (n + 1);
		if ((m + n) % 2 == 0) sz += 1 + (c->s1_jMaxes[y] = (n - abs(m)) / 2);
		else c->s1_jMaxes[y] = -1;
	}
	QH_MALLOC(c->s1_constfacs_base, (sz ? sz : 1) * sizeof(complex double));
	size_t cum = 0;
	for (size_t y = 0; y < ne; ++y) {
		int n = (int)sqrt((double)y), m = (int)y - n * (n + 1);
		if ((m + n) % 2 == 0) {
			c->s1_constfacs[y] = c->s1_constfacs_base + cum;
			for (int j = 0; j <= c->s1_jMaxes[y]; ++j)
				c->s1_constfacs[y][j] = -2 * qh_ipow(n + 1) * sqrt(M_PI)
					* qh_factorial((n - m) / 2) * qh_factorial((n + m) / 2) * qh_min1pow(j)
					/ (qh_factorial(j) * qh_factorial((n - m) / 2 - j) * qh_factorial((n + m) / 2 - j));
			cum += 1 + c->s1_jMaxes[y];
		} else c->s1_constfacs[y] = NULL;
	}
	/* 2D-in-3D constants for out-of-plane shifts */
	QH_MALLOC(c->S1_constfacs, (1 + ne) * sizeof(complex double *));
	size_t Ssz = 0;
	for (size_t y = 0; y < ne; ++y) {
		int n = (int)sqrt((double)y), m = (int)y - n * (n + 1);
		const int LM = n - abs(m);
		for (int j = 0; j <= LM; ++j)
			for (int s = j + (LM - j) % 2; s <= 2 * j && s <= LM; s += 2) ++Ssz;
	}
	QH_MALLOC(c->S1_constfacs_base, (Ssz ? Ssz : 1) * sizeof(complex double));
	cum = 0;
	for (size_t y = 0; y < ne; ++y) {
		int n = (int)sqrt((double)y), m = (int)y - n * (n + 1);
		const complex double yf = -2 * qh_ipow(n + 1) * sqrt(M_PI) * qh_factorial((n - m) / 2) * qh_factorial((n + m) / 2);
		c->S1_constfacs[y] = c->S1_constfacs_base + cum;
		const int LM = n - abs(m);
		for (int j = 0; j <= LM; ++j)
			for (int s = j + (LM - j) % 2; s <= 2 * j && s <= LM; s += 2)
				c->S1_constfacs_base[cum++] = yf * qh_min1pow(j) / qh_factorial(2 * j - s) / qh_factorial(s - j)
					/ qh_factorial_of_half(n - m - s) / qh_factorial_of_half(n + m - s);
	}
	c->S1_constfacs[ne] = c->S1_constfacs_base + Ssz;
	/* 1D z-axis constants */
	{
		const int sz_n = lMax / 2 + 1;
		size_t zsz = (lMax % 2) ? (size_t)(sz_n * sz_n + sz_n) : (size_t)(sz_n * sz_n);
		QH_MALLOC(c->s1_constfacs_1Dz_base, zsz * sizeof(complex double));
		QH_MALLOC(c->s1_constfacs_1Dz, (lMax + 1) * sizeof(complex double *));
		size_t zc = 0;
		for (int n = 0; n <= lMax; ++n) {
			c->s1_constfacs_1Dz[n] = c->s1_constfacs_1Dz_base + zc;
			for (int j = 0; j <= n / 2; ++j)
				c->s1_constfacs_1Dz[n][j] = -qh_ipow(n + 1) * qh_min1pow(j) * qh_factorial(n)
					/ (qh_factorial(j) * pow(2, 2 * j) * qh_factorial(n - 2 * j));
			zc += 1 + n / 2;
		}
	}
	c->legendre_csphase = csphase;
	c->legendre_normconv = 3; /* GSL_SF_LEGENDRE_NONE */
	const size_t nleg = (size_t)(lMax + 1) * (lMax + 2) / 2;
	QH_MALLOC(c->legendre0, nleg * sizeof(double));
	QH_MALLOC(c->legendre_plus1, nleg * sizeof(double));
	QH_MALLOC(c->legendre_minus1, nleg * sizeof(double));
	for (int l = 0; l <= lMax; ++l)
		for (int m = 0; m <= l; ++m) {
			const size_t i = (size_t)l * (l + 1) / 2 + m;
			c->legendre0[i] = qh_ferrers(l, m, 0.0, csphase);
			c->legendre_plus1[i] = (m == 0) ? 1.0 : 0.0;
			c->legendre_minus1[i] = (m == 0) ? qh_min1pow(l) : 0.0;
		}
	return c;
}

void qpms_ewald3_constants_free(qpms_ewald3_constants_t *c) {
	qh_e3_release(c);
	if (!c) return;
	free(c->legendre0); free(c->legendre_plus1); free(c->legendre_minus1);
	free(c->s1_constfacs); free(c->s1_constfacs_base);
	free(c->S1_constfacs); free(c->S1_constfacs_base);
	free(c->s1_constfacs_1Dz_base); free(c->s1_constfacs_1Dz);
	free(c->s1_jMaxes);
	free(c);
}
