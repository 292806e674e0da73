/* Oracle (test infrastructure): special functions, translation coefficients,
 * W / mode-problem builders, Mie coefficients, dense LU. See oracle.h.
 */
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;
typedef long double complex ldc;

static int g_threads = 1;
void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }

static inline int min1pow(int k) { return (k % 2) ? -1 : 1; }
static inline double complex ipow(int k) {
	switch (((k % 4) + 4) % 4) { case 0: return 1; case 1: return I; case 2: return -1; default: return -I; }
}
static inline int imin(int a, int b) { return a < b ? a : b; }

/* ---------------------------------------------------------------------- */
/* Legendre (translations.c:652-653 → GSL NONE)                             */
/* ---------------------------------------------------------------------- */
static ld ferrers_ld(int l, int m, ld x, int csphase) {
	/* m may be negative: P_l^{-k} = (-1)^k (l-k)!/(l+k)! P_l^k (Condon–Shortley). */
	if (m < 0) {
		int k = -m;
		ld f = 1;
		for (int i = l - k + 1; i <= l + k; ++i) f *= i;
		ld v = ferrers_ld(l, k, x, csphase) / f;
		return (csphase == -1 && (k % 2)) ? -v : v;
	}
	if (m > l) return 0;
	ld s = sqrtl((1 - x) * (1 + x));
	ld pmm = 1;
	for (int i = 1; i <= m; ++i) pmm *= (2 * i - 1) * s;
	if (csphase == -1 && (m % 2)) pmm = -pmm;
	if (l == m) return pmm;
	ld p1 = x * (2 * m + 1) * pmm, p0 = pmm;
	for (int ll = m + 2; ll <= l; ++ll) {
		ld pl = ((2 * ll - 1) * x * p1 - (ll + m - 1) * p0) / (ll - m);
		p0 = p1; p1 = pl;
	}
	return p1;
}

void or_legendre_array(int lmax, double x, int csphase, double *out) {
	for (int l = 0; l <= lmax; ++l)
		for (int m = 0; m <= l; ++m) out[l * (l + 1) / 2 + m] = (double)ferrers_ld(l, m, x, csphase);
}

/* ---------------------------------------------------------------------- */
/* Spherical Bessel: j by Miller's backward recurrence normalised with        */
/* sum_l (2l+1) j_l^2 = 1 (DLMF 10.60.5 at angle 0), y by upward recurrence.  */
/* ---------------------------------------------------------------------- */
void or_sph_bessel(int J, int lmax, double complex xd, double complex *out) {
	ldc x = xd;
	ld ax = cabsl(x);
	if (ax == 0) {
		for (int l = 0; l <= lmax; ++l) out[l] = (J == 1) ? (l == 0 ? 1 : 0) : NAN + I * NAN;
		return;
	}
	/* Hankel finite sums also for real arguments far above the order, where
	 * the downward recurrence below would need O(|x|) steps; j = Re h, y = Im h. */
	if ((J >= 3 && cimagl(x) != 0) || (cimagl(x) == 0 && ax > 1000 + 10 * lmax)) {
		/* spherical Hankel functions directly, DLMF 10.49.6:
		 * h^(1,2)_n(z) = (-+i)^{n+1} e^{+-iz}/z sum_k (+-i/(2z))^k (n+k)!/(k!(n-k)!) */
		const ld sg = (J == 4) ? -1 : 1;
		const ldc e = cexpl(sg * I * x) / x;
		for (int n = 0; n <= lmax; ++n) {
			ldc sum = 0, term = 1;
			for (int k = 0; k <= n; ++k) {
				if (k > 0) term *= sg * I / (2 * x) * (ld)(n + k) * (n - k + 1) / k;
				sum += term;
			}
			ldc ph = 1;
			for (int q = 0; q <= n; ++q) ph *= -sg * I;
			const ldc h = ph * e * sum;
			if (cimagl(x) != 0) out[n] = (double complex)h;
			else {
				const double jr = (double)creall(h), yr = (double)(sg * cimagl(h));
				out[n] = (J == 1) ? jr : (J == 2) ? yr : (J == 3) ? jr + I * yr : jr - I * yr;
			}
		}
		return;
	}
	ldc *j = malloc(sizeof(ldc) * (lmax + 2));
	ldc *y = malloc(sizeof(ldc) * (lmax + 2));
	ldc s = csinl(x), c = ccosl(x);
	y[0] = -c / x;
	if (lmax >= 1) y[1] = -c / (x * x) - s / x;
	for (int l = 1; l < lmax; ++l) y[l + 1] = (2 * l + 1) / x * y[l] - y[l - 1];
	int start = lmax + 40 + (int)(2 * ax);
	ldc *f = malloc(sizeof(ldc) * (start + 2));
	f[start + 1] = 0;
	f[start] = 1e-30L;
	for (int l = start; l >= 1; --l) {
		f[l - 1] = (2 * l + 1) / x * f[l] - f[l + 1];
		if (cabsl(f[l - 1]) > 1e1000L) {  /* rescale */
			for (int k = l - 1; k <= start + 1; ++k) f[k] *= 1e-1000L;
		}
	}
	ldc S = 0;
	for (int l = start; l >= 0; --l) S += (2 * l + 1) * f[l] * f[l];
	ldc sc = 1 / csqrtl(S);
	ldc j0 = s / x, j1 = s / (x * x) - c / x;
	/* fix the sign ambiguity of the square root with the larger of j0, j1 */
	if (cabsl(j0) >= cabsl(j1)) { if (creall(conjl(j0) * f[0] * sc) < 0) sc = -sc; }
	else { if (creall(conjl(j1) * f[1] * sc) < 0) sc = -sc; }
	for (int l = 0; l <= lmax; ++l) j[l] = f[l] * sc;
	for (int l = 0; l <= lmax; ++l) {
		switch (J) {
			case 1: out[l] = j[l]; break;
			case 2: out[l] = y[l]; break;
			case 3: out[l] = j[l] + I * y[l]; break;
			default: out[l] = j[l] - I * y[l]; break;
		}
	}
	/* real argument → real j, y exactly (bessel.c:52-83 uses real GSL arrays) */
	if (cimag(xd) == 0) {
		for (int l = 0; l <= lmax; ++l) {
			double jr = (double)creall(j[l]), yr = (double)creall(y[l]);
			switch (J) {
				case 1: out[l] = jr; break;
				case 2: out[l] = yr; break;
				case 3: out[l] = jr + I * yr; break;
				default: out[l] = jr - I * yr; break;
			}
		}
	}
	free(f); free(j); free(y);
}

/* ---------------------------------------------------------------------- */
/* Wigner 3j (Racah) and Gaunt (quadrature)                                 */
/* ---------------------------------------------------------------------- */
static ld factl(int n) { ld f = 1; for (int i = 2; i <= n; ++i) f *= i; return f; }

double or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
	if (m1 + m2 + m3 != 0) return 0;
	if (j3 < abs(j1 - j2) || j3 > j1 + j2) return 0;
	if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0;
	ld pre = sqrtl(factl(j1 + j2 - j3) * factl(j1 - j2 + j3) * factl(-j1 + j2 + j3) / factl(j1 + j2 + j3 + 1)
		* factl(j1 + m1) * factl(j1 - m1) * factl(j2 + m2) * factl(j2 - m2) * factl(j3 + m3) * factl(j3 - m3));
	ld sum = 0;
	for (int k = 0; k <= j1 + j2 + j3; ++k) {
		int a1 = j3 - j2 + k + m1, a2 = j3 - j1 + k - m2, a3 = j1 + j2 - j3 - k, a4 = j1 - k - m1, a5 = j2 - k + m2;
		if (a1 < 0 || a2 < 0 || a3 < 0 || a4 < 0 || a5 < 0) continue;
		sum += (k % 2 ? -1.0L : 1.0L) / (factl(k) * factl(a1) * factl(a2) * factl(a3) * factl(a4) * factl(a5));
	}
	return (double)(min1pow(j1 - j2 - m3) * pre * sum);
}

#define OR_GL_N 48
static ld gl_x[OR_GL_N], gl_w[OR_GL_N];
static pthread_once_t gl_once = PTHREAD_ONCE_INIT;
static void gl_init(void) {
	const int N = OR_GL_N;
	for (int i = 0; i < N; ++i) {
		ld x = cosl(M_PI * (i + 0.75L) / (N + 0.5L));
		for (int it = 0; it < 100; ++it) {
			ld p0 = 1, p1 = x;
			for (int k = 2; k <= N; ++k) { ld p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
			ld dp = N * (x * p1 - p0) / (x * x - 1);
			ld dx = p1 / dp;
			x -= dx;
			if (fabsl(dx) < 1e-19L) {
				p0 = 1; p1 = x;
				for (int k = 2; k <= N; ++k) { ld p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
				dp = N * (x * p1 - p0) / (x * x - 1);
				gl_x[i] = x;
				gl_w[i] = 2 / ((1 - x * x) * dp * dp);
				break;
			}
		}
	}
}

/* P_n^m P_nu^mu = sum_q a_q P_p^{m+mu}, p = n + nu - 2q (Xu 1996, eq. 2). */
double or_gaunt(int m, int n, int mu, int nu, int q) {
	pthread_once(&gl_once, gl_init);
	const int p = n + nu - 2 * q, M = m + mu;
	if (p < abs(M)) return 0;
	ld integral = 0;
	for (int i = 0; i < OR_GL_N; ++i) {
		ld x = gl_x[i];
		integral += gl_w[i] * ferrers_ld(n, m, x, -1) * ferrers_ld(nu, mu, x, -1) * ferrers_ld(p, M, x, -1);
	}
	/* orthogonality: int P_p^M P_p^M = 2/(2p+1) (p+M)!/(p-M)! */
	ld norm = (2 * p + 1) / 2.0L * factl(p - M) / factl(p + M);
	return (double)(integral * norm);
}

/* ---------------------------------------------------------------------- */
/* Multipliers (translations.c:117-133, 313-437)                            */
/* ---------------------------------------------------------------------- */
#define NORM_POWER 65536
#define NORM_SPHARM (65536 * 3)
#define NORM_NONE (65536 * 2)
#define NORM_BITS (65536 * 3)
#define CSPHASE 8

static double complex factor_noCS(int norm, int l, int m, int is_M) {
	double complex f;
	switch (norm & NORM_BITS) {
		case NORM_POWER: f = 1; break;
		case NORM_SPHARM: f = sqrt(l * (l + 1.0)); break;
		default: f = sqrt(l * (l + 1.0) * 4 * M_PI / (2 * l + 1)) * exp(0.5 * (lgamma(l + m + 1) - lgamma(l - m + 1))); break;
	}
	if (is_M) { if (norm & 32) f *= -1; if (norm & 16) f *= I; }
	else { if (norm & 128) f *= -1; if (norm & 64) f *= I; }
	if (norm & 1) f = 1 / f;
	return f;
}
static double normlogfac(int m, int n, int mu, int nu) {
	return -0.5 * (lgamma(n + m + 1) - lgamma(n - m + 1) + lgamma(nu - mu + 1) - lgamma(nu + mu + 1));
}
static double normfac_tr(int norm, int m, int n, int mu, int nu) {
	int cs = (norm & CSPHASE) ? -1 : 1;
	double f = (cs == 1) ? min1pow(m - mu) : 1.;
	f *= sqrt((n * (n + 1.)) / (nu * (nu + 1.)));
	f *= sqrt((2. * n + 1) / (2. * nu + 1));
	return f;
}
static int q_max(int m, int n, int mu, int nu) { return imin(n, imin(nu, (n + nu - abs(m + mu)) / 2)); }
static int qB_max(int M, int n, int mu, int nu) { return imin(n, imin(nu, (n + nu + 1 - abs(M + mu)) / 2)); }

static void mult_A(int norm, double complex *dest, int m, int n, int mu, int nu, int qmax) {
	double a0 = or_gaunt(-m, n, mu, nu, 0);
	double complex nf = normfac_tr(norm, m, n, mu, nu) * factor_noCS(norm, nu, mu, 0) / factor_noCS(norm, n, m, 0);
	nf *= min1pow(m);
	double ex = lgamma(2 * n + 1) - lgamma(n + 2) + lgamma(2 * nu + 3) - lgamma(nu + 2)
		+ lgamma(n + nu + m - mu + 1) - lgamma(n - m + 1) - lgamma(nu + mu + 1)
		+ lgamma(n + nu + 1) - lgamma(2 * (n + nu) + 1) + normlogfac(m, n, mu, nu);
	double complex presum = exp(ex) * nf / (4. * n) * ipow(n + nu);
	for (int q = 0; q <= qmax; ++q) {
		int p = n + nu - 2 * q, d = mu - m;
		double ppfac = (d >= 0) ? 1 : min1pow(d) * exp(lgamma(1 + p + d) - lgamma(1 + p - d));
		double sf = (n * (n + 1) + nu * (nu + 1) - p * (p + 1)) * min1pow(q) * or_gaunt(-m, n, mu, nu, q) / a0;
		dest[q] = presum * sf * ppfac * ipow(n - nu);
	}
}

static double cruzan_bfactor(int M, int n, int mu, int nu, int p) {
	double lp = 0.5 * (lgamma(n + M + 1) - lgamma(n - M + 1) + lgamma(nu + mu + 1) - lgamma(nu - mu + 1)
		+ lgamma(p - M - mu + 2) - lgamma(p + M + mu + 2));
	return min1pow(mu + M) * (2 * p + 3) * exp(lp) * or_wigner3j(n, nu, p + 1, M, mu, -M - mu)
		* or_wigner3j(n, nu, p, 0, 0, 0);
}

static void mult_B(int norm, double complex *dest, int m, int n, int mu, int nu, int Qmax) {
	double complex nf = normfac_tr(norm, m, n, mu, nu) * factor_noCS(norm, nu, mu, 1) / factor_noCS(norm, n, m, 0);
	double complex presum = min1pow(1 - m) * (2 * nu + 1) / (2. * (n * (n + 1)))
		* exp(lgamma(n + m + 1) - lgamma(n - m + 1) + lgamma(nu - mu + 1) - lgamma(nu + mu + 1) + normlogfac(m, n, mu, nu)) * nf;
	for (int q = 0; q <= Qmax; ++q) {
		int p = n + nu - 2 * q, d = mu - m;
		double ppfac = (d >= 0) ? 1 : min1pow(d) * exp(lgamma(2 + p + d) - lgamma(2 + p - d));
		double t = sqrt((double)((p + 1) * (p + 1) - (n - nu) * (n - nu)) * (double)((n + nu + 1) * (n + nu + 1) - (p + 1) * (p + 1)));
		dest[q] = presum * t * ppfac * cruzan_bfactor(-m, n, mu, nu, p) * ipow(p + 1) * ipow(n - nu);
	}
}

static void y2mn(int y, int *m, int *n) { *n = (int)sqrt((double)y + 1); *m = y - *n * (*n + 1) + 1; }

or_calc *or_calc_init(int lMax, int norm) {
	or_calc *c = calloc(1, sizeof(*c));
	c->lMax = lMax; c->nelem = lMax * (lMax + 2); c->norm = norm;
	const int ne = c->nelem, qs = lMax + 1;
	c->A = calloc((size_t)ne * ne * qs, sizeof(double complex));
	c->B = calloc((size_t)ne * ne * qs, sizeof(double complex));
	c->qA = calloc((size_t)ne * ne, sizeof(int));
	c->qB = calloc((size_t)ne * ne, sizeof(int));
	for (int y = 0; y < ne; ++y)
		for (int yu = 0; yu < ne; ++yu) {
			int m, n, mu, nu;
			y2mn(y, &m, &n); y2mn(yu, &mu, &nu);
			const int i = y * ne + yu;
			c->qA[i] = q_max(-m, n, mu, nu) + 1;
			c->qB[i] = qB_max(-m, n, mu, nu) + 1;
			mult_A(norm, c->A + (size_t)i * qs, m, n, mu, nu, c->qA[i] - 1);
			mult_B(norm, c->B + (size_t)i * qs, m, n, mu, nu, c->qB[i] - 1);
		}
	return c;
}

void or_calc_free(or_calc *c) {
	if (!c) return;
	free(c->A); free(c->B); free(c->qA); free(c->qB); free(c);
}

int or_calc_multipliers(const or_calc *c, int y, int yu, double complex *A, int *nA, double complex *B, int *nB) {
	const int i = y * c->nelem + yu, qs = c->lMax + 1;
	*nA = c->qA[i]; *nB = c->qB[i];
	memcpy(A, c->A + (size_t)i * qs, *nA * sizeof(double complex));
	memcpy(B, c->B + (size_t)i * qs, *nB * sizeof(double complex));
	return 0;
}

/* ---------------------------------------------------------------------- */
/* AB arrays with Kahan sums (translations.c:510-528, 549-567, 610-698)     */
/* ---------------------------------------------------------------------- */
static inline void ckahan(double complex *sum, double complex *comp, double complex in) {
	double complex y = in - *comp, t = *sum + y;
	*comp = (t - *sum) - y;
	*sum = t;
}

void or_get_AB_arrays(const or_calc *c, double complex kr, double theta, double phi, int r_ge_d, int J,
		double complex *A, double complex *B) {
	if (r_ge_d) J = 1;
	const int ne = c->nelem, qs = c->lMax + 1, L = 2 * c->lMax + 1;
	if (kr == 0 && J != 1) {
		for (int i = 0; i < ne * ne; ++i) { A[i] = NAN + I * NAN; B[i] = NAN + I * NAN; }
		return;
	}
	double leg[(L + 1) * (L + 2) / 2];
	double complex bes[L + 1];
	or_legendre_array(L, cos(theta), -1, leg);
	or_sph_bessel(J, L, kr, bes);
	for (int y = 0; y < ne; ++y)
		for (int yu = 0; yu < ne; ++yu) {
			int m, n, mu, nu;
			y2mn(y, &m, &n); y2mn(yu, &mu, &nu);
			const int i = y * ne + yu, ad = abs(mu - m);
			double complex sa = 0, ca = 0, sb = 0, cb = 0;
			for (int q = 0; q < c->qA[i]; ++q) {
				int p = n + nu - 2 * q;
				ckahan(&sa, &ca, leg[p * (p + 1) / 2 + ad] * bes[p] * c->A[(size_t)i * qs + q]);
			}
			for (int q = 0; q < c->qB[i]; ++q) {
				int p = n + nu - 2 * q + 1;
				ckahan(&sb, &cb, leg[p * (p + 1) / 2 + ad] * bes[p] * c->B[(size_t)i * qs + q]);
			}
			double complex e = cexp(I * (mu - m) * phi);
			A[i] = sa * e;
			B[i] = sb * e;
		}
}

/* translations.c:135-286 */
double complex or_trans_single_A(int norm, int m, int n, int mu, int nu, double complex kr, double theta,
		double phi, int J) {
	const int qmax = q_max(-m, n, mu, nu);
	double leg[(n + nu + 1) * (n + nu + 2) / 2];
	double complex bes[n + nu + 1];
	or_legendre_array(n + nu, cos(theta), -1, leg);
	or_sph_bessel(J, n + nu, kr, bes);
	double a0 = or_gaunt(-m, n, mu, nu, 0);
	double complex sum = 0;
	for (int q = 0; q <= qmax; ++q) {
		int p = n + nu - 2 * q, d = mu - m;
		double Pp = leg[p * (p + 1) / 2 + abs(d)];
		if (d < 0) Pp *= min1pow(mu - m) * exp(lgamma(1 + p + d) - lgamma(1 + p - d));
		sum += (n * (n + 1) + nu * (nu + 1) - p * (p + 1)) * min1pow(q) * (or_gaunt(-m, n, mu, nu, q) / a0) * bes[p] * Pp;
	}
	double ex = lgamma(2 * n + 1) - lgamma(n + 2) + lgamma(2 * nu + 3) - lgamma(nu + 2)
		+ lgamma(n + nu + m - mu + 1) - lgamma(n - m + 1) - lgamma(nu + mu + 1) + lgamma(n + nu + 1) - lgamma(2 * (n + nu) + 1);
	double complex presum = exp(ex) * cexp(I * (mu - m) * phi) * min1pow(m) * ipow(nu + n) / (4 * n);
	double complex nf = normfac_tr(norm, m, n, mu, nu) * factor_noCS(norm, nu, mu, 0) / factor_noCS(norm, n, m, 0);
	presum *= nf * exp(normlogfac(m, n, mu, nu)) * ipow(n - nu);
	return presum * sum;
}

double complex or_trans_single_B(int norm, int m, int n, int mu, int nu, double complex kr, double theta,
		double phi, int J) {
	const int q2max = q_max(-m - 1, n + 1, mu + 1, nu), Qmax = q_max(-m, n + 1, mu, nu);
	const int realQmax = qB_max(-m, n, mu, nu);
	double leg[(n + nu + 2) * (n + nu + 3) / 2];
	double complex bes[n + nu + 2];
	or_legendre_array(n + nu + 1, cos(theta), -1, leg);
	or_sph_bessel(J, n + nu + 1, kr, bes);
	double a2q0 = (mu == nu) ? 1 : or_gaunt(-m - 1, n + 1, mu + 1, nu, 0);
	double a3q0 = or_gaunt(-m, n + 1, mu, nu, 0);
	(void)q2max; (void)Qmax;
	double complex sum = 0;
	for (int q = 0; q <= realQmax; ++q) {
		int p = n + nu - 2 * q, d = mu - m;
		double a2 = (mu == nu) ? 0 : or_gaunt(-m - 1, n + 1, mu + 1, nu, q) / a2q0;
		double a3 = or_gaunt(-m, n + 1, mu, nu, q) / a3q0;
		double Pp = leg[(p + 1) * (p + 2) / 2 + abs(d)];
		if (d < 0) Pp *= min1pow(mu - m) * exp(lgamma(2 + p + d) - lgamma(2 + p - d));
		sum += ((2 * (n + 1) * (nu - mu) * a2 - (-nu * (nu + 1) - n * (n + 3) - 2 * mu * (n + 1) + p * (p + 3)) * a3)
			* min1pow(q) * bes[p + 1] * Pp);
	}
	double ex = lgamma(2 * n + 3) - lgamma(n + 2) + lgamma(2 * nu + 3) - lgamma(nu + 2)
		+ lgamma(n + nu + m - mu + 2) - lgamma(n - m + 1) - lgamma(nu + mu + 1) + lgamma(n + nu + 2) - lgamma(2 * (n + nu) + 3);
	double complex presum = exp(ex) * cexp(I * (mu - m) * phi) * min1pow(m) * ipow(nu + n + 1) / ((4 * n) * (n + 1) * (n + m + 1));
	double complex nf = normfac_tr(norm, m, n, mu, nu) * factor_noCS(norm, nu, mu, 1) / factor_noCS(norm, n, m, 0);
	presum *= exp(normlogfac(m, n, mu, nu)) * nf * ipow(n - nu);
	return presum * sum;
}

/* ---------------------------------------------------------------------- */
/* W and M builders (scatsystem.c:1180-1217, 1301-1339)                     */
/* ---------------------------------------------------------------------- */
typedef struct {
	const or_calc *c; int npart; const double *pos; double complex k; int J;
	int row0, nrows; double complex *W; const int *ttype; const double complex *tdiag; int modeM;
	int next; pthread_mutex_t *lock;
} wjob;

static void build_row(wjob *jb, int r) {
	const or_calc *c = jb->c;
	const int ne = c->nelem, nb = 2 * ne, np = jb->npart;
	const size_t ld = (size_t)nb * np;
	double complex *A = malloc(sizeof(double complex) * ne * ne * 2), *B = A + ne * ne;
	double complex *rows = jb->W + (size_t)(r - jb->row0) * nb * ld;
	for (int cc = 0; cc < np; ++cc) {
		if (cc == r) {
			for (int i = 0; i < nb; ++i)
				for (int j = 0; j < nb; ++j) rows[(size_t)i * ld + (size_t)cc * nb + j] = (jb->modeM && i == j) ? 1 : 0;
			continue;
		}
		double dx = jb->pos[3 * r] - jb->pos[3 * cc], dy = jb->pos[3 * r + 1] - jb->pos[3 * cc + 1], dz = jb->pos[3 * r + 2] - jb->pos[3 * cc + 2];
		double d = sqrt(dx * dx + dy * dy + dz * dz);
		or_get_AB_arrays(c, d * jb->k, acos(dz / d), atan2(dy, dx), 0, jb->J, A, B);
		for (int ti = 0; ti < 2; ++ti)
			for (int y = 0; y < ne; ++y) {
				const int i = ti * ne + y;
				double complex s = 1;
				if (jb->modeM) s = -jb->tdiag[(size_t)jb->ttype[r] * nb + i];
				for (int tj = 0; tj < 2; ++tj)
					for (int yu = 0; yu < ne; ++yu)
						rows[(size_t)i * ld + (size_t)cc * nb + tj * ne + yu] = s * ((ti == tj) ? A[y * ne + yu] : B[y * ne + yu]);
			}
	}
	free(A);
}

static void *wjob_worker(void *arg) {
	wjob *jb = arg;
	for (;;) {
		pthread_mutex_lock(jb->lock);
		int r = jb->next++;
		pthread_mutex_unlock(jb->lock);
		if (r >= jb->row0 + jb->nrows) break;
		build_row(jb, r);
	}
	return NULL;
}

static void run_wjob(wjob *jb) {
	pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
	jb->lock = &lock;
	jb->next = jb->row0;
	int nt = g_threads;
	pthread_t th[256];
	if (nt > 256) nt = 256;
	for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, wjob_worker, jb);
	for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
}

void or_build_W_rows(const or_calc *c, int npart, const double *pos, double complex k, int J,
		int row0, int nrows, double complex *W) {
	wjob jb = {c, npart, pos, k, J, row0, nrows, W, NULL, NULL, 0, 0, NULL};
	run_wjob(&jb);
}

void or_build_M(const or_calc *c, int npart, const double *pos, const int *ttype,
		const double complex *tdiag, double complex k, double complex *M) {
	wjob jb = {c, npart, pos, k, 3, 0, npart, M, ttype, tdiag, 1, 0, NULL};
	run_wjob(&jb);
}

/* ---------------------------------------------------------------------- */
/* Mie (tmatrices.c:500-561)                                                */
/* ---------------------------------------------------------------------- */
void or_mie(int lMax, double a, double complex ki, double complex ke, double complex mui,
		double complex mue, double complex *RV, double complex *RH) {
	double complex xi = ki * a, xe = ke * a, eti = ki / mui, ete = ke / mue;
	double complex zi[lMax + 2], ze[lMax + 2], zs[lMax + 2];
	or_sph_bessel(1, lMax + 1, xi, zi);
	or_sph_bessel(1, lMax + 1, xe, ze);
	or_sph_bessel(3, lMax + 1, xe, zs);
	for (int l = 0; l <= lMax; ++l) {
		double complex dzi = -zi[l + 1] + l / xi * zi[l];
		double complex dze = -ze[l + 1] + l / xe * ze[l];
		double complex dzs = -zs[l + 1] + l / xe * zs[l];
		double complex fi = zi[l] / xi + dzi, fs = zs[l] / xe + dzs, fe = ze[l] / xe + dze;
		RV[l] = (-eti * fe * zi[l] + ete * ze[l] * fi) / (-ete * fi * zs[l] + eti * zi[l] * fs);
		RH[l] = (-ete * fe * zi[l] + eti * ze[l] * fi) / (-eti * fi * zs[l] + ete * zi[l] * fs);
	}
}

/* ---------------------------------------------------------------------- */
/* LU (zgetrf / zgetrs semantics, row-major)                                */
/* ---------------------------------------------------------------------- */
typedef struct { int n, k, r0, r1; double complex *A; } lujob;
static void *lu_update(void *arg) {
	lujob *j = arg;
	const int n = j->n, k = j->k;
	double complex *A = j->A;
	const double complex *pr = A + (size_t)k * n;
	for (int i = j->r0; i < j->r1; ++i) {
		double complex *row = A + (size_t)i * n;
		const double complex l = row[k] / pr[k];
		row[k] = l;
		for (int c = k + 1; c < n; ++c) row[c] -= l * pr[c];
	}
	return NULL;
}

int or_zgetrf(int n, double complex *A, int *ipiv, int nthreads) {
	int info = 0;
	if (nthreads < 1) nthreads = 1;
	for (int k = 0; k < n; ++k) {
		int p = k;
		double best = -1;
		for (int i = k; i < n; ++i) {
			double v = fabs(creal(A[(size_t)i * n + k])) + fabs(cimag(A[(size_t)i * n + k]));
			if (v > best) { best = v; p = i; }
		}
		ipiv[k] = p + 1;
		if (best == 0) { if (!info) info = k + 1; continue; }
		if (p != k)
			for (int c = 0; c < n; ++c) {
				double complex t = A[(size_t)k * n + c];
				A[(size_t)k * n + c] = A[(size_t)p * n + c];
				A[(size_t)p * n + c] = t;
			}
		const int rows = n - k - 1;
		if (rows <= 0) continue;
		int nt = (rows > 64 * nthreads) ? nthreads : 1;
		if (nt == 1) { lujob j = {n, k, k + 1, n, A}; lu_update(&j); continue; }
		pthread_t th[256];
		lujob jb[256];
		if (nt > 256) nt = 256;
		for (int t = 0; t < nt; ++t) {
			jb[t].n = n; jb[t].k = k; jb[t].A = A;
			jb[t].r0 = k + 1 + (int)((long)rows * t / nt);
			jb[t].r1 = k + 1 + (int)((long)rows * (t + 1) / nt);
			pthread_create(&th[t], NULL, lu_update, &jb[t]);
		}
		for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
	}
	return info;
}

void or_zgetrs(int n, int nrhs, const double complex *LU, const int *ipiv, double complex *B) {
	for (int i = 0; i < n; ++i) {
		int p = ipiv[i] - 1;
		if (p != i)
			for (int c = 0; c < nrhs; ++c) {
				double complex t = B[(size_t)i * nrhs + c];
				B[(size_t)i * nrhs + c] = B[(size_t)p * nrhs + c];
				B[(size_t)p * nrhs + c] = t;
			}
	}
	for (int i = 0; i < n; ++i)
		for (int k = 0; k < i; ++k) {
			double complex l = LU[(size_t)i * n + k];
			for (int c = 0; c < nrhs; ++c) B[(size_t)i * nrhs + c] -= l * B[(size_t)k * nrhs + c];
		}
	for (int i = n - 1; i >= 0; --i) {
		for (int k = i + 1; k < n; ++k) {
			double complex u = LU[(size_t)i * n + k];
			for (int c = 0; c < nrhs; ++c) B[(size_t)i * nrhs + c] -= u * B[(size_t)k * nrhs + c];
		}
		for (int c = 0; c < nrhs; ++c) B[(size_t)i * nrhs + c] /= LU[(size_t)i * n + i];
	}
}
