/* Oracle (test infrastructure): 2D Ewald sums, restated from the reference.
 *
 *  or_xyweb_points   latticegens.c:449-567 (ring order, strict cut rules)
 *  or_gamma_inc      ewaldsf.c:44-205 (real / continued fraction / series
 *                    regimes and the non-principal branch formula)
 *  or_sr_integral    ewald.c:651-740, by composite Gauss–Legendre quadrature
 *                    in long double instead of GSL QAGIU
 *  or_sigma0         ewald.c:260-273
 *  or_sigma_short    ewald.c:742-920 (no auto-cutoff: all given points)
 *  or_sigma_long     ewald.c:286-503 (xy lattice, in-plane shift)
 *  or_e32_combine    translations.c:1024-1067
 */
#include "oracle.h"
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;
typedef long double complex ldc;

static inline int min1pow(int k) { return (k % 2) ? -1 : 1; }

/* ---------------------------------------------------------------------- */
/* Lattice                                                                   */
/* ---------------------------------------------------------------------- */
static void reduce(double *b1x, double *b1y, double *b2x, double *b2y) {
	double B1 = *b1x * *b1x + *b1y * *b1y;
	double mu = (*b1x * *b2x + *b1y * *b2y) / B1;
	*b2x -= round(mu) * *b1x; *b2y -= round(mu) * *b1y;
	double B2 = *b2x * *b2x + *b2y * *b2y;
	while (B2 < B1) {
		double tx = *b1x, ty = *b1y;
		*b1x = *b2x; *b1y = *b2y; *b2x = tx; *b2y = ty;
		B1 = B2;
		mu = (*b1x * *b2x + *b1y * *b2y) / B1;
		*b2x -= round(mu) * *b1x; *b2y -= round(mu) * *b1y;
		B2 = *b2x * *b2x + *b2y * *b2y;
	}
}

int or_xyweb_points(double b1x, double b1y, double b2x, double b2y, double rtol, double minR, int inc_minR,
		double maxR, int inc_maxR, double *xy, int cap) {
	/* latticegens.c:449-468: the walk uses the basis reduced once; the
	 * right-angle test (lattices2d.c:746-756) and the ring in-radius
	 * (lattices2d.c:841-848, via l2d_shortestBase3) each reduce it again. For a
	 * hexagonal basis whose Gram coefficient rounds to exactly 1/2 the second
	 * reduction flips the obtuse basis back to the acute one, the `<=` test then
	 * reports a right angle and the walk runs the 4-edge (rhombic) rings. */
	reduce(&b1x, &b1y, &b2x, &b2y);
	double c1x = b1x, c1y = b1y, c2x = b2x, c2y = b2y;
	reduce(&c1x, &c1y, &c2x, &c2y);
	double hx = c2x - c1x, hy = c2y - c1y;
	double c1s = c1x * c1x + c1y * c1y, c2s = c2x * c2x + c2y * c2y, hs = hx * hx + hy * hy;
	int rect = (hs - c2s - c1s <= rtol * (c1s + c2s));
	/* in-radius of the shortest-base hexagon (reduce, then obtuse test with rtol 0) */
	double o1x = c1x, o1y = c1y, o2x, o2y, o3x, o3y;
	if (hs - c2s - c1s > 0) { o3x = c2x; o3y = c2y; o2x = c2x + c1x; o2y = c2y + c1y; }
	else { o2x = c2x; o2y = c2y; o3x = c2x - c1x; o3y = c2y - c1y; }
	double r1 = hypot(o1x, o1y), r2 = hypot(o2x, o2y), r3 = hypot(o3x, o3y), p = (r1 + r2 + r3) / 2;
	double h = 2 * sqrt(p * (p - r1) * (p - r2) * (p - r3)) / r3;
	long layer = (long)ceil(minR / h);
	if (!inc_minR && layer * h <= minR) ++layer;
	long last = (long)floor(maxR / h);
	if (!inc_maxR && last * h >= maxR) --last;
	int n = 0;
	for (long l = layer; l <= last; ++l) {
		if (l == 0) {
			if (n < cap) { xy[2 * n] = 0; xy[2 * n + 1] = 0; }
			++n;
			continue;
		}
		/* walk the ring as the generator does */
		long i = l, j = 0;
		int nphase = rect ? 4 : 6;
		for (int ph = 0; ph < nphase;) {
			if (n < cap) { xy[2 * n] = i * b1x + j * b2x; xy[2 * n + 1] = i * b1y + j * b2y; }
			++n;
			if (rect) {
				switch (ph) {
					case 0: --i; ++j; if (i <= 0) ++ph; break;
					case 1: --i; --j; if (j <= 0) ++ph; break;
					case 2: ++i; --j; if (i >= 0) ++ph; break;
					default: ++i; ++j; if (j >= 0) ++ph; break;
				}
			} else {
				switch (ph) {
					case 0: --i; ++j; if (i <= 0) ++ph; break;
					case 1: --i; if (i + j <= 0) ++ph; break;
					case 2: --j; if (j <= 0) ++ph; break;
					case 3: ++i; --j; if (i >= 0) ++ph; break;
					case 4: ++i; if (i + j >= 0) ++ph; break;
					default: ++j; if (j >= 0) ++ph; break;
				}
			}
		}
	}
	return n;
}

/* ---------------------------------------------------------------------- */
/* Incomplete gamma (ewaldsf.c:44-205)                                       */
/* ---------------------------------------------------------------------- */
static ld gamma_ld(ld a) { return tgammal(a); }

/* continued fraction F(a,z): Gamma(a,z) = z^{a-1} e^{-z} F */
static ldc cf_F(ld a, ldc z) {
	const ld small = 1e-60L;
	ldc hn = 1, Cn = 1 / small, Dn = 1;
	for (int n = 2; n < 20000; ++n) {
		ldc an = (n % 2) ? 0.5L * (n - 1) / z : (0.5L * n - a) / z;
		Dn = 1 + an * Dn;
		if (cabsl(Dn) < small) Dn = small;
		Cn = 1 + an / Cn;
		if (cabsl(Cn) < small) Cn = small;
		Dn = 1 / Dn;
		ldc delta = Cn * Dn;
		hn *= delta;
		if (cabsl(delta - 1) < 1e-19L) break;
	}
	return hn;
}

static ldc series_G(ld a, ldc z) {
	ld ga = gamma_ld(a);
	ldc sum = 0, zk = 1;
	ld gak = gamma_ld(a + 1);
	for (int k = 0; k < 2000; ++k) {
		ldc term = zk / gak;
		sum += term;
		if (k > 5 && cabsl(term) < 1e-21L * cabsl(sum)) break;
		zk *= z;
		gak *= (a + k + 1);
	}
	return ga * (1 - cpowl(z, a) * cexpl(-z) * sum);
}

double complex or_gamma_inc(double a, double complex xd, int m) {
	ldc x = xd;
	ldc val;
	if (creal(xd) >= 0 && (cimag(xd) == 0 || fabs(cimag(xd)) < fabs(creal(xd)) * 1e-14)) {
		ld xr = creall(x);
		if (xr > 1) val = cf_F(a, xr) * expl((a - 1) * logl(xr) - xr);
		else val = series_G(a, xr);
	} else if (creal(xd) >= 0 && cabs(xd) > 0.5) {
		val = cf_F(a, x) * cexpl((a - 1) * clogl(x) - x);
	} else {
		val = series_G(a, x);
	}
	if (m) {
		ldc f = cexpl(2 * m * M_PI * a * I);
		val = f * val + (1 - f) * gamma_ld(a);
	}
	return (double complex)val;
}

/* ---------------------------------------------------------------------- */
/* Short-range integral                                                      */
/* ---------------------------------------------------------------------- */
#define GLN 40
static ld glx[GLN], glw[GLN];
static int gl_ready = 0;
static void gl40(void) {
	if (gl_ready) return;
	for (int i = 0; i < GLN; ++i) {
		ld x = cosl(M_PI * (i + 0.75L) / (GLN + 0.5L));
		for (int it = 0; it < 100; ++it) {
			ld p0 = 1, p1 = x;
			for (int k = 2; k <= GLN; ++k) { ld p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
			ld dp = GLN * (x * p1 - p0) / (x * x - 1), dx = p1 / dp;
			x -= dx;
			if (fabsl(dx) < 1e-19L) {
				p0 = 1; p1 = x;
				for (int k = 2; k <= GLN; ++k) { ld p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
				dp = GLN * (x * p1 - p0) / (x * x - 1);
				glx[i] = x; glw[i] = 2 / ((1 - x * x) * dp * dp);
				break;
			}
		}
	}
	gl_ready = 1;
}

/* int_eta^inf exp(-R^2 xi^2 + k^2/(4 xi^2)) xi^{2n} dxi, complex k */
static ldc sr_integral_c(ld R, ldc k, int n, ld eta) {
	gl40();
	ldc sum = 0;
	const ld w = 0.5L / R;          /* panel width: half the Gaussian scale */
	const int npanel = 80;          /* up to eta + 40/R: exp(-1600) */
	for (int p = 0; p < npanel; ++p) {
		ld a = eta + p * w, b = a + w;
		ld c = 0.5L * (a + b), hw = 0.5L * (b - a);
		for (int i = 0; i < GLN; ++i) {
			ld xi = c + hw * glx[i];
			ldc f = cexpl(-R * R * xi * xi + k * k / (4 * xi * xi)) * powl(xi, 2 * n);
			sum += glw[i] * hw * f;
		}
	}
	return sum;
}

double or_sr_integral(double r, double k, int n, double eta) { return (double)creall(sr_integral_c(r, k, n, eta)); }

/* ---------------------------------------------------------------------- */
/* σ sums                                                                    */
/* ---------------------------------------------------------------------- */
static void ferrers0(int L, int csphase, double *leg) {
	or_legendre_array(L, 0.0, csphase, leg);
}

int or_sigma0(int lMax_e, double eta, double complex k, double complex *out) {
	(void)lMax_e;
	double complex z = -(k / (2 * eta)) * (k / (2 * eta));
	int m = (creal(z) < 0 && !signbit(cimag(z))) ? -1 : 0;
	double complex g = or_gamma_inc(-0.5, z, m);
	*out = g * 1.0 / 2 / sqrt(M_PI);   /* P_0^0(0) = 1 */
	return 0;
}

static inline void ck(double complex *s, double complex *c, double complex in) {
	double complex y = in - *c, t = *s + y;
	*c = (t - *s) - y;
	*s = t;
}

int or_sigma_short(int L, int csphase, double eta, double complex k, int npts, const double *Rxy,
		double betax, double betay, double sx, double sy, double complex *out) {
	const int nsc = (L + 1) * (L + 1);
	double complex comp[nsc];
	double leg[(L + 1) * (L + 2) / 2];
	ferrers0(L, csphase, leg);
	for (int i = 0; i < nsc; ++i) out[i] = comp[i] = 0;
	const int kreal = cimag(k) == 0;
	for (int ip = 0; ip < npts; ++ip) {
		double rx = Rxy[2 * ip] - sx, ry = Rxy[2 * ip + 1] - sy;
		double r = sqrt(rx * rx + ry * ry);
		double phi = atan2(ry, rx);
		double complex eb = cexp(I * (betax * rx + betay * ry));
		for (int n = 0; n <= L; ++n) {
			double complex prefacn = -I * (kreal ? pow(2. / creal(k), n + 1) : cpow(2. / k, n + 1)) * M_2_SQRTPI / 2;
			double Rn = pow(r, n);
			double complex In = (double complex)sr_integral_c(r, k, n, eta);
			for (int m = -n; m <= n; ++m) {
				if ((m + n) % 2) continue;
				double complex e = cexp(I * m * phi);
				double complex s = prefacn * Rn * leg[n * (n + 1) / 2 + abs(m)] * In * eb * e * ((m < 0) ? min1pow(m) : 1);
				ck(&out[n * (n + 1) + m], &comp[n * (n + 1) + m], s);
			}
		}
	}
	return 0;
}

/* clilgamma (ewald.h:157-170) */
static double complex clilg(double complex z) {
	double complex a1 = z - 1, a2 = z + 1;
	if (creal(a2) < 0 && cimag(a2) <= 0) a2 = -csqrt(a2); else a2 = csqrt(a2);
	if (creal(a1) < 0 && cimag(a1) >= 0) a1 = -csqrt(a1); else a1 = csqrt(a1);
	return a1 * a2;
}

static double complex cpow0(double complex x, int z) {
	if (x == 0 && z == 0) return 1;
	if (x == 0 && z > 0) return 0;
	return cpow(x, z);
}

int or_sigma_long(int L, int csphase, double eta, double complex k, double area, int npts, const double *Kxy,
		double betax, double betay, double sx, double sy, double complex *out) {
	const int nsc = (L + 1) * (L + 1);
	double complex comp[nsc];
	double leg[(L + 1) * (L + 2) / 2];
	ferrers0(L, csphase, leg);
	for (int i = 0; i < nsc; ++i) out[i] = comp[i] = 0;
	for (int ip = 0; ip < npts; ++ip) {
		double Kx = Kxy[2 * ip], Ky = Kxy[2 * ip + 1];
		double bx = Kx + betax, by = Ky + betay;
		double rb = sqrt(bx * bx + by * by), arg = atan2(by, bx);
		double complex phasefac = cexp(I * (Kx * sx + Ky * sy));
		double complex gam = clilg(rb / k);
		double complex x = gam * k / (2 * eta);
		double complex x2 = x * x;
		double complex G[L / 2 + 1];
		for (int j = 0; j <= L / 2; ++j)
			G[j] = or_gamma_inc(0.5 - j, x2, (creal(x2) < 0 && !signbit(cimag(x2))) ? -1 : 0);
		for (int n = 0; n <= L; ++n)
			for (int m = -n; m <= n; ++m) {
				if ((m + n) % 2) continue;
				double complex js = 0, jc = 0;
				for (int j = 0; j <= (n - abs(m)) / 2; ++j) {
					/* s1_constfacs (ewald.c:142-146) */
					double fac = 1;
					for (int i = 2; i <= (n - m) / 2; ++i) fac *= i;
					for (int i = 2; i <= (n + m) / 2; ++i) fac *= i;
					double den = 1;
					for (int i = 2; i <= j; ++i) den *= i;
					for (int i = 2; i <= (n - m) / 2 - j; ++i) den *= i;
					for (int i = 2; i <= (n + m) / 2 - j; ++i) den *= i;
					double complex ipn1 = cpow(I, n + 1);
					double complex cf = -2 * ipn1 * sqrt(M_PI) * fac * min1pow(j) / den;
					double complex s = cpow0(rb / k, n - 2 * j) * cexp(I * m * arg) * leg[n * (n + 1) / 2 + abs(m)]
						* ((m < 0) ? min1pow(m) : 1) * cpow(gam, 2 * j - 1) * cf * G[j];
					ck(&js, &jc, s);
				}
				ck(&out[n * (n + 1) + m], &comp[n * (n + 1) + m], js * phasefac);
			}
	}
	double complex cf = 1 / (k * k * area);
	for (int i = 0; i < nsc; ++i) out[i] *= cf;
	return 0;
}

void or_e32_combine(const or_calc *c, const double complex *sig, int bqmax_mode, double complex *A, double complex *B) {
	const int ne = c->nelem, qs = c->lMax + 1;
	for (int y = 0; y < ne; ++y)
		for (int yu = 0; yu < ne; ++yu) {
			int n = (int)sqrt((double)y + 1), m = y - n * (n + 1) + 1;
			int nu = (int)sqrt((double)yu + 1), mu = yu - nu * (nu + 1) + 1;
			const int i = y * ne + yu;
			double complex sa = 0, ca = 0, sb = 0, cb = 0;
			for (int q = 0; q < c->qA[i]; ++q) {
				int p = n + nu - 2 * q;
				ck(&sa, &ca, c->A[(size_t)i * qs + q] * sig[p * (p + 1) + mu - m]);
			}
			const int nb = bqmax_mode ? c->qB[i] : c->qA[i];
			for (int q = 0; q < nb; ++q) {
				int p = n + nu - 2 * q + 1;
				ck(&sb, &cb, c->B[(size_t)i * qs + q] * sig[p * (p + 1) + mu - m]);
			}
			A[i] = sa;
			B[i] = sb;
		}
}
