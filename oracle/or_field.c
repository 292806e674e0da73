/* Oracle (test infrastructure): regular-VSWF expansion of a plane wave,
 * restating vswf.c:445-530 (qpms_planewave2vswf_fill_sph,
 * qpms_incfield_planewave) with the dual vector spherical harmonics of
 * vswf.c:347-381 and legendre.c:54-110 for the default normalisation
 * (power-normalised, Condon–Shortley phase). The θ-derivative of the
 * Ferrers functions is taken from the ladder identity
 *   dP_l^m/dθ = ½ [P_l^{m+1} − (l+m)(l−m+1) P_l^{m−1}]
 * instead of the x-derivative form.
 *
 * Also the field of a coefficient vector at given points: VSWF evaluation
 * (vswf.c:230-288, 611-655) summed over particles (scatsystem.c:2013-2037),
 * and the reference's translation-based variant of the same sum
 * (scatsystem.c:2048-2109). */
#include <math.h>
#include <complex.h>
#include <stdlib.h>
#include "oracle.h"

typedef long double ld;

/* Ferrers P_l^m(cos θ) with CS phase, any integer m with |m| <= l (else 0). */
static ld ferrers_any(int l, int m, ld x, ld s) {
	if (m > l || -m > l) return 0;
	const int am = m < 0 ? -m : m;
	ld pmm = 1;
	for (int k = 1; k <= am; ++k) pmm = -pmm * (2 * k - 1) * s;
	ld val;
	if (l == am) val = pmm;
	else {
		ld p0 = pmm, p1 = x * (2 * am + 1) * pmm;
		for (int k = am + 2; k <= l; ++k) {
			ld pk = ((2 * k - 1) * x * p1 - (k + am - 1) * p0) / (k - am);
			p0 = p1; p1 = pk;
		}
		val = p1;
	}
	if (m < 0) {   /* DLMF 14.9.3 */
		ld r = 1;
		for (int k = l - am + 1; k <= l + am; ++k) r *= k;
		val *= ((am & 1) ? -1 : 1) / r;
	}
	return val;
}

static void pitau(int l, int m, double theta, ld *leg, ld *pi, ld *tau) {
	const ld x = cosl(theta), s = sinl(theta);
	const ld PI_ = 3.14159265358979323846264338327950288L;
	if (fabsl((ld)cos(theta)) == 1) {     /* legendre.c:65-80 */
		const ld ct = cos(theta);
		*leg = *pi = *tau = 0;
		if (m == 0) *leg = ((l % 2) ? ct : 1) * sqrtl((2 * l + 1) / (4 * PI_ * l * (l + 1)));
		const ld fl = 0.25L * sqrtl((2 * l + 1) / PI_);
		const int lpar = (l % 2) ? -1 : 1;
		const ld cs = -1;
		if (m == 1 || m == -1) *pi = -((ct > 0) ? -1 : lpar) * fl * cs;
		if (m == 1) *tau = ((ct > 0) ? 1 : lpar) * fl * cs;
		if (m == -1) *tau = -((ct > 0) ? 1 : lpar) * fl * cs;
		return;
	}
	const int am = m < 0 ? -m : m;
	ld fr = 1;
	for (int k = l - am + 1; k <= l + am; ++k) fr *= k;
	ld N = sqrtl((2 * l + 1) / (4 * PI_ * fr) / (l * (l + 1)));
	const ld P = ferrers_any(l, am, x, s);
	const ld dP = 0.5L * (ferrers_any(l, am + 1, x, s) - (ld)(l + am) * (l - am + 1) * ferrers_any(l, am - 1, x, s));
	if (m < 0 && (am & 1)) N = -N;
	*leg = N * P;
	*tau = N * dP;
	*pi = m / s * (*leg);
}

void or_planewave(int lMax, int npart, const double *pos, const double *kvec, const double complex *E,
		double complex *out) {
	const int ne = lMax * (lMax + 2);
	const double kn = sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
	const double theta = kn ? acos(kvec[2] / kn) : M_PI_2, phi = atan2(kvec[1], kvec[0]);
	const double ct = cos(theta), st = sin(theta), cf = cos(phi), sf = sin(phi);
	const double complex Et = ct * cf * E[0] + ct * sf * E[1] - st * E[2];
	const double complex Ef = -sf * E[0] + cf * E[1];
	for (int p = 0; p < npart; ++p) {
		const double complex ph = cexp(I * (kvec[0] * pos[3 * p] + kvec[1] * pos[3 * p + 1] + kvec[2] * pos[3 * p + 2]));
		for (int l = 1; l <= lMax; ++l)
			for (int m = -l; m <= l; ++m) {
				const int y = l * (l + 1) + m - 1;
				ld leg, pi, tau;
				pitau(l, m, theta, &leg, &pi, &tau);
				const double complex e = cexp(-I * m * phi);
				const double complex d = m ? -I * e : 0;
				const double complex a1t = (double)pi * d, a1f = -(double)tau * e;
				const double complex a2t = (double)tau * e, a2f = (double)pi * d;
				const double complex il = cpow(I, l);
				out[(size_t)p * 2 * ne + y] = -4 * M_PI * il * I * (a2t * Et + a2f * Ef) * ph;
				out[(size_t)p * 2 * ne + ne + y] = 4 * M_PI * il * (a1t * Et + a1f * Ef) * ph;
			}
	}
}

/* VSWFs of the default normalisation at kr = k r (complex), direction (θ, φ),
 * restating qpms_vswf_fill_csph (vswf.c:230-288): spherical components
 * (r, θ, φ) of N_lm (electric) and M_lm (magnetic), y = l(l+1) + m - 1.
 * At kr = 0 (regular waves only) z_l/kr is 1/3 for l = 1 and 0 otherwise. */
static void or_vswf_at(int lMax, int J, double complex kr, double theta, double phi,
		double complex (*N)[3], double complex (*M)[3]) {
	double complex z[lMax + 1];
	if (kr == 0) {
		for (int l = 0; l <= lMax; ++l) z[l] = (J == 1) ? (l == 0) : NAN;
	} else
		or_sph_bessel(J, lMax, kr, z);
	for (int l = 1; l <= lMax; ++l) {
		const double complex besfac = (kr != 0) ? z[l] / kr : ((l == 1) ? 1.0 / 3 : 0);
		const double complex besder = z[l - 1] - l * besfac;
		for (int m = -l; m <= l; ++m) {
			const int y = l * (l + 1) + m - 1;
			ld leg, pi, tau;
			pitau(l, m, theta, &leg, &pi, &tau);
			const double complex e = cexp(I * m * phi);
			const double complex d = m ? I * e : 0;   /* normalisation.h:197-212 */
			N[y][0] = l * (l + 1) * (double)leg * besfac * e;
			N[y][1] = (double)tau * besder * e;
			N[y][2] = (double)pi * besder * d;
			M[y][0] = 0;
			M[y][1] = (double)pi * z[l] * d;
			M[y][2] = -(double)tau * z[l] * e;
		}
	}
}

/* Spherical (r, θ, φ) components at (θ, φ) to cartesian (vectors.h:272-291). */
static void or_sph2cart_vec(const long double complex v[3], double theta, double phi, long double complex out[3]) {
	const ld st = sin(theta), ct = cos(theta), sf = sin(phi), cf = cos(phi);
	out[0] = st * cf * v[0] + ct * cf * v[1] - sf * v[2];
	out[1] = st * sf * v[0] + ct * sf * v[1] + cf * v[2];
	out[2] = ct * v[0] - st * v[1];
}

/* cart2sph (vectors.h:134-140). */
static void or_cart2sph(const double d[3], double *r, double *theta, double *phi) {
	*r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
	*theta = *r ? acos(d[2] / *r) : M_PI_2;
	*phi = atan2(d[1], d[0]);
}

void or_scattered_E(int lMax, int npart, const double *pos, double complex k, int J, const double complex *cvf,
		int npts, const double *where, double complex *out) {
	const int ne = lMax * (lMax + 2);
	double complex (*N)[3] = malloc(sizeof(*N) * ne), (*M)[3] = malloc(sizeof(*M) * ne);
	for (int t = 0; t < npts; ++t) {
		long double complex acc[3] = {0, 0, 0};
		for (int p = 0; p < npart; ++p) {
			const double d[3] = {where[3 * t] - pos[3 * p], where[3 * t + 1] - pos[3 * p + 1],
				where[3 * t + 2] - pos[3 * p + 2]};
			double r, th, ph;
			or_cart2sph(d, &r, &th, &ph);
			or_vswf_at(lMax, J, k * r, th, ph, N, M);
			const double complex *c = cvf + (size_t)p * 2 * ne;
			long double complex es[3] = {0, 0, 0}, ec[3];
			for (int y = 0; y < ne; ++y)
				for (int q = 0; q < 3; ++q)
					es[q] += (long double complex)c[y] * N[y][q] + (long double complex)c[ne + y] * M[y][q];
			or_sph2cart_vec(es, th, ph, ec);
			for (int q = 0; q < 3; ++q) acc[q] += ec[q];
		}
		for (int q = 0; q < 3; ++q) out[3 * t + q] = (double complex)acc[q];
	}
	free(N); free(M);
}

void or_scattered_E_alt(const or_calc *c, int npart, const double *pos, double complex k, int J,
		const double complex *cvf, int npts, const double *where, double complex *out) {
	const int lMax = c->lMax, ne = c->nelem;
	/* regular electric dipoles at the origin, evaluated at θ = π/2, φ = 0 */
	double complex N[3][3], Mg[3][3];
	or_vswf_at(1, 1, 0, M_PI_2, 0, N, Mg);
	long double complex dip[3][3];
	for (int j = 0; j < 3; ++j) {
		long double complex v[3] = {N[j][0], N[j][1], N[j][2]};
		or_sph2cart_vec(v, M_PI_2, 0, dip[j]);
	}
	double complex *A = malloc(sizeof(*A) * ne * ne), *B = malloc(sizeof(*B) * ne * ne);
	for (int t = 0; t < npts; ++t) {
		long double complex acc[3] = {0, 0, 0};
		for (int p = 0; p < npart; ++p) {
			/* destination: the dipoles at the particle; source: the point */
			const double d[3] = {pos[3 * p] - where[3 * t], pos[3 * p + 1] - where[3 * t + 1],
				pos[3 * p + 2] - where[3 * t + 2]};
			double r, th, ph;
			or_cart2sph(d, &r, &th, &ph);
			or_get_AB_arrays(c, k * r, th, ph, 0, J, A, B);
			const double complex *cv = cvf + (size_t)p * 2 * ne;
			for (int i = 0; i < 2 * ne; ++i) {
				const int srcy = i % ne, electric = i < ne;
				for (int j = 0; j < 3; ++j) {
					/* dipoles are electric: same type -> A, else B (translations_inlines.h:8-35) */
					const double complex s = electric ? A[j * ne + srcy] : B[j * ne + srcy];
					for (int q = 0; q < 3; ++q) acc[q] += (long double complex)(cv[i] * s) * dip[j][q];
				}
			}
		}
		for (int q = 0; q < 3; ++q) out[3 * t + q] = (double complex)acc[q];
	}
	(void)lMax;
	free(A); free(B);
}
