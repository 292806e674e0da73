/* Oracle (test infrastructure): regular-VSWF expansion of a plane wave,
 * restating vswf.c:445-530 (qpms_planewave2vswf_fill_sph,
 * qpms_incfield_planewave) with the dual vector spherical harmonics of
 * vswf.c:347-381 and legendre.c:54-110 for the default normalisation
 * (power-normalised, Condon–Shortley phase). The θ-derivative of the
 * Ferrers functions is taken from the ladder identity
 *   dP_l^m/dθ = ½ [P_l^{m+1} − (l+m)(l−m+1) P_l^{m−1}]
 * instead of the x-derivative form. */
#include <math.h>
#include <complex.h>
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
