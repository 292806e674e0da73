/* CPU oracle for the multiple-scattering hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain C99 restatement of the reference algorithms (citations are
 * `path:line` into kostyfisik/qpms), used as the parity checker for the GPU
 * library and as the timed CPU baseline. Nothing in the product library links
 * or calls it: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load liboracle.so.
 *
 * Numerics: long double internally where cheap; Kahan-compensated sums as in
 * the reference (kahansum.h:24-29). How each function is pinned (golden
 * vectors, compiled reference pieces, closed forms, invariants) is recorded
 * in tests/test_oracle_*.py.
 */
#ifndef QPMS_ORACLE_H
#define QPMS_ORACLE_H
#include <complex.h>
#include <stddef.h>

/* Ferrers P_l^m(x), 0 <= m <= l <= lmax, unnormalised, CS phase (-1)^m iff
 * csphase == -1, index l(l+1)/2 + m (gsl_sf_legendre_array_e NONE,
 * translations.c:652-653). */
void or_legendre_array(int lmax, double x, int csphase, double *out);

/* Spherical Bessel kind J (1 j, 2 y, 3 h+, 4 h-) for l = 0..lmax at complex
 * x (bessel.c:86-147). */
void or_sph_bessel(int J, int lmax, double complex x, double complex *out);

/* Gaunt coefficient a_q(m,n,mu,nu) (gaunt.c:396, Xu 1996 eq. 4), by
 * Gauss–Legendre quadrature of the Ferrers triple product. */
double or_gaunt(int m, int n, int mu, int nu, int q);
/* Wigner 3j by Racah's formula in long double. */
double or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);

/* Translation calculator tables (translations.c:313-508). */
typedef struct or_calc {
	int lMax, nelem, norm;
	double complex *A, *B;   /* [(y*nelem + yu) * (lMax+1) + q] */
	int *qA, *qB;            /* term counts per (y, yu) */
} or_calc;
or_calc *or_calc_init(int lMax, int norm);
void or_calc_free(or_calc *c);
/* Raw multiplier access (for table comparison). */
int or_calc_multipliers(const or_calc *c, int y, int yu, double complex *A, int *nA, double complex *B, int *nB);

/* AB arrays (translations.c:610-698): A[y*nelem + yu]. */
void or_get_AB_arrays(const or_calc *c, double complex kr, double theta, double phi, int r_ge_d, int J,
		double complex *A, double complex *B);
/* Single-coefficient formulas (translations.c:135-286) with directly
 * evaluated Gaunt coefficients: an independent second formula. */
double complex or_trans_single_A(int norm, int m, int n, int mu, int nu, double complex kr, double theta,
		double phi, int J);
double complex or_trans_single_B(int norm, int m, int n, int mu, int nu, double complex kr, double theta,
		double phi, int J);

/* Full W (scatsystem.c:1180-1217) for npart particles with the default bspec
 * of c->lMax: rows [row0, row0+nrows) of particles, ld = 2*nelem*npart. */
void or_build_W_rows(const or_calc *c, int npart, const double *pos, double complex k, int J,
		int row0, int nrows, double complex *W);
/* Mode-problem M = I - diag(t) W (scatsystem.c:1301-1339, intended finite
 * semantics); tdiag[ttype[p]*nb + i]. */
void or_build_M(const or_calc *c, int npart, const double *pos, const int *ttype,
		const double complex *tdiag, double complex k, double complex *M);

/* Mie reflection coefficients RV[l], RH[l], l = 0..lMax (tmatrices.c:500-561). */
void or_mie(int lMax, double a, double complex ki, double complex ke, double complex mui,
		double complex mue, double complex *RV, double complex *RH);

/* Dense LU with partial pivoting, row-major, ipiv 1-based (zgetrf), and a
 * solve with nrhs row-major right-hand sides (zgetrs); pthreads over rows. */
int or_zgetrf(int n, double complex *A, int *ipiv, int nthreads);
void or_zgetrs(int n, int nrhs, const double complex *LU, const int *ipiv, double complex *B);

/* Ewald (ewald.c): xyWeb point enumeration and the σ sums. */
int or_xyweb_points(double b1x, double b1y, double b2x, double b2y, double rtol, double minR, int inc_minR,
		double maxR, int inc_maxR, double *xy, int cap);
int or_sigma0(int lMax_e, double eta, double complex k, double complex *out);
int or_sigma_short(int lMax_e, int csphase, double eta, double complex k, int npts, const double *Rxy,
		double betax, double betay, double sx, double sy, double complex *out);
int or_sigma_long(int lMax_e, int csphase, double eta, double complex k, double unitcell_area, int npts,
		const double *Kxy, double betax, double betay, double sx, double sy, double complex *out);
double complex or_gamma_inc(double a, double complex x, int m);
double or_sr_integral(double r, double k, int n, double eta);

/* e32 combine (translations.c:1024-1067); bqmax_mode 0 = reference (B bounded
 * by A's qmax), 1 = B's own bound. */
void or_e32_combine(const or_calc *c, const double complex *sigma_total, int bqmax_mode,
		double complex *A, double complex *B);

/* Plane-wave regular-VSWF coefficients at npart points (vswf.c:445-530):
 * out[p*2*nelem + y] electric, out[p*2*nelem + nelem + y] magnetic; real
 * wave vector kvec, complex cartesian amplitude E[3]. */
void or_planewave(int lMax, int npart, const double *pos, const double *kvec, const double complex *E,
		double complex *out);

/* Field of the coefficient vector cvf (npart blocks of 2*nelem, default
 * basis) at npts cartesian points: out[3*t + (x, y, z)], waves of kind J at
 * wavenumber k (qpms_scatsys_scattered_E, scatsystem.c:2013-2037). */
void or_scattered_E(int lMax, int npart, const double *pos, double complex k, int J, const double complex *cvf,
		int npts, const double *where, double complex *out);
/* The same sum as the reference's __alt variant computes it (scatsystem.c:
 * 2048-2109): translation arrays with destination = regular electric dipoles
 * at the particle, source = the point, times the dipoles' value at r = 0. */
void or_scattered_E_alt(const or_calc *c, int npart, const double *pos, double complex k, int J,
		const double complex *cvf, int npts, const double *where, double complex *out);

#endif
