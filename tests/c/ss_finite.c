/* Finite-cluster plane-wave scattering through the public C ABI, in the
 * order a caller of the reference library uses it (misc/finiterectlat-scatter.py:
 * 80-111 through qpms_c.pyx): scatsys from a sphere T-matrix generator,
 * W, M = I - T W, LU, incident plane wave, T a, solve, extinction.
 *
 * usage: ss_finite OUT nx ny pitch radius epsin_re epsin_im eps_bg eV lMax
 *                  kx ky kz Ex_re Ex_im Ey_re Ey_im Ez_re Ez_im [POINTS]
 * Particles at ((i - (nx-1)/2) pitch, (j - (ny-1)/2) pitch, 0). Writes to
 * OUT: int64 L, then complex128 W[L*L], M[L*L], a[L], Ta[L], f[L], T diag
 * [nb], and float64 sigma_ext, wavenumber (re, im). With POINTS (float64
 * x, y, z per point) it then appends int64 npts and the scattered field of f
 * (complex128 x, y, z per point) from qpms_scatsys_scattered_E,
 * qpms_scatsysw_scattered_E, qpms_scatsysw_scattered_E__alt and
 * qpms_hip_scatsys_scattered_E_points, in that order. */
#include <complex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "qpms_hip.h"

#define EV_RAD_PER_S (1.602176634e-19 / 1.054571817e-34)

int main(int argc, char **argv) {
	if (argc != 20 && argc != 21) {
		fprintf(stderr, "usage: %s OUT nx ny pitch radius epsin_re epsin_im eps_bg eV lMax kx ky kz "
			"Ex_re Ex_im Ey_re Ey_im Ez_re Ez_im [POINTS]\n", argv[0]);
		return 2;
	}
	const char *out = argv[1];
	const int nx = atoi(argv[2]), ny = atoi(argv[3]);
	const double pitch = atof(argv[4]), radius = atof(argv[5]);
	const complex double eps_in = atof(argv[6]) + I * atof(argv[7]);
	const double eps_bg = atof(argv[8]), eV = atof(argv[9]);
	const int lMax = atoi(argv[10]);
	const double kdir[3] = {atof(argv[11]), atof(argv[12]), atof(argv[13])};
	const complex double E[3] = {atof(argv[14]) + I * atof(argv[15]), atof(argv[16]) + I * atof(argv[17]),
		atof(argv[18]) + I * atof(argv[19])};
	if (qpms_hip_init() != 0) return 3;

	qpms_vswf_set_spec_t *bspec = qpms_vswf_set_spec_from_lMax(lMax, QPMS_NORMALISATION_DEFAULT);
	qpms_epsmu_t bg = {eps_bg, 1}, in = {eps_in, 1};
	qpms_tmatrix_generator_sphere_param_t sp = {{qpms_epsmu_const_g, &bg}, {qpms_epsmu_const_g, &in}, radius};
	qpms_tmatrix_generator_t gen = {qpms_tmatrix_generator_sphere, &sp};
	qpms_tmatrix_function_t tmf = {bspec, &gen};
	qpms_ss_derived_tmatrix_t dtm;
	memset(&dtm, 0, sizeof dtm);
	dtm.tmgi = 0;
	dtm.op.typ = QPMS_TMATRIX_OPERATION_NOOP;
	const int np = nx * ny;
	qpms_particle_tid_t *p = calloc(np, sizeof *p);
	for (int i = 0; i < nx; ++i)
		for (int j = 0; j < ny; ++j) {
			p[i * ny + j].pos.x = (i - 0.5 * (nx - 1)) * pitch;
			p[i * ny + j].pos.y = (j - 0.5 * (ny - 1)) * pitch;
			p[i * ny + j].pos.z = 0;
			p[i * ny + j].tmatrix_id = 0;
		}
	qpms_scatsys_t ss0;
	memset(&ss0, 0, sizeof ss0);
	ss0.lattice_dimension = 0;
	ss0.medium.function = qpms_epsmu_const_g;
	ss0.medium.params = &bg;
	ss0.tmg = &tmf; ss0.tmg_count = 1;
	ss0.tm = &dtm; ss0.tm_count = 1;
	ss0.p = p; ss0.p_count = np;

	const complex double omega = eV * EV_RAD_PER_S;
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_apply_symmetry(&ss0, NULL, omega, NULL);
	const qpms_scatsys_t *ss = ssw->ss;
	const size_t L = ss->fecv_size;
	const complex double k = ssw->wavenumber;

	complex double *W = qpms_scatsys_build_translation_matrix_full(NULL, ss, k);
	complex double *M = qpms_scatsys_build_modeproblem_matrix_full(NULL, ssw);
	qpms_ss_LU lu = qpms_scatsysw_build_modeproblem_matrix_full_LU(NULL, NULL, ssw);

	qpms_incfield_planewave_params_t pw;
	memset(&pw, 0, sizeof pw);
	pw.use_cartesian = true;
	pw.k.cart.x = creal(k) * kdir[0]; pw.k.cart.y = creal(k) * kdir[1]; pw.k.cart.z = creal(k) * kdir[2];
	pw.E.cart.x = E[0]; pw.E.cart.y = E[1]; pw.E.cart.z = E[2];
	complex double *a = qpms_scatsys_incident_field_vector_full(NULL, ss, qpms_incfield_planewave, &pw, false);
	complex double *Ta = qpms_scatsysw_apply_Tmatrices_full(NULL, a, ssw);
	complex double *f = qpms_scatsys_scatter_solve(NULL, Ta, lu);
	complex double dot = 0;
	for (size_t i = 0; i < L; ++i) dot += conj(a[i]) * f[i];
	const double sigma_ext = -creal(dot) / creal(k * conj(k));   /* real k here */

	FILE *fo = fopen(out, "wb");
	if (!fo) { perror(out); return 4; }
	const long long LL = (long long)L;
	const size_t nb = bspec->n;
	complex double *td = malloc(nb * sizeof *td);
	for (size_t i = 0; i < nb; ++i) td[i] = ssw->tm[0]->m[i * nb + i];
	const double kk[2] = {creal(k), cimag(k)};
	int ok = fwrite(&LL, sizeof LL, 1, fo) == 1
		&& fwrite(W, sizeof *W, L * L, fo) == L * L && fwrite(M, sizeof *M, L * L, fo) == L * L
		&& fwrite(a, sizeof *a, L, fo) == L && fwrite(Ta, sizeof *Ta, L, fo) == L && fwrite(f, sizeof *f, L, fo) == L
		&& fwrite(td, sizeof *td, nb, fo) == nb && fwrite(&sigma_ext, sizeof sigma_ext, 1, fo) == 1
		&& fwrite(kk, sizeof *kk, 2, fo) == 2;
	if (ok && argc == 21) {
		FILE *fp = fopen(argv[20], "rb");
		if (!fp) { perror(argv[20]); return 4; }
		fseek(fp, 0, SEEK_END);
		const long long npts = ftell(fp) / (long long)sizeof(cart3_t);
		fseek(fp, 0, SEEK_SET);
		cart3_t *w = malloc((size_t)npts * sizeof *w);
		ccart3_t *e = malloc((size_t)npts * 4 * sizeof *e);
		ok = fread(w, sizeof *w, (size_t)npts, fp) == (size_t)npts;
		fclose(fp);
		for (long long t = 0; ok && t < npts; ++t) {
			e[t] = qpms_scatsys_scattered_E(ss, QPMS_HANKEL_PLUS, k, f, w[t]);
			e[npts + t] = qpms_scatsysw_scattered_E(ssw, QPMS_HANKEL_PLUS, f, w[t]);
			e[2 * npts + t] = qpms_scatsysw_scattered_E__alt(ssw, QPMS_HANKEL_PLUS, f, w[t]);
		}
		if (ok) qpms_hip_scatsys_scattered_E_points(ss, QPMS_HANKEL_PLUS, k, f, (size_t)npts, w, e + 3 * npts);
		ok = ok && fwrite(&npts, sizeof npts, 1, fo) == 1
			&& fwrite(e, sizeof *e, (size_t)npts * 4, fo) == (size_t)npts * 4;
		free(w); free(e);
	}
	fclose(fo);
	printf("L=%zu sigma_ext=%.17g\n", L, sigma_ext);
	free(td); free(a); free(Ta); free(f); free(W); free(M);
	qpms_ss_LU_free(lu);
	qpms_tmatrix_t **tm = ssw->tm;
	qpms_scatsys_t *ssm = (qpms_scatsys_t *)ss;
	for (qpms_ss_tmi_t i = 0; i < ss->tm_count; ++i) qpms_tmatrix_free(tm[i]);
	free(tm);
	free(ssw);
	qpms_scatsys_free(ssm);
	qpms_vswf_set_spec_free(bspec);
	free(p);
	return ok ? 0 : 5;
}
