/* 2D-periodic mode problem through the public C ABI (the path of
 * qpms_c.pyx's ScatteringSystem for lattices): scatsys with a 2D lattice,
 * qpms_scatsyswk_build_translation_matrix_full and
 * qpms_scatsyswk_build_modeproblem_matrix_full at a Bloch vector, then the
 * FP64 LU of M and log|det M|.
 *
 * usage: ss_periodic OUT a lMax radius epsin_re epsin_im eps_bg eV beta_x beta_y [POINTS]
 * Hexagonal lattice b1 = (a, 0), b2 = (a/2, a sqrt3/2) with particles at
 * (0,0,0) and (a/sqrt3, 0, 0). Writes int64 L, double eta, double
 * wavenumber (re, im), double lattice_basis[2][3], complex128 W[L*L], M[L*L],
 * T diag[nb], LU[L*L], int32 ipiv[L]. With POINTS (float64 x, y, z per
 * point) it appends int64 npts and qpms_scatsyswk_scattered_E of the
 * coefficient vector c_i = exp(0.7 i I) / (1 + i) at each point (complex128
 * x, y, z). */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "qpms_hip.h"

#define EV_RAD_PER_S (1.602176634e-19 / 1.054571817e-34)

int main(int argc, char **argv) {
	if (argc != 11 && argc != 12) {
		fprintf(stderr, "usage: %s OUT a lMax radius epsin_re epsin_im eps_bg eV beta_x beta_y [POINTS]\n", argv[0]);
		return 2;
	}
	const double a = atof(argv[2]);
	const int lMax = atoi(argv[3]);
	const double radius = atof(argv[4]);
	const complex double eps_in = atof(argv[5]) + I * atof(argv[6]);
	const double eps_bg = atof(argv[7]), eV = atof(argv[8]);
	const double beta[3] = {atof(argv[9]), atof(argv[10]), 0};
	if (qpms_hip_init() != 0) return 3;

	qpms_vswf_set_spec_t *bspec = qpms_vswf_set_spec_from_lMax(lMax, QPMS_NORMALISATION_DEFAULT);
	qpms_epsmu_t bg = {eps_bg, 1}, in = {eps_in, 1};
	qpms_tmatrix_generator_sphere_param_t sp = {{qpms_epsmu_const_g, &bg}, {qpms_epsmu_const_g, &in}, radius};
	qpms_tmatrix_generator_t gen = {qpms_tmatrix_generator_sphere, &sp};
	qpms_tmatrix_function_t tmf = {bspec, &gen};
	qpms_ss_derived_tmatrix_t dtm;
	memset(&dtm, 0, sizeof dtm);
	dtm.op.typ = QPMS_TMATRIX_OPERATION_NOOP;
	qpms_particle_tid_t p[2];
	memset(p, 0, sizeof p);
	p[1].pos.x = a / sqrt(3.0);
	qpms_scatsys_t ss0;
	memset(&ss0, 0, sizeof ss0);
	ss0.lattice_dimension = 2;
	ss0.medium.function = qpms_epsmu_const_g;
	ss0.medium.params = &bg;
	ss0.tmg = &tmf; ss0.tmg_count = 1;
	ss0.tm = &dtm; ss0.tm_count = 1;
	ss0.p = p; ss0.p_count = 2;
	ss0.per.lattice_basis[0].x = a;
	ss0.per.lattice_basis[1].x = a / 2; ss0.per.lattice_basis[1].y = a * sqrt(3.0) / 2;

	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_apply_symmetry(&ss0, NULL, eV * EV_RAD_PER_S, NULL);
	const qpms_scatsys_t *ss = ssw->ss;
	const size_t L = ss->fecv_size;
	qpms_scatsys_at_omega_k_t sswk;
	sswk.ssw = ssw;
	memcpy(sswk.k, beta, sizeof beta);
	sswk.eta = qpms_ss_adjusted_eta(ss, ssw->wavenumber, beta);
	complex double *W = qpms_scatsyswk_build_translation_matrix_full(NULL, &sswk);
	complex double *M = qpms_scatsyswk_build_modeproblem_matrix_full(NULL, &sswk);
	complex double *LU = malloc(L * L * sizeof *LU);
	memcpy(LU, M, L * L * sizeof *LU);
	qpms_ss_LU lu = qpms_scatsysw_modeproblem_matrix_full_factorise(LU, NULL, NULL, &sswk);

	FILE *fo = fopen(argv[1], "wb");
	if (!fo) { perror(argv[1]); return 4; }
	const long long LL = (long long)L;
	const size_t nb = bspec->n;
	complex double *td = malloc(nb * sizeof *td);
	for (size_t i = 0; i < nb; ++i) td[i] = ssw->tm[0]->m[i * nb + i];
	double hdr[1 + 2 + 6] = {sswk.eta, creal(ssw->wavenumber), cimag(ssw->wavenumber),
		ss->per.lattice_basis[0].x, ss->per.lattice_basis[0].y, ss->per.lattice_basis[0].z,
		ss->per.lattice_basis[1].x, ss->per.lattice_basis[1].y, ss->per.lattice_basis[1].z};
	int ok = fwrite(&LL, sizeof LL, 1, fo) == 1 && fwrite(hdr, sizeof hdr, 1, fo) == 1
		&& fwrite(W, sizeof *W, L * L, fo) == L * L && fwrite(M, sizeof *M, L * L, fo) == L * L
		&& fwrite(td, sizeof *td, nb, fo) == nb && fwrite(lu.a, sizeof *lu.a, L * L, fo) == L * L
		&& fwrite(lu.ipiv, sizeof *lu.ipiv, L, fo) == L;
	if (ok && argc == 12) {
		FILE *fp = fopen(argv[11], "rb");
		if (!fp) { perror(argv[11]); return 4; }
		fseek(fp, 0, SEEK_END);
		const long long npts = ftell(fp) / (long long)sizeof(cart3_t);
		fseek(fp, 0, SEEK_SET);
		cart3_t *w = malloc((size_t)npts * sizeof *w);
		ccart3_t *e = malloc((size_t)npts * sizeof *e);
		complex double *cv = malloc(L * sizeof *cv);
		for (size_t i = 0; i < L; ++i) cv[i] = cexp(0.7 * I * i) / (1.0 + i);
		ok = fread(w, sizeof *w, (size_t)npts, fp) == (size_t)npts;
		fclose(fp);
		for (long long t = 0; ok && t < npts; ++t) e[t] = qpms_scatsyswk_scattered_E(&sswk, QPMS_HANKEL_PLUS, cv, w[t]);
		ok = ok && fwrite(&npts, sizeof npts, 1, fo) == 1 && fwrite(e, sizeof *e, (size_t)npts, fo) == (size_t)npts;
		free(w); free(e); free(cv);
	}
	fclose(fo);
	printf("L=%zu eta=%.17g\n", L, sswk.eta);
	free(td); free(W); free(M);
	qpms_ss_LU_free(lu);
	for (qpms_ss_tmi_t i = 0; i < ss->tm_count; ++i) qpms_tmatrix_free(ssw->tm[i]);
	free(ssw->tm);
	qpms_scatsys_free((qpms_scatsys_t *)ss);
	free(ssw);
	qpms_vswf_set_spec_free(bspec);
	return ok ? 0 : 5;
}
