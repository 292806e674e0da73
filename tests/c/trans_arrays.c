/* Drives the remaining translation and eigenmode entry points of the public
 * C ABI as a C caller of the reference library would:
 *   qpms_trans_calculator_get_trans_array / _lc3p   (translations.c:701-812)
 *   qpms_trans_calculator_get_AB_arrays_e32 and get_trans_array_e32(_e)
 *                                                   (translations.c:724-1094)
 *   qpms_scatsys_build_translation_matrix_full with particles of two
 *   different bases (scatsystem.c:1180-1217)
 *   qpms_scatsys_finite_find_eigenmodes             (scatsystem.c:2301-2325)
 *   qpms_scatsys_periodic_find_eigenmodes           (scatsystem.c:2351-2370)
 *
 * usage: trans_arrays OUTDIR
 * Writes raw little-endian arrays into OUTDIR, one file per quantity (the
 * test reads them with the shapes it knows). */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "qpms_hip.h"

#define EV_RAD_PER_S (1.602176634e-19 / 1.054571817e-34)

static const char *g_dir;

static void dump(const char *name, const void *p, size_t bytes) {
	char path[4096];
	snprintf(path, sizeof path, "%s/%s.bin", g_dir, name);
	FILE *f = fopen(path, "wb");
	if (!f || fwrite(p, 1, bytes, f) != bytes) { fprintf(stderr, "cannot write %s\n", path); exit(4); }
	fclose(f);
}

/* A basis of lMax in the reverse of the default order (magnetic block first,
 * m from +n down to -n): every element moves. */
static qpms_vswf_set_spec_t *reversed_spec(int lMax) {
	qpms_vswf_set_spec_t *s = qpms_vswf_set_spec_from_lMax(lMax, QPMS_NORMALISATION_DEFAULT);
	for (size_t i = 0; i < s->n / 2; ++i) {
		qpms_uvswfi_t t = s->ilist[i];
		s->ilist[i] = s->ilist[s->n - 1 - i];
		s->ilist[s->n - 1 - i] = t;
	}
	return s;
}

/* The electric waves of lMax only (n = nelem). */
static qpms_vswf_set_spec_t *electric_spec(int lMax) {
	qpms_vswf_set_spec_t *s = qpms_vswf_set_spec_from_lMax(lMax, QPMS_NORMALISATION_DEFAULT);
	size_t k = 0;
	for (size_t i = 0; i < s->n; ++i)
		if ((s->ilist[i] & 3) == QPMS_VSWF_ELECTRIC) s->ilist[k++] = s->ilist[i];
	s->n = k;
	s->lMax_M = 0;
	return s;
}

/* M(omega) of the 2D-periodic system at Bloch vector k: I - T(omega) W(omega, k)
 * at the adjusted Ewald parameter, through the public builder (the callback a
 * caller of beyn_solve writes for the periodic problem, scatsystem.c:2333-2349). */
static int periodic_M(complex double *target, size_t m, complex double omega, void *params) {
	const struct { const qpms_scatsys_t *ss; const double *k; } *p = params;
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_at_omega(p->ss, omega);
	if (m != p->ss->fecv_size) return 1;
	qpms_scatsys_at_omega_k_t sswk;
	sswk.ssw = ssw;
	sswk.k[0] = p->k[0]; sswk.k[1] = p->k[1]; sswk.k[2] = p->k[2];
	sswk.eta = qpms_ss_adjusted_eta(p->ss, ssw->wavenumber, p->k);
	qpms_scatsyswk_build_modeproblem_matrix_full(target, &sswk);
	qpms_scatsys_at_omega_free(ssw);
	return 0;
}

static void dump_spec(const char *name, const qpms_vswf_set_spec_t *s) {
	dump(name, s->ilist, s->n * sizeof(qpms_uvswfi_t));
}

int main(int argc, char **argv) {
	if (argc != 2) { fprintf(stderr, "usage: %s OUTDIR\n", argv[0]); return 2; }
	g_dir = argv[1];
	if (qpms_hip_init() != 0) return 3;

	/* ---- 1. get_trans_array and _lc3p, lMax 3 calculator ---------------- */
	qpms_trans_calculator *c3 = qpms_trans_calculator_init(3, QPMS_NORMALISATION_DEFAULT);
	qpms_vswf_set_spec_t *dst = qpms_vswf_set_spec_from_lMax(3, QPMS_NORMALISATION_DEFAULT);
	qpms_vswf_set_spec_t *src1 = reversed_spec(3), *src2 = electric_spec(2);
	dump_spec("spec_dst", dst);
	dump_spec("spec_src1", src1);
	dump_spec("spec_src2", src2);
	const double kd[4][4] = {{2.0, 0.0, 0.5, 0.5}, {7.0, 0.3, 1.2, -2.0}, {15.0, 0.0, M_PI, 1.0}, {0.8, 0.0, 2.0, 0.3}};
	complex double *T1 = malloc(4 * dst->n * src1->n * sizeof(complex double));
	complex double *T2 = malloc(4 * dst->n * src2->n * sizeof(complex double));
	for (int i = 0; i < 4; ++i) {
		csph_t k = {kd[i][0] + I * kd[i][1], kd[i][2], kd[i][3]};
		/* row-major dst x src (deststride = src->n, srcstride = 1) */
		qpms_trans_calculator_get_trans_array(c3, T1 + i * dst->n * src1->n, dst, src1->n, src1, 1, k, false,
			QPMS_HANKEL_PLUS);
		/* column-major: deststride 1, srcstride dst->n, regular waves */
		qpms_trans_calculator_get_trans_array(c3, T2 + i * dst->n * src2->n, dst, 1, src2, dst->n, k, false,
			QPMS_BESSEL_REGULAR);
	}
	dump("trans1", T1, 4 * dst->n * src1->n * sizeof(complex double));
	dump("trans2", T2, 4 * dst->n * src2->n * sizeof(complex double));
	const cart3_t pd = {120e-9, -40e-9, 75e-9}, ps = {-310e-9, 205e-9, -15e-9};
	const complex double klc = 1.52 * 2.1 * EV_RAD_PER_S / SPEED_OF_LIGHT * (1 + 0.01 * I);
	complex double *Tl = malloc(dst->n * src1->n * sizeof(complex double));
	qpms_trans_calculator_get_trans_array_lc3p(c3, Tl, dst, src1->n, src1, 1, klc, pd, ps, QPMS_HANKEL_PLUS);
	dump("trans_lc3p", Tl, dst->n * src1->n * sizeof(complex double));
	const double lcin[8] = {pd.x, pd.y, pd.z, ps.x, ps.y, ps.z, creal(klc), cimag(klc)};
	dump("lc3p_in", lcin, sizeof lcin);

	/* ---- 2. e32: AB arrays and trans arrays on the hexagonal lattice ---- */
	qpms_trans_calculator *c2 = qpms_trans_calculator_init(2, QPMS_NORMALISATION_DEFAULT);
	const double a = 375e-9;
	const cart2_t b1 = {a, 0}, b2 = {a / 2, a * sqrt(3) / 2};
	const double area = a * a * sqrt(3) / 2, eta = 2.5 / a;
	const cart2_t beta = {0.31 * 2 * M_PI / a, -0.17 * 2 * M_PI / a};
	const cart3_t shift = {a / sqrt(3), 0, 0};
	const complex double ke = 1.52 * 1.9 * EV_RAD_PER_S / SPEED_OF_LIGHT;
	const double maxR = 64 * sqrt(area), maxK = 2048 * 2 * M_PI / maxR;
	const size_t ne2 = c2->nelem;
	complex double *A = malloc(ne2 * ne2 * sizeof(complex double)), *B = malloc(ne2 * ne2 * sizeof(complex double));
	qpms_trans_calculator_get_AB_arrays_e32(c2, A, NULL, B, NULL, ne2, 1, eta, ke, b1, b2, beta, shift, maxR, maxK);
	dump("e32_A", A, ne2 * ne2 * sizeof(complex double));
	dump("e32_B", B, ne2 * ne2 * sizeof(complex double));
	qpms_vswf_set_spec_t *d2 = qpms_vswf_set_spec_from_lMax(2, QPMS_NORMALISATION_DEFAULT), *s2r = reversed_spec(2);
	complex double *Te = malloc(d2->n * s2r->n * sizeof(complex double));
	double *Terr = malloc(d2->n * s2r->n * sizeof(double));
	qpms_trans_calculator_get_trans_array_e32(c2, Te, Terr, d2, s2r->n, s2r, 1, eta, ke, b1, b2, beta, shift, maxR,
		maxK);
	dump("e32_trans", Te, d2->n * s2r->n * sizeof(complex double));
	/* the parts of the _e variant add up to the whole */
	complex double *Tp = calloc(d2->n * s2r->n, sizeof(complex double));
	complex double *Tq = malloc(d2->n * s2r->n * sizeof(complex double));
	const qpms_ewald_part parts[3] = {QPMS_EWALD_SHORT_RANGE, QPMS_EWALD_LONG_RANGE, QPMS_EWALD_0TERM};
	for (int q = 0; q < 3; ++q) {
		qpms_trans_calculator_get_trans_array_e32_e(c2, Tq, NULL, d2, s2r->n, s2r, 1, eta, ke, b1, b2, beta, shift,
			maxR, maxK, parts[q]);
		for (size_t i = 0; i < d2->n * s2r->n; ++i) Tp[i] += Tq[i];
	}
	dump("e32_trans_parts", Tp, d2->n * s2r->n * sizeof(complex double));

	/* ---- 3. W of particles with two different bases --------------------- */
	qpms_vswf_set_spec_t *bs2 = qpms_vswf_set_spec_from_lMax(2, QPMS_NORMALISATION_DEFAULT);
	qpms_vswf_set_spec_t *bs3 = qpms_vswf_set_spec_from_lMax(3, QPMS_NORMALISATION_DEFAULT);
	qpms_epsmu_t bg = {1.52 * 1.52, 1}, in1 = {-12 + 1.0 * I, 1}, in2 = {-8 + 0.5 * I, 1};
	qpms_tmatrix_generator_sphere_param_t sp1 = {{qpms_epsmu_const_g, &bg}, {qpms_epsmu_const_g, &in1}, 50e-9};
	qpms_tmatrix_generator_sphere_param_t sp2 = {{qpms_epsmu_const_g, &bg}, {qpms_epsmu_const_g, &in2}, 70e-9};
	qpms_tmatrix_generator_t g1 = {qpms_tmatrix_generator_sphere, &sp1}, g2 = {qpms_tmatrix_generator_sphere, &sp2};
	qpms_tmatrix_function_t tmf[2] = {{bs2, &g1}, {bs3, &g2}};
	qpms_ss_derived_tmatrix_t dtm[2];
	memset(dtm, 0, sizeof dtm);
	dtm[0].tmgi = 0; dtm[1].tmgi = 1;
	const int np = 6;
	qpms_particle_tid_t p[6];
	for (int i = 0; i < np; ++i) {
		p[i].pos.x = (i % 3) * 400e-9 + 13e-9 * i;
		p[i].pos.y = (i / 3) * 450e-9 - 7e-9 * i;
		p[i].pos.z = (i % 2) * 60e-9;
		p[i].tmatrix_id = (i % 2) ? 1 : 0;   /* alternate lMax 2 and lMax 3 */
	}
	qpms_scatsys_t ss0;
	memset(&ss0, 0, sizeof ss0);
	ss0.medium.function = qpms_epsmu_const_g;
	ss0.medium.params = &bg;
	ss0.tmg = tmf; ss0.tmg_count = 2;
	ss0.tm = dtm; ss0.tm_count = 2;
	ss0.p = p; ss0.p_count = np;
	const complex double om = 2.05 * EV_RAD_PER_S;
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_apply_symmetry(&ss0, NULL, om, NULL);
	const qpms_scatsys_t *ss = ssw->ss;
	const size_t L = ss->fecv_size;
	complex double *W = qpms_scatsys_build_translation_matrix_full(NULL, ss, ssw->wavenumber);
	dump("mixed_W", W, L * L * sizeof(complex double));
	double mixin[6 * 4 + 2];
	for (int i = 0; i < np; ++i) {
		mixin[4 * i] = ss->p[i].pos.x; mixin[4 * i + 1] = ss->p[i].pos.y; mixin[4 * i + 2] = ss->p[i].pos.z;
		mixin[4 * i + 3] = ss->tm[ss->p[i].tmatrix_id].tmgi == 0 ? 2 : 3;
	}
	mixin[24] = creal(ssw->wavenumber);
	mixin[25] = cimag(ssw->wavenumber);
	dump("mixed_in", mixin, sizeof mixin);
	free(W);
	qpms_scatsys_at_omega_free(ssw);

	/* ---- 4. finite eigenmodes of a 2 x 2 cluster ------------------------ */
	qpms_vswf_set_spec_t *be = qpms_vswf_set_spec_from_lMax(2, QPMS_NORMALISATION_DEFAULT);
	qpms_epsmu_t inb = {-12 + 1.0 * I, 1};
	qpms_tmatrix_generator_sphere_param_t spb = {{qpms_epsmu_const_g, &bg}, {qpms_epsmu_const_g, &inb}, 70e-9};
	qpms_tmatrix_generator_t gb = {qpms_tmatrix_generator_sphere, &spb};
	qpms_tmatrix_function_t tmb = {be, &gb};
	qpms_ss_derived_tmatrix_t dtb;
	memset(&dtb, 0, sizeof dtb);
	qpms_particle_tid_t pb[4];
	for (int i = 0; i < 4; ++i) {
		pb[i].pos.x = ((i / 2) - 0.5) * 400e-9; pb[i].pos.y = ((i % 2) - 0.5) * 400e-9; pb[i].pos.z = 0;
		pb[i].tmatrix_id = 0;
	}
	qpms_scatsys_t sb0;
	memset(&sb0, 0, sizeof sb0);
	sb0.medium.function = qpms_epsmu_const_g;
	sb0.medium.params = &bg;
	sb0.tmg = &tmb; sb0.tmg_count = 1;
	sb0.tm = &dtb; sb0.tm_count = 1;
	sb0.p = pb; sb0.p_count = 4;
	qpms_scatsys_at_omega_t *sbw = qpms_scatsys_apply_symmetry(&sb0, NULL, 1.12 * EV_RAD_PER_S, NULL);
	const complex double centre = (1.12 - 0.25 * I) * EV_RAD_PER_S;
	/* absolute singular-value cut: 1e-5 of the scale 2 pi r_Re sqrt(m l) of A0 (m = l = 64) */
	const double rank_tol = 1e-5 * 2 * M_PI * 0.1 * EV_RAD_PER_S * 64;
	beyn_result_t *r = qpms_scatsys_finite_find_eigenmodes(sbw->ss, QPMS_NO_IRREP, centre, 0.1 * EV_RAD_PER_S,
		0.05 * EV_RAD_PER_S, 48, rank_tol, 1, 0.0);
	const double nres[2] = {(double)r->neig, (double)r->vlen};
	dump("eig_n", nres, sizeof nres);
	dump("eig_val", r->eigval, r->neig * sizeof(complex double));
	dump("eig_vec", r->eigvec, r->neig * r->vlen * sizeof(complex double));
	double bin[4 * 3];
	for (int i = 0; i < 4; ++i) { bin[3 * i] = pb[i].pos.x; bin[3 * i + 1] = pb[i].pos.y; bin[3 * i + 2] = 0; }
	dump("eig_pos", bin, sizeof bin);
	beyn_result_free(r);
	qpms_scatsys_at_omega_free(sbw);

	/* ---- 5. periodic eigenmodes: the wrapper against beyn_solve itself --- */
	qpms_vswf_set_spec_t *bp = qpms_vswf_set_spec_from_lMax(2, QPMS_NORMALISATION_DEFAULT);
	qpms_tmatrix_generator_sphere_param_t spp = {{qpms_epsmu_const_g, &bg}, {qpms_epsmu_const_g, &inb}, 50e-9};
	qpms_tmatrix_generator_t gp = {qpms_tmatrix_generator_sphere, &spp};
	qpms_tmatrix_function_t tmp = {bp, &gp};
	qpms_ss_derived_tmatrix_t dtp;
	memset(&dtp, 0, sizeof dtp);
	qpms_particle_tid_t pp[2];
	memset(pp, 0, sizeof pp);
	pp[1].pos.x = a / sqrt(3);
	qpms_scatsys_t sp0;
	memset(&sp0, 0, sizeof sp0);
	sp0.lattice_dimension = 2;
	sp0.medium.function = qpms_epsmu_const_g;
	sp0.medium.params = &bg;
	sp0.tmg = &tmp; sp0.tmg_count = 1;
	sp0.tm = &dtp; sp0.tm_count = 1;
	sp0.p = pp; sp0.p_count = 2;
	sp0.per.lattice_basis[0].x = b1.x; sp0.per.lattice_basis[0].y = b1.y;
	sp0.per.lattice_basis[1].x = b2.x; sp0.per.lattice_basis[1].y = b2.y;
	qpms_scatsys_at_omega_t *spw = qpms_scatsys_apply_symmetry(&sp0, NULL, 2.0 * EV_RAD_PER_S, NULL);
	const double kb[3] = {0.1 * 2 * M_PI / a, 0.05 * 2 * M_PI / a, 0};
	const complex double pc = (2.0 - 0.08 * I) * EV_RAD_PER_S;
	const double prr = 0.25 * EV_RAD_PER_S, pri = 0.1 * EV_RAD_PER_S;
	const size_t pm = spw->ss->fecv_size;   /* 32: l = m probe columns */
	const double ptol = 1e-5 * 2 * M_PI * prr * (double)pm;
	beyn_result_t *rw = qpms_scatsys_periodic_find_eigenmodes(spw->ss, kb, pc, prr, pri, 40, ptol, 1, 0.0);
	struct per_param { const qpms_scatsys_t *ss; const double *k; } ppar = {spw->ss, kb};
	beyn_contour_t *pcont = beyn_contour_ellipse(pc, prr, pri, 40);
	beyn_result_t *rd = beyn_solve(pm, pm, periodic_M, NULL, &ppar, pcont, ptol, 1, 0.0);
	const double pn[4] = {(double)rw->neig, (double)rd->neig, (double)pm, ptol};
	dump("per_n", pn, sizeof pn);
	dump("per_wrap_val", rw->eigval, rw->neig * sizeof(complex double));
	dump("per_direct_val", rd->eigval, rd->neig * sizeof(complex double));
	dump("per_wrap_sv", rw->ranktest_SV, pm * sizeof(double));
	dump("per_direct_sv", rd->ranktest_SV, pm * sizeof(double));
	beyn_result_free(rw);
	beyn_result_free(rd);
	free(pcont);

	qpms_trans_calculator_free(c3);
	qpms_trans_calculator_free(c2);
	printf("trans_arrays ok\n");
	return 0;
}
