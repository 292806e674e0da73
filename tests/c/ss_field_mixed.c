/* Scattered field of a cluster whose particles carry different bases, through
 * the public C ABI: two sphere T-matrix generators with bases of lMax 2 and
 * lMax 3 (default order), three particles, qpms_scatsys_scattered_E and
 * qpms_hip_scatsys_scattered_E_points of the coefficient vector
 * c_i = exp(0.37 i I) / (1 + 0.1 i) (fecv order) at the given points.
 *
 * usage: ss_field_mixed OUT POINTS
 * POINTS: float64 x, y, z per point. Writes to OUT: int64 fecv_size, float64
 * wavenumber (re, im), then complex128 x, y, z per point from the single-point
 * call, then from the many-point call. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "qpms_hip.h"

#define EV_RAD_PER_S (1.602176634e-19 / 1.054571817e-34)

int main(int argc, char **argv) {
	if (argc != 3) { fprintf(stderr, "usage: %s OUT POINTS\n", argv[0]); return 2; }
	if (qpms_hip_init() != 0) return 3;
	FILE *fp = fopen(argv[2], "rb");
	if (!fp) { perror(argv[2]); return 4; }
	fseek(fp, 0, SEEK_END);
	const long long npts = ftell(fp) / (long long)sizeof(cart3_t);
	fseek(fp, 0, SEEK_SET);
	cart3_t *w = malloc((size_t)npts * sizeof *w);
	if (fread(w, sizeof *w, (size_t)npts, fp) != (size_t)npts) return 4;
	fclose(fp);

	qpms_vswf_set_spec_t *b2 = qpms_vswf_set_spec_from_lMax(2, QPMS_NORMALISATION_DEFAULT);
	qpms_vswf_set_spec_t *b3 = qpms_vswf_set_spec_from_lMax(3, QPMS_NORMALISATION_DEFAULT);
	qpms_epsmu_t bg = {2.31, 1}, in = {-12 + 1 * I, 1};
	qpms_tmatrix_generator_sphere_param_t sp = {{qpms_epsmu_const_g, &bg}, {qpms_epsmu_const_g, &in}, 50e-9};
	qpms_tmatrix_generator_t gen = {qpms_tmatrix_generator_sphere, &sp};
	qpms_tmatrix_function_t tmf[2] = {{b2, &gen}, {b3, &gen}};
	qpms_ss_derived_tmatrix_t dtm[2];
	memset(dtm, 0, sizeof dtm);
	dtm[0].tmgi = 0; dtm[0].op.typ = QPMS_TMATRIX_OPERATION_NOOP;
	dtm[1].tmgi = 1; dtm[1].op.typ = QPMS_TMATRIX_OPERATION_NOOP;
	qpms_particle_tid_t p[3];
	memset(p, 0, sizeof p);
	p[0].pos.x = -400e-9; p[0].tmatrix_id = 0;
	p[1].pos.y = 350e-9; p[1].pos.z = 40e-9; p[1].tmatrix_id = 1;
	p[2].pos.x = 300e-9; p[2].pos.y = -200e-9; p[2].tmatrix_id = 0;
	qpms_scatsys_t ss0;
	memset(&ss0, 0, sizeof ss0);
	ss0.medium.function = qpms_epsmu_const_g;
	ss0.medium.params = &bg;
	ss0.tmg = tmf; ss0.tmg_count = 2;
	ss0.tm = dtm; ss0.tm_count = 2;
	ss0.p = p; ss0.p_count = 3;
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_apply_symmetry(&ss0, NULL, 2.0 * EV_RAD_PER_S, NULL);
	const qpms_scatsys_t *ss = ssw->ss;
	const size_t L = ss->fecv_size;
	complex double *cv = malloc(L * sizeof *cv);
	for (size_t i = 0; i < L; ++i) cv[i] = cexp(0.37 * I * i) / (1.0 + 0.1 * i);
	ccart3_t *e = malloc((size_t)npts * 2 * sizeof *e);
	for (long long t = 0; t < npts; ++t) e[t] = qpms_scatsys_scattered_E(ss, QPMS_HANKEL_PLUS, ssw->wavenumber, cv, w[t]);
	qpms_hip_scatsys_scattered_E_points(ss, QPMS_HANKEL_PLUS, ssw->wavenumber, cv, (size_t)npts, w, e + npts);
	FILE *fo = fopen(argv[1], "wb");
	if (!fo) { perror(argv[1]); return 4; }
	const long long LL = (long long)L;
	const double kk[2] = {creal(ssw->wavenumber), cimag(ssw->wavenumber)};
	const int ok = fwrite(&LL, sizeof LL, 1, fo) == 1 && fwrite(kk, sizeof kk, 1, fo) == 1
		&& fwrite(e, sizeof *e, (size_t)npts * 2, fo) == (size_t)npts * 2;
	fclose(fo);
	for (qpms_ss_tmi_t i = 0; i < ss->tm_count; ++i) qpms_tmatrix_free(ssw->tm[i]);
	free(ssw->tm);
	qpms_scatsys_free((qpms_scatsys_t *)ss);
	free(ssw);
	qpms_vswf_set_spec_free(b2);
	qpms_vswf_set_spec_free(b3);
	free(cv); free(e); free(w);
	return ok ? 0 : 5;
}
