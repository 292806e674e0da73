/* The plain drop-in path of one serving request: the public scattering-system
 * API called one function after another with host buffers and FP64 LU, as a
 * C caller of the reference library answers it (misc/finiterectlat-scatter.py
 * for plane-wave solves; scatsystem.c:2301-2370 with beyn_solve for
 * eigenfrequencies; the periodic mode-problem builder and zgetrf for
 * determinants). Used as the baseline of the serving loop and as its oracle
 * on the device side. */
#include "qpms_internal.h"
#include "../../include/qpms_hip_serve.h"
#include <complex.h>
#include <math.h>
#include <string.h>

void qh_probe_host(long n, unsigned long long seed, double *V);

typedef struct {
	qpms_vswf_set_spec_t *bspec;
	qpms_epsmu_t bg, in;
	qpms_tmatrix_generator_sphere_param_t sp;
	qpms_tmatrix_generator_t gen;
	qpms_tmatrix_function_t tmf;
	qpms_ss_derived_tmatrix_t dtm;
	qpms_particle_tid_t *p;
	qpms_scatsys_t ss0;
} qh_plain_sys;

static void plain_sys_init(qh_plain_sys *q, const qpms_hip_request_t *r, int periodic) {
	memset(q, 0, sizeof(*q));
	q->bspec = qpms_vswf_set_spec_from_lMax(r->lMax, QPMS_NORMALISATION_DEFAULT);
	q->bg.eps = r->eps_bg; q->bg.mu = 1;
	q->in.eps = r->eps_in; q->in.mu = 1;
	q->sp.outside.function = qpms_epsmu_const_g; q->sp.outside.params = &q->bg;
	q->sp.inside.function = qpms_epsmu_const_g; q->sp.inside.params = &q->in;
	q->sp.radius = r->radius;
	q->gen.function = qpms_tmatrix_generator_sphere; q->gen.params = &q->sp;
	q->tmf.spec = q->bspec; q->tmf.gen = &q->gen;
	q->dtm.tmgi = 0; q->dtm.op.typ = QPMS_TMATRIX_OPERATION_NOOP;
	QH_CALLOC(q->p, r->npart, sizeof(*q->p));
	for (size_t i = 0; i < r->npart; ++i) {
		q->p[i].pos.x = r->pos[3 * i]; q->p[i].pos.y = r->pos[3 * i + 1]; q->p[i].pos.z = r->pos[3 * i + 2];
	}
	q->ss0.lattice_dimension = periodic ? 2 : 0;
	q->ss0.medium.function = qpms_epsmu_const_g; q->ss0.medium.params = &q->bg;
	q->ss0.tmg = &q->tmf; q->ss0.tmg_count = 1;
	q->ss0.tm = &q->dtm; q->ss0.tm_count = 1;
	q->ss0.p = q->p; q->ss0.p_count = (qpms_ss_pi_t)r->npart;
	if (periodic) {
		q->ss0.per.lattice_basis[0].x = r->lattice[0][0]; q->ss0.per.lattice_basis[0].y = r->lattice[0][1];
		q->ss0.per.lattice_basis[1].x = r->lattice[1][0]; q->ss0.per.lattice_basis[1].y = r->lattice[1][1];
	}
}

static void plain_ssw_free(qpms_scatsys_at_omega_t *ssw) {
	qpms_scatsys_t *ss = (qpms_scatsys_t *)ssw->ss;
	for (qpms_ss_tmi_t i = 0; i < ss->tm_count; ++i) qpms_tmatrix_free(ssw->tm[i]);
	free(ssw->tm);
	free(ssw);
	qpms_scatsys_free(ss);
}

static void plain_sys_free(qh_plain_sys *q) {
	qpms_vswf_set_spec_free(q->bspec);
	free(q->p);
}

struct qh_plain_beyn { const qpms_scatsys_t *ss; };
static int plain_beyn_M(complex double *target, size_t m, complex double omega, void *params) {
	const struct qh_plain_beyn *p = params;
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_at_omega(p->ss, omega);
	QH_ENSURE(m == p->ss->fecv_size, "matrix size mismatch");
	qpms_scatsysw_build_modeproblem_matrix_full(target, ssw);
	qpms_scatsys_at_omega_free(ssw);
	return 0;
}

int qpms_hip_request_plain(const qpms_hip_request_t *r, qpms_hip_response_t *out) {
	memset(out, 0, sizeof(*out));
	out->kind = r->kind;
	qh_plain_sys q;
	plain_sys_init(&q, r, r->kind == QPMS_HIP_REQ_PERIODIC_DET);
	qpms_scatsys_at_omega_t *ssw = qpms_scatsys_apply_symmetry(&q.ss0, NULL, r->omega, NULL);
	const qpms_scatsys_t *ss = ssw->ss;
	const size_t L = ss->fecv_size;
	out->dim = L;
	if (r->kind == QPMS_HIP_REQ_SCATTER) {
		qpms_ss_LU lu = qpms_scatsysw_build_modeproblem_matrix_full_LU(NULL, NULL, ssw);
		const double k = creal(ssw->wavenumber);
		qpms_incfield_planewave_params_t pw;
		memset(&pw, 0, sizeof pw);
		pw.use_cartesian = true;
		pw.k.cart.x = k * r->kdir[0]; pw.k.cart.y = k * r->kdir[1]; pw.k.cart.z = k * r->kdir[2];
		pw.E.cart.x = r->E[0]; pw.E.cart.y = r->E[1]; pw.E.cart.z = r->E[2];
		complex double *a = qpms_scatsys_incident_field_vector_full(NULL, ss, qpms_incfield_planewave, &pw, false);
		complex double *Ta = qpms_scatsysw_apply_Tmatrices_full(NULL, a, ssw);
		complex double *f = qpms_scatsys_scatter_solve(NULL, Ta, lu);
		complex double dot = 0;
		for (size_t i = 0; i < L; ++i) dot += conj(a[i]) * f[i];
		out->sigma_ext = -creal(dot) / (k * k);
		free(a); free(Ta); free(f);
		qpms_ss_LU_free(lu);
	} else if (r->kind == QPMS_HIP_REQ_PERIODIC_DET) {
		qpms_scatsys_at_omega_k_t sswk;
		sswk.ssw = ssw;
		sswk.k[0] = r->beta[0]; sswk.k[1] = r->beta[1]; sswk.k[2] = 0;
		sswk.eta = qpms_ss_adjusted_eta(ss, ssw->wavenumber, sswk.k);
		qpms_ss_LU lu = qpms_scatsyswk_build_modeproblem_matrix_full_LU(NULL, NULL, &sswk);
		double la = 0, ph = 0;
		for (size_t i = 0; i < L; ++i) {
			const complex double u = lu.a[i * L + i];
			la += log(cabs(u));
			ph += carg(u);
			if (lu.ipiv[i] != (int)i + 1) ph += M_PI;
		}
		ph = fmod(ph, 2 * M_PI);
		if (ph > M_PI) ph -= 2 * M_PI;
		if (ph <= -M_PI) ph += 2 * M_PI;
		out->logabsdet = la;
		out->argdet = ph;
		qpms_ss_LU_free(lu);
	} else {
		const size_t l = (size_t)(r->beyn_l > 0 ? r->beyn_l : 32) > L ? L : (size_t)(r->beyn_l > 0 ? r->beyn_l : 32);
		const int N = r->beyn_nodes > 0 ? r->beyn_nodes : 32;
		const double rank_tol = r->rank_tol > 0 ? r->rank_tol : 1e-4;
		complex double *V;
		QH_MALLOC(V, L * l * sizeof(complex double));
		qh_probe_host((long)(L * l), r->beyn_seed, (double *)V);
		beyn_contour_t *c = beyn_contour_ellipse(CMPLX(r->omega, r->beyn_centre_im), r->beyn_rRe, r->beyn_rIm,
			(size_t)N);
		struct qh_plain_beyn p = {ss};
		beyn_result_t *br = beyn_solve_v(L, l, plain_beyn_M, NULL, &p, c, rank_tol, 0, 0.0, V);
		out->neig = (int)br->neig;
		for (size_t i = 0; i < br->neig && i < QPMS_HIP_SERVE_MAX_EIG; ++i) out->eigval[i] = br->eigval[i];
		for (size_t i = 0; i < l && i < QPMS_HIP_SERVE_MAX_EIG; ++i) out->sv[i] = br->ranktest_SV[i];
		beyn_result_free(br);
		free(c);
		free(V);
	}
	plain_ssw_free(ssw);
	plain_sys_free(&q);
	return 0;
}
