/* Scattering system: construction (trivial symmetry group), T-matrices at ω,
 * translation / mode-problem matrix builders, LU and solve.
 *
 * Reference: scatsystem.c:233-563 (construction), 606-640 (at ω),
 * 1155-1357 (builders), 2197-2273 (LU / solve); tmatrices.c:39-69, 500-591
 * (T-matrices, Mie); vswf.c:76-97 (bspec); materials.h:34-48.
 *
 * The builders run on the GPU: positions and T-matrix diagonals go to HBM,
 * k_w_build writes the matrix, the result is copied to the caller's host
 * buffer. Dense LU / solve run through the k_linalg kernels.
 */
#include "qpms_internal.h"
#include <complex.h>
#include <math.h>
#include <string.h>

#define QH_LEN_RTOL 1e-13             /* scatsystem.c:35 */
#define QH_TM_ATOL 1e-12              /* scatsystem.c:36 */
#define QH_TM_RTOL 1e-12              /* scatsystem.c:37 */
#define QH_EWALD_MAX_EXPONENT 16.     /* scatsystem.c:42 */

int qh_reciprocalBasis1(cart2_t b1, cart2_t b2, cart2_t *rb1, cart2_t *rb2);
void qh_trans_array_from_AB(_Complex double *t, const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		const _Complex double *A, const _Complex double *B, int nelem_AB);

/* ---------------------------------------------------------------------- */
/* bspec, T-matrix, materials                                               */
/* ---------------------------------------------------------------------- */
qpms_epsmu_t qpms_epsmu_const_g(complex double omega, const void *epsmu) {
	(void)omega;
	return *(const qpms_epsmu_t *)epsmu;
}

static inline complex double qh_wavenumber(complex double omega, qpms_epsmu_t em) {
	return csqrt(em.eps * em.mu) * omega / SPEED_OF_LIGHT;   /* materials.h:40-48 */
}

qpms_vswf_set_spec_t *qpms_vswf_set_spec_from_lMax(qpms_l_t lMax, qpms_normalisation_t norm) {
	qpms_vswf_set_spec_t *c;
	QH_MALLOC(c, sizeof(*c));
	c->n = c->capacity = 2 * qpms_lMax2nelem(lMax);
	QH_MALLOC(c->ilist, sizeof(qpms_uvswfi_t) * c->capacity);
	size_t i = 0;
	for (int it = 0; it < 2; ++it)
		for (int n = 1; n <= lMax; ++n)
			for (int m = -n; m <= n; ++m)
				c->ilist[i++] = qpms_tmn2uvswfi(it ? QPMS_VSWF_MAGNETIC : QPMS_VSWF_ELECTRIC, m, n);
	c->norm = norm;
	c->lMax = c->lMax_M = c->lMax_N = lMax;
	c->lMax_L = -1;
	return c;
}

void qpms_vswf_set_spec_free(qpms_vswf_set_spec_t *s) {
	if (s) free(s->ilist);
	free(s);
}

qpms_tmatrix_t *qpms_tmatrix_init(const qpms_vswf_set_spec_t *bspec) {
	qpms_tmatrix_t *t;
	QH_MALLOC(t, sizeof(*t));
	t->spec = bspec;
	QH_CALLOC(t->m, bspec->n * bspec->n, sizeof(complex double));
	t->owns_m = true;
	return t;
}

void qpms_tmatrix_free(qpms_tmatrix_t *t) {
	if (t && t->owns_m) free(t->m);
	free(t);
}

/* Mie coefficients on the GPU (k_mie). */
complex double *qpms_mie_coefficients_reflection(complex double *target, const qpms_vswf_set_spec_t *bspec,
		double a, complex double k_i, complex double k_e, complex double mu_i, complex double mu_e,
		qpms_bessel_t J_ext, qpms_bessel_t J_scat) {
	QH_ENSURE(J_ext != J_scat, "J_ext and J_scat must not be equal. Perhaps you want J_ext = 1 and J_scat = 3 anyways.");
	QH_ENSURE(J_ext == QPMS_BESSEL_REGULAR && J_scat == QPMS_HANKEL_PLUS,
		"only J_ext = REGULAR, J_scat = HANKEL_PLUS are implemented");
	if (!target) QH_MALLOC(target, bspec->n * sizeof(complex double));
	const int lMax = bspec->lMax;
	double prm[9] = {a, creal(k_i), cimag(k_i), creal(k_e), cimag(k_e),
		creal(mu_i), cimag(mu_i), creal(mu_e), cimag(mu_e)};
	const size_t outb = 2 * (size_t)(lMax + 1) * sizeof(complex double);
	char *d = qh_dmalloc(sizeof prm + outb);
	qh_h2d(d, prm, sizeof prm, NULL);
	qh_launch_mie(lMax, 1, (const double *)d, (double *)(d + sizeof prm), NULL);
	complex double *R;
	QH_MALLOC(R, outb);
	qh_d2h(R, d + sizeof prm, outb, NULL);
	qh_sync(NULL);
	qh_dfree(d);
	for (size_t i = 0; i < bspec->n; ++i) {
		qpms_uvswfi_t u = bspec->ilist[i];
		int t = (int)(u & 3);
		long ysc = (long)(u / 4);
		int l = (int)sqrt((double)ysc);
		QH_ENSURE(l <= lMax && (t == QPMS_VSWF_ELECTRIC || t == QPMS_VSWF_MAGNETIC), "invalid bspec entry");
		target[i] = (t == QPMS_VSWF_ELECTRIC) ? R[l] : R[lMax + 1 + l];
	}
	free(R);
	return target;
}

qpms_errno_t qpms_tmatrix_generator_sphere(qpms_tmatrix_t *t, complex double omega, const void *params) {
	const qpms_tmatrix_generator_sphere_param_t *p = params;
	qpms_epsmu_t out = p->outside.function(omega, p->outside.params);
	qpms_epsmu_t in = p->inside.function(omega, p->inside.params);
	complex double *mc = qpms_mie_coefficients_reflection(NULL, t->spec, p->radius,
		qh_wavenumber(omega, in), qh_wavenumber(omega, out), in.mu, out.mu,
		QPMS_BESSEL_REGULAR, QPMS_HANKEL_PLUS);
	const size_t n = t->spec->n;
	memset(t->m, 0, n * n * sizeof(complex double));
	for (size_t i = 0; i < n; ++i) t->m[n * i + i] = mc[i];
	free(mc);
	return QPMS_SUCCESS;
}

/* ---------------------------------------------------------------------- */
/* Construction                                                             */
/* ---------------------------------------------------------------------- */
static qpms_tmatrix_t *qh_tmatrix_from_function(qpms_tmatrix_function_t f, complex double omega) {
	qpms_tmatrix_t *t = qpms_tmatrix_init(f.spec);
	QH_ENSURE(f.gen->function(t, omega, f.gen->params) == QPMS_SUCCESS, "T-matrix generator failed");
	return t;
}

static int qh_tmatrix_isclose(const qpms_tmatrix_t *a, const qpms_tmatrix_t *b) {
	if (a->spec != b->spec && a->spec->n != b->spec->n) return 0;
	if (a->spec->n != b->spec->n) return 0;
	for (size_t i = 0; i < a->spec->n; ++i) if (a->spec->ilist[i] != b->spec->ilist[i]) return 0;
	for (size_t i = 0; i < a->spec->n * a->spec->n; ++i)
		if (cabs(a->m[i] - b->m[i]) > QH_TM_ATOL + QH_TM_RTOL * 0.5 * (cabs(a->m[i]) + cabs(b->m[i])))
			return 0;
	return 1;
}

static void qh_process_lattice_2d(qpms_scatsys_t *ss) {
	/* scatsystem.c:189-218 */
	cart3_t *lb = ss->per.lattice_basis;
	double n0 = sqrt(lb[0].x * lb[0].x + lb[0].y * lb[0].y + lb[0].z * lb[0].z);
	cart3_t on0 = {lb[0].x / n0, lb[0].y / n0, lb[0].z / n0};
	double d01 = on0.x * lb[1].x + on0.y * lb[1].y + on0.z * lb[1].z;
	cart3_t on1 = {lb[1].x - d01 * on0.x, lb[1].y - d01 * on0.y, lb[1].z - d01 * on0.z};
	double n1 = sqrt(on1.x * on1.x + on1.y * on1.y + on1.z * on1.z);
	on1.x /= n1; on1.y /= n1; on1.z /= n1;
	cart2_t b2d[2] = {{n0, 0}, {d01, on1.x * lb[1].x + on1.y * lb[1].y + on1.z * lb[1].z}};
	l2d_reduceBasis(b2d[0], b2d[1], &b2d[0], &b2d[1]);
	ss->per.unitcell_volume = l2d_unitcell_area(b2d[0], b2d[1]);
	ss->per.eta = 2. / M_2_SQRTPI / sqrt(ss->per.unitcell_volume);
	cart2_t r2d[2];
	QH_ENSURE(qh_reciprocalBasis1(b2d[0], b2d[1], &r2d[0], &r2d[1]) == QPMS_SUCCESS, "degenerate lattice basis");
	for (int i = 0; i < 2; ++i) {
		lb[i].x = b2d[i].x * on0.x + b2d[i].y * on1.x;
		lb[i].y = b2d[i].x * on0.y + b2d[i].y * on1.y;
		lb[i].z = b2d[i].x * on0.z + b2d[i].y * on1.z;
		ss->per.reciprocal_basis[i].x = r2d[i].x * on0.x + r2d[i].y * on1.x;
		ss->per.reciprocal_basis[i].y = r2d[i].x * on0.y + r2d[i].y * on1.y;
		ss->per.reciprocal_basis[i].z = r2d[i].x * on0.z + r2d[i].y * on1.z;
	}
}

qpms_scatsys_at_omega_t *qpms_scatsys_apply_symmetry(const qpms_scatsys_t *orig,
		const struct qpms_finite_group_t *sym, complex double omega,
		const struct qpms_tolerance_spec_t *tol) {
	(void)tol;
	if (sym) {
		/* qpms_finite_group_t begins with {char *name; int order; ...} (groups.h:53-56) */
		const struct { char *name; int order; } *g = (const void *)sym;
		QH_ENSURE(g->order == 1, "only the trivial symmetry group is supported (got order %d)", g->order);
	}
	QH_ENSURE(orig->lattice_dimension == 0 || orig->lattice_dimension == 2,
		"only finite and 2D-periodic systems are supported");
	double lenscale;
	{
		double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
		for (qpms_ss_pi_t i = 0; i < orig->p_count; ++i) {
			const double v[3] = {orig->p[i].pos.x, orig->p[i].pos.y, orig->p[i].pos.z};
			for (int a = 0; a < 3; ++a) { mn[a] = fmin(mn[a], v[a]); mx[a] = fmax(mx[a], v[a]); }
		}
		lenscale = (fabs(mx[0]) + fabs(mx[1]) + fabs(mx[2]) + (mx[0] - mn[0]) + (mx[1] - mn[1]) + (mx[2] - mn[2])) / 3;
	}
	qpms_scatsys_t *ss;
	QH_CALLOC(ss, 1, sizeof(*ss));
	ss->lattice_dimension = orig->lattice_dimension;
	if (ss->lattice_dimension) {
		ss->per = orig->per;
		qh_process_lattice_2d(ss);
		for (int i = 0; i < 2; ++i) {
			cart3_t b = ss->per.lattice_basis[i];
			lenscale = fmax(lenscale, sqrt(b.x * b.x + b.y * b.y + b.z * b.z));
		}
	}
	ss->lenscale = lenscale;
	ss->sym = sym;
	ss->medium = orig->medium;
	ss->tmg_count = orig->tmg_count;
	QH_MALLOC(ss->tmg, ss->tmg_count * sizeof(*ss->tmg));
	memcpy(ss->tmg, orig->tmg, ss->tmg_count * sizeof(*ss->tmg));

	/* T-matrices at ω, duplicates (by value) merged as the reference does. */
	qpms_tmatrix_t **tmg_at;
	QH_MALLOC(tmg_at, ss->tmg_count * sizeof(*tmg_at));
	for (qpms_ss_tmgi_t i = 0; i < ss->tmg_count; ++i) tmg_at[i] = qh_tmatrix_from_function(orig->tmg[i], omega);
	ss->tm_capacity = orig->tm_count;
	QH_MALLOC(ss->tm, ss->tm_capacity * sizeof(*ss->tm));
	qpms_scatsys_at_omega_t *ssw;
	QH_MALLOC(ssw, sizeof(*ssw));
	ssw->ss = ss;
	ssw->omega = omega;
	ssw->medium = ss->medium.function(omega, ss->medium.params);
	ssw->wavenumber = qh_wavenumber(omega, ssw->medium);
	QH_MALLOC(ssw->tm, ss->tm_capacity * sizeof(*ssw->tm));
	qpms_ss_tmi_t *remap;
	QH_MALLOC(remap, (orig->tm_count + 1) * sizeof(*remap));
	ss->tm_count = 0;
	ss->max_bspecn = 0;
	for (qpms_ss_tmi_t i = 0; i < orig->tm_count; ++i) {
		QH_ENSURE(orig->tm[i].op.typ == QPMS_TMATRIX_OPERATION_NOOP, "only NOOP T-matrix operations are supported");
		const qpms_tmatrix_t *src = tmg_at[orig->tm[i].tmgi];
		qpms_tmatrix_t *ti = qpms_tmatrix_init(src->spec);
		memcpy(ti->m, src->m, src->spec->n * src->spec->n * sizeof(complex double));
		qpms_ss_tmi_t j;
		for (j = 0; j < ss->tm_count; ++j) if (qh_tmatrix_isclose(ti, ssw->tm[j])) break;
		if (j == ss->tm_count) {
			ss->tm[j] = orig->tm[i];
			ssw->tm[j] = ti;
			if (ti->spec->n > ss->max_bspecn) ss->max_bspecn = ti->spec->n;
			++ss->tm_count;
		} else qpms_tmatrix_free(ti);
		remap[i] = j;
	}
	for (qpms_ss_tmgi_t i = 0; i < ss->tmg_count; ++i) qpms_tmatrix_free(tmg_at[i]);
	free(tmg_at);

	ss->p_count = ss->p_capacity = orig->p_count;
	QH_MALLOC(ss->p, ss->p_count * sizeof(*ss->p));
	for (qpms_ss_pi_t i = 0; i < ss->p_count; ++i) {
		ss->p[i] = orig->p[i];
		ss->p[i].tmatrix_id = remap[orig->p[i].tmatrix_id];
	}
	free(remap);
	/* Trivial group: one irrep, every particle its own orbit. */
	QH_MALLOC(ss->fecv_pstarts, ss->p_count * sizeof(size_t));
	ss->fecv_size = 0;
	qpms_l_t lMax = 0;
	qpms_normalisation_t norm = QPMS_NORMALISATION_UNDEF;
	for (qpms_ss_pi_t i = 0; i < ss->p_count; ++i) {
		const qpms_vswf_set_spec_t *b = ssw->tm[ss->p[i].tmatrix_id]->spec;
		ss->fecv_pstarts[i] = ss->fecv_size;
		ss->fecv_size += b->n;
		if (b->lMax > lMax) lMax = b->lMax;
		if (norm == QPMS_NORMALISATION_UNDEF) norm = b->norm;
		QH_ENSURE(norm == b->norm, "all bspecs must share one normalisation");
	}
	QH_MALLOC(ss->saecv_sizes, sizeof(size_t));
	ss->saecv_sizes[0] = ss->fecv_size;
	ss->c = qpms_trans_calculator_init(lMax, norm);
	return ssw;
}

void qpms_scatsys_free(qpms_scatsys_t *ss) {
	if (ss) {
		free(ss->tm);
		free(ss->tmg);
		free(ss->p);
		free(ss->fecv_pstarts);
		free(ss->saecv_sizes);
		qpms_trans_calculator_free(ss->c);
	}
	free(ss);
}

qpms_scatsys_at_omega_t *qpms_scatsys_at_omega(const qpms_scatsys_t *ss, complex double omega) {
	qpms_scatsys_at_omega_t *ssw;
	QH_MALLOC(ssw, sizeof(*ssw));
	ssw->omega = omega;
	ssw->ss = ss;
	ssw->medium = ss->medium.function(omega, ss->medium.params);
	ssw->wavenumber = qh_wavenumber(omega, ssw->medium);
	QH_CALLOC(ssw->tm, ss->tm_count, sizeof(*ssw->tm));
	for (qpms_ss_tmi_t i = 0; i < ss->tm_count; ++i)
		ssw->tm[i] = qh_tmatrix_from_function(ss->tmg[ss->tm[i].tmgi], omega);
	return ssw;
}

void qpms_scatsys_at_omega_free(qpms_scatsys_at_omega_t *ssw) {
	if (ssw && ssw->tm)
		for (qpms_ss_tmi_t i = 0; i < ssw->ss->tm_count; ++i) qpms_tmatrix_free(ssw->tm[i]);
	if (ssw) free(ssw->tm);
	free(ssw);
}

double qpms_ss_adjusted_eta(const qpms_scatsys_t *ss, complex double wavenumber, const double k[3]) {
	QH_ENSURE(ss->lattice_dimension > 0, "This method is applicable only to periodic systems.");
	const double eta_default = ss->per.eta;
	const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
	const double kappa2 = cabs(wavenumber) * cabs(wavenumber);
	if (kappa2 < k2) return eta_default;
	const int maxj = ss->c->lMax;
	const double mina = 0.5 * (ss->lattice_dimension - 1) - maxj;
	const double eta_min = sqrt(fabs((kappa2 - k2) * (mina - 1.) / QH_EWALD_MAX_EXPONENT));
	return fmax(eta_default, eta_min);
}

/* ---------------------------------------------------------------------- */
/* Builders                                                                 */
/* ---------------------------------------------------------------------- */
static const qpms_vswf_set_spec_t *qh_bspec_pi(const qpms_scatsys_t *ss, qpms_ss_pi_t pi) {
	return ss->tmg[ss->tm[ss->p[pi].tmatrix_id].tmgi].spec;   /* scatsystem.h:236-238 */
}

/* True if every particle uses the default bspec of the calculator's lMax. */
static int qh_uniform_default_bspec(const qpms_scatsys_t *ss) {
	const int lMax = ss->c->lMax, ne = (int)ss->c->nelem;
	for (qpms_ss_pi_t pi = 0; pi < ss->p_count; ++pi) {
		const qpms_vswf_set_spec_t *b = qh_bspec_pi(ss, pi);
		if (b->lMax != lMax || b->n != (size_t)(2 * ne) || b->lMax_L >= 0) return 0;
		size_t i = 0;
		for (int it = 0; it < 2; ++it)
			for (int n = 1; n <= lMax; ++n)
				for (int m = -n; m <= n; ++m)
					if (b->ilist[i++] != qpms_tmn2uvswfi(it ? QPMS_VSWF_MAGNETIC : QPMS_VSWF_ELECTRIC, m, n)) return 0;
	}
	return 1;
}

static double *qh_upload_positions(const qpms_scatsys_t *ss) {
	const size_t np = ss->p_count;
	double *h;
	QH_MALLOC(h, 3 * np * sizeof(double));
	for (size_t i = 0; i < np; ++i) { h[3 * i] = ss->p[i].pos.x; h[3 * i + 1] = ss->p[i].pos.y; h[3 * i + 2] = ss->p[i].pos.z; }
	double *d = qh_dmalloc(3 * np * sizeof(double));
	qh_h2d(d, h, 3 * np * sizeof(double), NULL);
	qh_sync(NULL);
	free(h);
	return d;
}

/* General bspecs: A/B of every ordered pair on the GPU, block scatter on the host. */
static void qh_build_W_general(complex double *target, const qpms_scatsys_t *ss, complex double k, qpms_bessel_t J) {
	const size_t np = ss->p_count, L = ss->fecv_size, ne = ss->c->nelem;
	const size_t npairs = np * np;
	complex double *kr;
	double *th, *ph;
	QH_MALLOC(kr, npairs * sizeof(*kr));
	QH_MALLOC(th, npairs * sizeof(double));
	QH_MALLOC(ph, npairs * sizeof(double));
	for (size_t r = 0; r < np; ++r)
		for (size_t c = 0; c < np; ++c) {
			cart3_t a = ss->p[r].pos, b = ss->p[c].pos;
			double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z, d = sqrt(dx * dx + dy * dy + dz * dz);
			size_t i = r * np + c;
			if (r == c) d = 1, dz = 0, dx = 1, dy = 0;
			kr[i] = d * k; th[i] = d ? acos(dz / d) : M_PI_2; ph[i] = atan2(dy, dx);
		}
	const size_t abbytes = npairs * ne * ne * sizeof(complex double);
	char *d = qh_dmalloc(npairs * (16 + 16) + 2 * abbytes);
	complex double *dkr = (complex double *)d;
	double *dth = (double *)(d + 16 * npairs), *dph = (double *)(d + 24 * npairs);
	complex double *dA = (complex double *)(d + 32 * npairs), *dB = (complex double *)(d + 32 * npairs + abbytes);
	qh_h2d(dkr, kr, npairs * 16, NULL);
	qh_h2d(dth, th, npairs * 8, NULL);
	qh_h2d(dph, ph, npairs * 8, NULL);
	qpms_trans_calculator_get_AB_arrays_batch_dev(ss->c, npairs, dkr, dth, dph, 0, J, dA, dB, NULL);
	complex double *hA;
	QH_MALLOC(hA, 2 * abbytes);
	qh_d2h(hA, dA, 2 * abbytes, NULL);
	qh_sync(NULL);
	qh_dfree(d);
	memset(target, 0, L * L * sizeof(complex double));
	for (size_t r = 0; r < np; ++r)
		for (size_t c = 0; c < np; ++c) {
			if (r == c) continue;
			const size_t i = r * np + c;
			qh_trans_array_from_AB(target + ss->fecv_pstarts[r] * L + ss->fecv_pstarts[c],
				qh_bspec_pi(ss, r), L, qh_bspec_pi(ss, c), 1,
				hA + i * ne * ne, hA + npairs * ne * ne + i * ne * ne, (int)ne);
		}
	free(hA); free(kr); free(th); free(ph);
}

complex double *qpms_scatsys_build_translation_matrix_e_full(complex double *target,
		const qpms_scatsys_t *ss, complex double k, qpms_bessel_t J) {
	QH_ENSURE(ss->lattice_dimension == 0, "This method is applicable only to nonperiodic systems.");
	const size_t L = ss->fecv_size;
	if (!target) QH_MALLOC(target, L * L * sizeof(complex double));
	if (qh_uniform_default_bspec(ss)) {
		double *dpos = qh_upload_positions(ss);
		complex double *dW = qh_dmalloc(L * L * sizeof(complex double));
		qh_launch_w_build(qh_tables_get(ss->c), ss->p_count, dpos, creal(k), cimag(k), (int)J,
			0, ss->p_count, (double *)dW, L, NULL);
		qh_d2h(target, dW, L * L * sizeof(complex double), NULL);
		qh_sync(NULL);
		qh_dfree(dW);
		qh_dfree(dpos);
	} else {
		qh_build_W_general(target, ss, k, J);
	}
	return target;
}

complex double *qpms_scatsys_build_translation_matrix_full(complex double *target,
		const qpms_scatsys_t *ss, complex double k) {
	return qpms_scatsys_build_translation_matrix_e_full(target, ss, k, QPMS_HANKEL_PLUS);
}

/* Extension: W of a cluster of default-bspec particles, all on the device. */
int qpms_hip_build_translation_matrix_dev(const qpms_trans_calculator *c, size_t npart,
		const cart3_t *pos_dev, complex double k, qpms_bessel_t J, complex double *W_dev, void *stream) {
	const size_t L = npart * 2 * c->nelem;
	return qh_launch_w_build(qh_tables_get(c), npart, (const double *)pos_dev, creal(k), cimag(k), (int)J,
		0, npart, (double *)W_dev, L, stream);
}

int qpms_hip_build_translation_matrix_rows_dev(const qpms_trans_calculator *c, size_t npart,
		const cart3_t *pos_dev, complex double k, qpms_bessel_t J,
		size_t row_p0, size_t row_np, complex double *W_dev, size_t ld, void *stream) {
	return qh_launch_w_build(qh_tables_get(c), npart, (const double *)pos_dev, creal(k), cimag(k), (int)J,
		row_p0, row_np, (double *)W_dev, ld, stream);
}

int qpms_hip_build_modeproblem_matrix_dev(const qpms_trans_calculator *c, size_t npart,
		const cart3_t *pos_dev, const int *ttype_dev, const complex double *tdiag_dev,
		complex double k, complex double *M_dev, void *stream) {
	const size_t L = npart * 2 * c->nelem;
	return qh_launch_modeproblem(qh_tables_get(c), npart, (const double *)pos_dev, ttype_dev,
		(const double *)tdiag_dev, creal(k), cimag(k), (double *)M_dev, L, stream);
}

/* True if every T-matrix at ω is diagonal. */
static int qh_all_T_diagonal(const qpms_scatsys_at_omega_t *ssw) {
	for (qpms_ss_tmi_t t = 0; t < ssw->ss->tm_count; ++t) {
		const size_t n = ssw->tm[t]->spec->n;
		for (size_t i = 0; i < n; ++i)
			for (size_t j = 0; j < n; ++j)
				if (i != j && ssw->tm[t]->m[i * n + j] != 0) return 0;
	}
	return 1;
}

/* Finite M = I − T_R S(R←C), S from the reference's HANKEL_PLUS translation
 * (scatsystem.c:1301-1339 with k == NULL; the reference itself aborts in
 * qpms_ss_ensure_periodic at :1293, so this is the loop body's semantics). */
complex double *qpms_scatsysw_build_modeproblem_matrix_full(complex double *target,
		const qpms_scatsys_at_omega_t *ssw) {
	const qpms_scatsys_t *ss = ssw->ss;
	QH_ENSURE(ss->lattice_dimension == 0,
		"This method is applicable only to nonperiodic systems. Use qpms_scatsyswk_build_modeproblem_matrix_full() instead.");
	const size_t L = ss->fecv_size, np = ss->p_count;
	if (!target) QH_MALLOC(target, L * L * sizeof(complex double));
	if (qh_uniform_default_bspec(ss) && qh_all_T_diagonal(ssw)) {
		const size_t nb = 2 * ss->c->nelem;
		complex double *tdiag;
		int *ttype;
		QH_MALLOC(tdiag, ss->tm_count * nb * sizeof(complex double));
		QH_MALLOC(ttype, np * sizeof(int));
		for (qpms_ss_tmi_t t = 0; t < ss->tm_count; ++t)
			for (size_t i = 0; i < nb; ++i) tdiag[t * nb + i] = ssw->tm[t]->m[i * nb + i];
		for (size_t i = 0; i < np; ++i) ttype[i] = ss->p[i].tmatrix_id;
		double *dpos = qh_upload_positions(ss);
		complex double *dT = qh_dmalloc(ss->tm_count * nb * sizeof(complex double));
		int *dty = qh_dmalloc(np * sizeof(int));
		complex double *dM = qh_dmalloc(L * L * sizeof(complex double));
		qh_h2d(dT, tdiag, ss->tm_count * nb * sizeof(complex double), NULL);
		qh_h2d(dty, ttype, np * sizeof(int), NULL);
		qh_launch_modeproblem(qh_tables_get(ss->c), np, dpos, dty, (const double *)dT,
			creal(ssw->wavenumber), cimag(ssw->wavenumber), (double *)dM, L, NULL);
		qh_d2h(target, dM, L * L * sizeof(complex double), NULL);
		qh_sync(NULL);
		qh_dfree(dM); qh_dfree(dT); qh_dfree(dty); qh_dfree(dpos);
		free(tdiag); free(ttype);
	} else {
		/* General T: W on the GPU, then M = I − T W per block row on the GPU. */
		complex double *W = qpms_scatsys_build_translation_matrix_e_full(NULL, ss, ssw->wavenumber, QPMS_HANKEL_PLUS);
		complex double *dW = qh_dmalloc(L * L * sizeof(complex double));
		complex double *dM = qh_dmalloc(L * L * sizeof(complex double));
		qh_h2d(dW, W, L * L * sizeof(complex double), NULL);
		for (size_t r = 0; r < np; ++r) {
			const qpms_tmatrix_t *T = ssw->tm[ss->p[r].tmatrix_id];
			const size_t nr = T->spec->n;
			complex double *dT = qh_dmalloc(nr * nr * sizeof(complex double));
			qh_h2d(dT, T->m, nr * nr * sizeof(complex double), NULL);
			qh_launch_block_tw(nr, L, (const double *)dT, (const double *)(dW + ss->fecv_pstarts[r] * L),
				(double *)(dM + ss->fecv_pstarts[r] * L), ss->fecv_pstarts[r], NULL);
			qh_sync(NULL);
			qh_dfree(dT);
		}
		qh_d2h(target, dM, L * L * sizeof(complex double), NULL);
		qh_sync(NULL);
		qh_dfree(dW); qh_dfree(dM);
		free(W);
	}
	return target;
}

complex double *qpms_scatsys_build_modeproblem_matrix_full(complex double *target,
		const qpms_scatsys_at_omega_t *ssw) {
	return qpms_scatsysw_build_modeproblem_matrix_full(target, ssw);
}

/* ---------------------------------------------------------------------- */
/* LU / solve (scatsystem.c:2197-2273)                                      */
/* ---------------------------------------------------------------------- */
void qpms_ss_LU_free(qpms_ss_LU lu) {
	free(lu.a);
	free(lu.ipiv);
}

qpms_ss_LU qpms_scatsysw_modeproblem_matrix_full_factorise(complex double *mp, int *target_piv,
		const qpms_scatsys_at_omega_t *ssw, const qpms_scatsys_at_omega_k_t *sswk) {
	if (sswk) {
		QH_ENSURE(sswk->ssw == ssw || !ssw, "inconsistent ssw / sswk");
		ssw = sswk->ssw;
		QH_ENSURE(ssw->ss->lattice_dimension > 0, "periodic sswk given for a finite system");
	} else {
		QH_ENSURE(ssw->ss->lattice_dimension == 0, "finite ssw given for a periodic system");
	}
	QH_ENSURE(mp, "A non-NULL pointer to the pre-calculated mode matrix is required");
	const size_t n = ssw->ss->fecv_size;
	if (!target_piv) QH_MALLOC(target_piv, n * sizeof(int));
	complex double *dA = qh_dmalloc(n * n * sizeof(complex double));
	int *dp = qh_dmalloc(n * sizeof(int));
	qh_h2d(dA, mp, n * n * sizeof(complex double), NULL);
	int info = qpms_hip_zgetrf_dev(n, dA, dp, NULL);
	qh_d2h(mp, dA, n * n * sizeof(complex double), NULL);
	qh_d2h(target_piv, dp, n * sizeof(int), NULL);
	qh_sync(NULL);
	qh_dfree(dA); qh_dfree(dp);
	QH_ENSURE(info >= 0, "zgetrf failed (info %d)", info);
	qpms_ss_LU lu;
	lu.a = mp; lu.ipiv = target_piv; lu.ssw = ssw; lu.sswk = sswk; lu.full = true; lu.iri = -1;
	return lu;
}

qpms_ss_LU qpms_scatsysw_build_modeproblem_matrix_full_LU(complex double *target, int *target_piv,
		const qpms_scatsys_at_omega_t *ssw) {
	target = qpms_scatsysw_build_modeproblem_matrix_full(target, ssw);
	return qpms_scatsysw_modeproblem_matrix_full_factorise(target, target_piv, ssw, NULL);
}

complex double *qpms_scatsys_scatter_solve(complex double *f, const complex double *a_inc, qpms_ss_LU lu) {
	QH_ENSURE(lu.full, "irrep-packed LU is not supported");
	const size_t n = lu.ssw->ss->fecv_size;
	if (!f) QH_MALLOC(f, n * sizeof(complex double));
	complex double *dA = qh_dmalloc(n * n * sizeof(complex double) + n * sizeof(complex double));
	complex double *db = dA + n * n;
	int *dp = qh_dmalloc(n * sizeof(int));
	qh_h2d(dA, lu.a, n * n * sizeof(complex double), NULL);
	qh_h2d(db, a_inc, n * sizeof(complex double), NULL);
	qh_h2d(dp, lu.ipiv, n * sizeof(int), NULL);
	qpms_hip_zgetrs_dev(n, 1, dA, dp, db, NULL);
	qh_d2h(f, db, n * sizeof(complex double), NULL);
	qh_sync(NULL);
	qh_dfree(dA); qh_dfree(dp);
	return f;
}

/* ---------------------------------------------------------------------- */
/* Incident field and T application (vswf.c:480-530, scatsystem.c:1970-2011) */
/* ---------------------------------------------------------------------- */
int qh_launch_planewave(int lMax, size_t npart, const double *pos_dev, const double kvec[3],
		const double *E_dev, double *out_dev, long ld, void *stream);
int qh_launch_apply_t(size_t npart, int nb, const int *ttype_dev, const double *T_dev, const double *a_dev,
		double *out_dev, void *stream);

/* Real k and cartesian E of the parameter structure (vswf.c:483-496). */
static void qh_planewave_kE(const qpms_incfield_planewave_params_t *p, double kv[3], double E[6]) {
	complex double kx, ky, kz, ex, ey, ez;
	if (p->use_cartesian) {
		kx = p->k.cart.x; ky = p->k.cart.y; kz = p->k.cart.z;
		ex = p->E.cart.x; ey = p->E.cart.y; ez = p->E.cart.z;
	} else {
		const double st = sin(p->k.sph.theta), ct = cos(p->k.sph.theta);
		const double sf = sin(p->k.sph.phi), cf = cos(p->k.sph.phi);
		const complex double r = p->k.sph.r;
		kx = r * st * cf; ky = r * st * sf; kz = r * ct;
		/* csphvec2ccart at the wave direction (vectors.h:283-298) */
		const complex double er = p->E.sph.rc, et = p->E.sph.thetac, ef = p->E.sph.phic;
		ex = st * cf * er + ct * cf * et - sf * ef;
		ey = st * sf * er + ct * sf * et + cf * ef;
		ez = ct * er - st * et;
	}
	QH_ENSURE(cimag(kx) == 0 && cimag(ky) == 0 && cimag(kz) == 0,
		"Complex-valued wave vector not implemented correctly; cf. docs.");
	kv[0] = creal(kx); kv[1] = creal(ky); kv[2] = creal(kz);
	E[0] = creal(ex); E[1] = cimag(ex); E[2] = creal(ey); E[3] = cimag(ey); E[4] = creal(ez); E[5] = cimag(ez);
}

/* Default-basis coefficients (electric block, magnetic block) at np points. */
static void qh_planewave_points(int lMax, size_t np, const double *pos_host, const double kv[3], const double E[6],
		complex double *out) {
	const size_t nb = 2 * (size_t)qpms_lMax2nelem(lMax);
	double *dpos = qh_dmalloc(3 * np * sizeof(double) + 6 * sizeof(double));
	double *dE = dpos + 3 * np;
	complex double *dout = qh_dmalloc(np * nb * sizeof(complex double));
	qh_h2d(dpos, pos_host, 3 * np * sizeof(double), NULL);
	qh_h2d(dE, E, 6 * sizeof(double), NULL);
	qh_launch_planewave(lMax, np, dpos, kv, dE, (double *)dout, (long)nb, NULL);
	qh_d2h(out, dout, np * nb * sizeof(complex double), NULL);
	qh_sync(NULL);
	qh_dfree(dout); qh_dfree(dpos);
}

/* Scatter default-basis coefficients into bspec order (vswf.c:504-526). */
static void qh_planewave_to_bspec(complex double *target, const qpms_vswf_set_spec_t *bspec, const complex double *c,
		size_t nelem, bool add) {
	if (!add) memset(target, 0, bspec->n * sizeof(complex double));
	for (size_t i = 0; i < bspec->n; ++i) {
		const qpms_uvswfi_t ui = bspec->ilist[i];
		const int t = (int)(ui % 4);
		const size_t ysc = (size_t)(ui / 4);
		if (ysc == 0) { target[i] = 0; continue; }
		const size_t y = ysc - 1;
		QH_ENSURE(t != QPMS_VSWF_LONGITUDINAL, "longitudinal incident-field components are not supported");
		target[i] += (t == QPMS_VSWF_ELECTRIC) ? c[y] : c[nelem + y];
	}
}

qpms_errno_t qpms_incfield_planewave(complex double *target, const qpms_vswf_set_spec_t *bspec,
		const cart3_t evalpoint, const void *args, bool add) {
	QH_ENSURE(bspec->norm == QPMS_NORMALISATION_DEFAULT, "only the default VSWF normalisation is supported");
	double kv[3], E[6];
	qh_planewave_kE((const qpms_incfield_planewave_params_t *)args, kv, E);
	const size_t ne = qpms_lMax2nelem(bspec->lMax);
	complex double *c;
	QH_MALLOC(c, 2 * ne * sizeof(complex double));
	const double pos[3] = {evalpoint.x, evalpoint.y, evalpoint.z};
	qh_planewave_points(bspec->lMax, 1, pos, kv, E, c);
	qh_planewave_to_bspec(target, bspec, c, ne, add);
	free(c);
	return QPMS_SUCCESS;
}

complex double *qpms_scatsys_incident_field_vector_full(complex double *target_full, const qpms_scatsys_t *ss,
		qpms_incfield_t f, const void *args, bool add) {
	if (!target_full) QH_CALLOC(target_full, ss->fecv_size, sizeof(complex double));
	if (f == qpms_incfield_planewave && qh_uniform_default_bspec(ss) && ss->c->normalisation == QPMS_NORMALISATION_DEFAULT) {
		/* every particle at once on the GPU */
		double kv[3], E[6];
		qh_planewave_kE((const qpms_incfield_planewave_params_t *)args, kv, E);
		const size_t np = ss->p_count, nb = 2 * ss->c->nelem;
		double *pos;
		QH_MALLOC(pos, 3 * np * sizeof(double));
		for (size_t i = 0; i < np; ++i) {
			pos[3 * i] = ss->p[i].pos.x; pos[3 * i + 1] = ss->p[i].pos.y; pos[3 * i + 2] = ss->p[i].pos.z;
		}
		complex double *c;
		QH_MALLOC(c, np * nb * sizeof(complex double));
		qh_planewave_points(ss->c->lMax, np, pos, kv, E, c);
		for (size_t i = 0; i < np * nb; ++i) target_full[i] = add ? target_full[i] + c[i] : c[i];
		free(c); free(pos);
		return target_full;
	}
	for (qpms_ss_pi_t pi = 0; pi < ss->p_count; ++pi)
		QH_ENSURE(f(target_full + ss->fecv_pstarts[pi], qh_bspec_pi(ss, pi), ss->p[pi].pos, args, add) == QPMS_SUCCESS,
			"incident field evaluation failed");
	return target_full;
}

complex double *qpms_scatsysw_apply_Tmatrices_full(complex double *target_full, const complex double *inc_full,
		const qpms_scatsys_at_omega_t *ssw) {
	const qpms_scatsys_t *ss = ssw->ss;
	if (!target_full) QH_CALLOC(target_full, ss->fecv_size, sizeof(complex double));
	/* group particles by T (tm index); every particle's block is nb x nb */
	for (qpms_ss_tmi_t ti = 0; ti < ss->tm_count; ++ti) {
		const qpms_tmatrix_t *T = ssw->tm[ti];
		const size_t nb = T->spec->n;
		size_t cnt = 0;
		for (qpms_ss_pi_t pi = 0; pi < ss->p_count; ++pi) cnt += (ss->p[pi].tmatrix_id == ti);
		if (!cnt) continue;
		complex double *a, *o;
		int *ty;
		QH_MALLOC(a, cnt * nb * sizeof(complex double));
		QH_MALLOC(o, cnt * nb * sizeof(complex double));
		QH_CALLOC(ty, cnt, sizeof(int));
		size_t j = 0;
		for (qpms_ss_pi_t pi = 0; pi < ss->p_count; ++pi)
			if (ss->p[pi].tmatrix_id == ti) memcpy(a + nb * j++, inc_full + ss->fecv_pstarts[pi], nb * sizeof(complex double));
		complex double *dT = qh_dmalloc((nb * nb + 2 * cnt * nb) * sizeof(complex double));
		complex double *da = dT + nb * nb, *dout = da + cnt * nb;
		int *dty = qh_dmalloc(cnt * sizeof(int));
		qh_h2d(dT, T->m, nb * nb * sizeof(complex double), NULL);
		qh_h2d(da, a, cnt * nb * sizeof(complex double), NULL);
		qh_h2d(dty, ty, cnt * sizeof(int), NULL);
		qh_launch_apply_t(cnt, (int)nb, dty, (const double *)dT, (const double *)da, (double *)dout, NULL);
		qh_d2h(o, dout, cnt * nb * sizeof(complex double), NULL);
		qh_sync(NULL);
		j = 0;
		for (qpms_ss_pi_t pi = 0; pi < ss->p_count; ++pi)
			if (ss->p[pi].tmatrix_id == ti) memcpy(target_full + ss->fecv_pstarts[pi], o + nb * j++, nb * sizeof(complex double));
		qh_dfree(dT); qh_dfree(dty);
		free(a); free(o); free(ty);
	}
	return target_full;
}

/* ------------------------------------------------------------------------ */
/* Scattered field at points (scatsystem.c:2013-2109)                        */
/* ------------------------------------------------------------------------ */
int qh_launch_scattered_e(int lMax, size_t npts, const double *where_dev, size_t npart, const double *pos_dev,
		double k_re, double k_im, int kind, const double *cvf_dev, int parity, double *out_dev, void *stream);

/* The coefficient vector in the kernel's layout: npart blocks of 2 nelem of
 * the calculator's lMax (electric, then magnetic), zeros where a particle's
 * basis has no entry. */
static complex double *qh_pad_cvf(const qpms_scatsys_t *ss, const complex double *cvf) {
	QH_ENSURE(ss->c->normalisation == QPMS_NORMALISATION_DEFAULT,
		"the scattered field is implemented for the default VSWF normalisation only");
	const size_t ne = ss->c->nelem, np = ss->p_count;
	complex double *c;
	QH_CALLOC(c, np * 2 * ne, sizeof(complex double));
	for (qpms_ss_pi_t pi = 0; pi < ss->p_count; ++pi) {
		const qpms_vswf_set_spec_t *b = qh_bspec_pi(ss, pi);
		QH_ENSURE(b->norm == QPMS_NORMALISATION_DEFAULT, "the scattered field is implemented for the default VSWF normalisation only");
		for (size_t i = 0; i < b->n; ++i) {
			const qpms_uvswfi_t u = b->ilist[i];
			const int t = (int)(u & 3);
			const long y = (long)(u / 4) - 1;
			QH_ENSURE((t == QPMS_VSWF_ELECTRIC || t == QPMS_VSWF_MAGNETIC) && y >= 0 && (size_t)y < ne,
				"scattered field: basis entry %llu is not an electric or magnetic wave of degree <= %d",
				(unsigned long long)u, (int)ss->c->lMax);
			c[pi * 2 * ne + (t == QPMS_VSWF_ELECTRIC ? 0 : ne) + (size_t)y] = cvf[ss->fecv_pstarts[pi] + i];
		}
	}
	return c;
}

/* The reference's __alt variants translate with the vector particle - point
 * (destination = the dipoles at the particle, source = the point), which is
 * the direct sum with electric coefficients times (-1)^(l+1) and magnetic ones
 * times (-1)^l. QPMS_HIP_SCATTERED_E_ALT=direct gives the direct sum instead. */
static int qh_alt_parity(void) {
	const char *e = getenv("QPMS_HIP_SCATTERED_E_ALT");
	if (!e || !strcmp(e, "reference")) return 1;
	QH_ENSURE(!strcmp(e, "direct"), "QPMS_HIP_SCATTERED_E_ALT must be 'reference' or 'direct', not '%s'", e);
	return 0;
}

int qpms_hip_scattered_E_dev(int lMax, size_t npart, const double *pos_dev, _Complex double k, qpms_bessel_t btyp,
		const double *cvf_dev, size_t npts, const double *where_dev, double *out_dev, int parity, void *stream) {
	QH_ENSURE(lMax >= 1 && lMax <= 4, "lMax %d outside 1..4", lMax);
	QH_ENSURE(btyp >= QPMS_BESSEL_REGULAR && btyp <= QPMS_HANKEL_MINUS, "invalid Bessel type %d", (int)btyp);
	QH_ENSURE(npts == 0 || (where_dev && out_dev && (npart == 0 || (pos_dev && cvf_dev))),
		"scattered field: NULL device buffer for %zu points, %zu particles", npts, npart);
	return qh_launch_scattered_e(lMax, npts, where_dev, npart, pos_dev, creal(k), cimag(k), (int)btyp, cvf_dev,
		parity, out_dev, stream);
}

static void qh_scattered_E_points(const qpms_scatsys_t *ss, qpms_bessel_t btyp, complex double k,
		const complex double *cvf, size_t npts, const cart3_t *where, ccart3_t *out, int parity) {
	if (npts == 0) return;
	complex double *c = qh_pad_cvf(ss, cvf);
	const size_t np = ss->p_count, nb = 2 * ss->c->nelem;
	double *dpos = qh_upload_positions(ss);
	double *dc = qh_dmalloc(np * nb * sizeof(complex double));
	double *dw = qh_dmalloc(npts * sizeof(cart3_t));
	double *dout = qh_dmalloc(npts * sizeof(ccart3_t));
	qh_h2d(dc, c, np * nb * sizeof(complex double), NULL);
	qh_h2d(dw, where, npts * sizeof(cart3_t), NULL);
	qpms_hip_scattered_E_dev((int)ss->c->lMax, np, dpos, k, btyp, dc, npts, dw, dout, parity, NULL);
	qh_d2h(out, dout, npts * sizeof(ccart3_t), NULL);
	qh_sync(NULL);
	qh_dfree(dout); qh_dfree(dw); qh_dfree(dc); qh_dfree(dpos);
	free(c);
}

void qpms_hip_scatsys_scattered_E_points(const qpms_scatsys_t *ss, qpms_bessel_t btyp, _Complex double k,
		const _Complex double *cvf, size_t npts, const cart3_t *where, ccart3_t *out) {
	qh_scattered_E_points(ss, btyp, k, cvf, npts, where, out, 0);
}

ccart3_t qpms_scatsys_scattered_E(const qpms_scatsys_t *ss, qpms_bessel_t btyp, _Complex double k,
		const _Complex double *cvf, const cart3_t where) {
	ccart3_t r;
	qh_scattered_E_points(ss, btyp, k, cvf, 1, &where, &r, 0);
	return r;
}

ccart3_t qpms_scatsysw_scattered_E(const qpms_scatsys_at_omega_t *ssw, qpms_bessel_t btyp,
		const _Complex double *cvf, const cart3_t where) {
	QH_ENSURE(ssw->ss->lattice_dimension == 0, "qpms_scatsysw_scattered_E(): applicable only to nonperiodic systems.");
	return qpms_scatsys_scattered_E(ssw->ss, btyp, ssw->wavenumber, cvf, where);
}

ccart3_t qpms_scatsys_scattered_E__alt(const qpms_scatsys_t *ss, qpms_bessel_t btyp, _Complex double k,
		const _Complex double *cvf, const cart3_t where) {
	QH_ENSURE(ss->lattice_dimension == 0, "qpms_scatsys_scattered_E__alt(): applicable only to nonperiodic systems.");
	ccart3_t r;
	qh_scattered_E_points(ss, btyp, k, cvf, 1, &where, &r, qh_alt_parity());
	return r;
}

ccart3_t qpms_scatsysw_scattered_E__alt(const qpms_scatsys_at_omega_t *ssw, qpms_bessel_t btyp,
		const _Complex double *cvf, const cart3_t where) {
	return qpms_scatsys_scattered_E__alt(ssw->ss, btyp, ssw->wavenumber, cvf, where);
}

/* scatsystem.c:2113-2188: periodic systems, through the lattice-summed
 * translation arrays (destination = regular electric dipoles, source = the
 * particle's basis, shift = point - particle) times the dipoles' value at the
 * origin. Points outside the lattice plane need out-of-plane Ewald shifts,
 * which the e32 arrays reject. */
ccart3_t qpms_scatsyswk_scattered_E(const qpms_scatsys_at_omega_k_t *sswk, qpms_bessel_t btyp,
		const _Complex double *cvf, const cart3_t where) {
	QH_ENSURE(btyp == QPMS_HANKEL_PLUS, "Only scattered field with first kind Hankel functions currently implemented.");
	const qpms_scatsys_t *ss = sswk->ssw->ss;
	QH_ENSURE(ss->lattice_dimension == 2, "Only 2D-periodic lattices implemented");
	QH_ENSURE(sswk->k[2] == 0 && ss->per.lattice_basis[0].z == 0 && ss->per.lattice_basis[1].z == 0,
		"lattice and Bloch vector must lie in the xy plane");
	QH_ENSURE(ss->c->normalisation == QPMS_NORMALISATION_DEFAULT,
		"the scattered field is implemented for the default VSWF normalisation only");
	const complex double k = sswk->ssw->wavenumber;
	/* the regular electric dipoles N_{1,-1..1} at r = 0 (theta = pi/2, phi = 0) */
	qpms_uvswfi_t dil[3] = {qpms_tmn2uvswfi(QPMS_VSWF_ELECTRIC, -1, 1), qpms_tmn2uvswfi(QPMS_VSWF_ELECTRIC, 0, 1),
		qpms_tmn2uvswfi(QPMS_VSWF_ELECTRIC, 1, 1)};
	qpms_vswf_set_spec_t dipspec;
	memset(&dipspec, 0, sizeof dipspec);
	dipspec.n = 3; dipspec.ilist = dil; dipspec.lMax = 1; dipspec.lMax_M = 0; dipspec.lMax_N = 1;
	dipspec.lMax_L = -1; dipspec.norm = ss->c->normalisation;
	ccart3_t dip[3];
	{
		/* three unit coefficient vectors of lMax 1 (6 complex each), the origin
		 * (3 doubles, padded to 4 so that the results stay 16-byte aligned) */
		double h[3 * 12 + 4] = {0};
		for (int j = 0; j < 3; ++j) h[12 * j + 2 * j] = 1;
		double *d = qh_dmalloc(sizeof h + 3 * sizeof(ccart3_t));
		double *dcv = d, *dw = d + 36, *dout = d + 40;
		qh_h2d(d, h, sizeof h, NULL);
		for (int j = 0; j < 3; ++j)
			qpms_hip_scattered_E_dev(1, 1, dw, 0, QPMS_BESSEL_REGULAR, dcv + 12 * j, 1, dw, dout + 6 * j, 0, NULL);
		qh_d2h(dip, dout, 3 * sizeof(ccart3_t), NULL);
		qh_sync(NULL);
		qh_dfree(d);
	}
	const double maxR = sqrt(ss->per.unitcell_volume) * 64;   /* scatsystem.c:2160-2161 */
	const double maxK = 2048 * 2 * M_PI / maxR;
	const cart2_t b1 = {ss->per.lattice_basis[0].x, ss->per.lattice_basis[0].y};
	const cart2_t b2 = {ss->per.lattice_basis[1].x, ss->per.lattice_basis[1].y};
	const cart2_t beta = {sswk->k[0], sswk->k[1]};
	complex double *s;
	QH_MALLOC(s, ss->max_bspecn * 3 * sizeof(complex double));
	complex double e[3] = {0, 0, 0};
	for (qpms_ss_pi_t pi = 0; pi < ss->p_count; ++pi) {
		const qpms_vswf_set_spec_t *bspec = qh_bspec_pi(ss, pi);
		const cart3_t pp = ss->p[pi].pos;
		const cart3_t sh = {where.x - pp.x, where.y - pp.y, where.z - pp.z};
		QH_ENSURE(qpms_trans_calculator_get_trans_array_e32(ss->c, s, NULL, &dipspec, 1, bspec, 3, sswk->eta, k,
			b1, b2, beta, sh, maxR, maxK) == QPMS_SUCCESS, "lattice-summed translation array failed");
		const complex double *cv = cvf + ss->fecv_pstarts[pi];
		for (size_t i = 0; i < bspec->n; ++i)
			for (int j = 0; j < 3; ++j) {
				const complex double f = cv[i] * s[3 * i + j];
				e[0] += f * dip[j].x; e[1] += f * dip[j].y; e[2] += f * dip[j].z;
			}
	}
	free(s);
	ccart3_t r = {e[0], e[1], e[2]};
	return r;
}
