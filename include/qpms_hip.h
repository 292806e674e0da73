/* qpms_hip.h — C ABI of the MI355X (gfx950) multiple-scattering hot path.
 *
 * This header is the drop-in boundary. Every function below that carries a
 * reference citation keeps the reference's name, argument order, argument
 * meaning and error behaviour (abort with a `file:line, func: msg` line on
 * stderr); every struct that the reference's callers read keeps the field
 * order and types of the reference declaration it cites, so that code built
 * against the reference headers can link against libqpms_hip.so.
 *
 * All reference entry points take and return HOST pointers, exactly like the
 * reference. The computation behind them runs on the GPU: the host copies the
 * inputs to HBM, launches the hand-written HIP kernels and copies the results
 * back. The `*_dev` extensions take device pointers and a hipStream_t (passed
 * as `void *`, NULL = the library's own stream) and never touch host memory.
 *
 * Citations are `path:line` into kostyfisik/qpms.
 */
#ifndef QPMS_HIP_H
#define QPMS_HIP_H

#ifndef __cplusplus
#include <complex.h>
#endif
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Basic types — qpms/qpms_types.h                                           */
/* ------------------------------------------------------------------------ */

typedef int qpms_lm_t;              /* qpms_types.h:25 */
typedef int qpms_l_t;               /* qpms_types.h:27 */
typedef qpms_lm_t qpms_m_t;         /* qpms_types.h:31 */
typedef size_t qpms_y_t;            /* qpms_types.h:44 */
typedef size_t qpms_y_sc_t;         /* qpms_types.h:59 */
typedef unsigned long long qpms_uvswfi_t; /* qpms_types.h:81 */
typedef int32_t qpms_ss_tmi_t;      /* qpms_types.h:310 */
typedef int32_t qpms_ss_tmgi_t;     /* qpms_types.h:313 */
typedef int32_t qpms_ss_pi_t;       /* qpms_types.h:316 */
typedef int qpms_gmi_t;             /* qpms_types.h:320 */
typedef int qpms_iri_t;             /* qpms_types.h:323 */
#define QPMS_NO_IRREP ((qpms_iri_t) -1)

/* qpms_types.h:62-66 */
typedef enum {
	QPMS_VSWF_ELECTRIC = 2,
	QPMS_VSWF_MAGNETIC = 1,
	QPMS_VSWF_LONGITUDINAL = 0
} qpms_vswf_type_t;

/* qpms_types.h:85-92 */
typedef enum {
	QPMS_SUCCESS = 0,
	QPMS_ERROR = -1,
	QPMS_FAILURE = QPMS_ERROR,
	QPMS_ENOMEM = 8,
	QPMS_ESING = 21,
	QPMS_NAN_ENCOUNTERED = 1024
} qpms_errno_t;

/* qpms_types.h:104-168 (values must match bit for bit) */
typedef enum {
	QPMS_NORMALISATION_UNDEF = 0,
	QPMS_NORMALISATION_INVERSE = 1,
	QPMS_NORMALISATION_REVERSE_AZIMUTHAL_PHASE = 2,
	QPMS_NORMALISATION_SPHARM_REAL = 4,
	QPMS_NORMALISATION_CSPHASE = 8,
	QPMS_NORMALISATION_M_I = 16,
	QPMS_NORMALISATION_M_MINUS = 32,
	QPMS_NORMALISATION_N_I = 64,
	QPMS_NORMALISATION_N_MINUS = 128,
	QPMS_NORMALISATION_L_I = 256,
	QPMS_NORMALISATION_L_MINUS = 512,
	QPMS_NORMALISATION_NORM_BITSTART = 65536,
	QPMS_NORMALISATION_NORM_POWER = 65536 * 1,
	QPMS_NORMALISATION_NORM_SPHARM = 65536 * 3,
	QPMS_NORMALISATION_NORM_NONE = 65536 * 2,
	QPMS_NORMALISATION_NORM_BITS = 65536 * 3,
	QPMS_NORMALISATION_CONVENTION_KRISTENSSON_REAL = 65536 | 4,
	QPMS_NORMALISATION_CONVENTION_KRISTENSSON = 65536 | 8,
	QPMS_NORMALISATION_CONVENTION_SCUFF = 65536 | 8 | 16 | 128,
	QPMS_NORMALISATION_DEFAULT = 65536 | 8
} qpms_normalisation_t;

/* qpms_types.h:176-182 */
typedef enum {
	QPMS_BESSEL_REGULAR = 1,
	QPMS_BESSEL_SINGULAR = 2,
	QPMS_HANKEL_PLUS = 3,
	QPMS_HANKEL_MINUS = 4,
	QPMS_BESSEL_UNDEF = 0
} qpms_bessel_t;

typedef struct cart3_t { double x, y, z; } cart3_t;            /* qpms_types.h:186 */
typedef struct cart2_t { double x, y; } cart2_t;                /* qpms_types.h:202 */
typedef struct sph_t { double r, theta, phi; } sph_t;           /* qpms_types.h:207 */
typedef struct csph_t { _Complex double r; double theta, phi; } csph_t; /* qpms_types.h:212 */
typedef struct pol_t { double r, phi; } pol_t;                  /* qpms_types.h:223 */
typedef struct ccart3_t { _Complex double x, y, z; } ccart3_t;  /* qpms_types.h:191 */
typedef struct csphvec_t { _Complex double rc, thetac, phic; } csphvec_t; /* qpms_types.h:218 */
typedef union anycoord_point_t {                               /* qpms_types.h:229 */
	double z; cart3_t cart3; cart2_t cart2; sph_t sph; pol_t pol;
} anycoord_point_t;
typedef struct qpms_quat_t { _Complex double a, b; } qpms_quat_t; /* qpms_types.h:261 */
typedef struct qpms_irot3_t { qpms_quat_t rot; short det; } qpms_irot3_t; /* qpms_types.h:275 */

/* qpms_types.h:290-307 */
typedef struct qpms_vswf_set_spec_t {
	size_t n;
	qpms_uvswfi_t *ilist;
	qpms_l_t lMax;
	qpms_l_t lMax_M, lMax_N, lMax_L;
	size_t capacity;
	qpms_normalisation_t norm;
} qpms_vswf_set_spec_t;

/* qpms_types.h:338-349 */
typedef struct qpms_tmatrix_t {
	const qpms_vswf_set_spec_t *spec;
	_Complex double *m;
	bool owns_m;
} qpms_tmatrix_t;

/* qpms_types.h:422-425 */
typedef struct qpms_epsmu_t { _Complex double eps; _Complex double mu; } qpms_epsmu_t;

/* ------------------------------------------------------------------------ */
/* Index helpers — qpms/indexing.h (same formulas, needed by callers)         */
/* ------------------------------------------------------------------------ */
static inline qpms_y_t qpms_mn2y(qpms_m_t m, qpms_l_t n) { return n * (n + 1) + m - 1; }
static inline qpms_y_t qpms_mn2y_sc(qpms_m_t m, qpms_l_t n) { return n * (n + 1) + m; }
static inline qpms_y_t qpms_lMax2nelem(qpms_l_t lmax) { return lmax * ((qpms_y_t)lmax + 2); }
static inline qpms_y_t qpms_lMax2nelem_sc(qpms_l_t lmax) { return lmax * ((qpms_y_t)lmax + 2) + 1; }
static inline qpms_uvswfi_t qpms_tmn2uvswfi(qpms_vswf_type_t t, qpms_m_t m, qpms_l_t n) {
	return t + 4 * qpms_mn2y_sc(m, n);
}

/* ------------------------------------------------------------------------ */
/* Materials, T-matrix generators — materials.h, tmatrices.h                 */
/* ------------------------------------------------------------------------ */
#ifndef SPEED_OF_LIGHT
#define SPEED_OF_LIGHT (2.99792458e8) /* materials.h:11 */
#endif

/* materials.h:16-25 */
typedef struct qpms_epsmu_generator_t {
	qpms_epsmu_t (*function)(_Complex double omega, const void *params);
	const void *params;
} qpms_epsmu_generator_t;

/* materials.h:34: params point to a qpms_epsmu_t; frequency ignored. */
qpms_epsmu_t qpms_epsmu_const_g(_Complex double omega, const void *epsmu);

/* tmatrices.h:258-264 */
typedef struct qpms_tmatrix_generator_t {
	qpms_errno_t (*function)(qpms_tmatrix_t *t, _Complex double omega, const void *params);
	const void *params;
} qpms_tmatrix_generator_t;

/* tmatrices.h:526-536 */
typedef struct qpms_tmatrix_function_t {
	const qpms_vswf_set_spec_t *spec;
	const qpms_tmatrix_generator_t *gen;
} qpms_tmatrix_function_t;

/* tmatrices.h:367-371 */
typedef struct qpms_tmatrix_generator_sphere_param_t {
	qpms_epsmu_generator_t outside;
	qpms_epsmu_generator_t inside;
	double radius;
} qpms_tmatrix_generator_sphere_param_t;

/* tmatrices.h:545-556; only NOOP is accepted by this library. */
typedef enum { QPMS_TMATRIX_OPERATION_NOOP = 0 } qpms_tmatrix_operation_kind_t;

/* tmatrices.h:609-623. The union members are not interpreted here; the
 * storage reproduces the reference's size (largest member: 40 bytes). */
typedef struct qpms_tmatrix_operation_t {
	qpms_tmatrix_operation_kind_t typ;
	union {
		struct { _Complex double *m; size_t m_size; bool owns_m; } lrmatrix;
		qpms_irot3_t irot3;
		struct { size_t n; void *ops; void *opmem; size_t opmem_size; bool *ops_owned; } compose_chain;
		const void *finite_group;
	} op;
} qpms_tmatrix_operation_t;

qpms_vswf_set_spec_t *qpms_vswf_set_spec_from_lMax(qpms_l_t lMax, qpms_normalisation_t norm); /* vswf.h, vswf.c:76-92 */
void qpms_vswf_set_spec_free(qpms_vswf_set_spec_t *s);                                        /* vswf.c:94-97 */
qpms_tmatrix_t *qpms_tmatrix_init(const qpms_vswf_set_spec_t *bspec);                         /* tmatrices.c:39-47 */
void qpms_tmatrix_free(qpms_tmatrix_t *t);                                                    /* tmatrices.c:66-69 */
/* tmatrices.c:580-591: Lorentz–Mie T-matrix of a sphere (params: qpms_tmatrix_generator_sphere_param_t). */
qpms_errno_t qpms_tmatrix_generator_sphere(qpms_tmatrix_t *t, _Complex double omega, const void *params);
/* tmatrices.c:500-561: Mie reflection coefficients; target of length bspec->n (NULL: allocated). */
_Complex double *qpms_mie_coefficients_reflection(_Complex double *target, const qpms_vswf_set_spec_t *bspec,
		double a, _Complex double k_i, _Complex double k_e, _Complex double mu_i, _Complex double mu_e,
		qpms_bessel_t J_ext, qpms_bessel_t J_scat);

/* ------------------------------------------------------------------------ */
/* Lattice point generators — lattices.h:120-248, latticegens.c:426-580      */
/* ------------------------------------------------------------------------ */
typedef enum {                                     /* lattices.h:33-48 */
	LAT1D = 1, LAT2D = 2, LAT3D = 4, SPACE1D = 8, SPACE2D = 16, SPACE3D = 32,
	LAT_1D_IN_3D = 33, LAT_2D_IN_3D = 34, LAT_3D_IN_3D = 40,
	LAT_ZONLY = 64, LAT_XYONLY = 128,
	LAT_1D_IN_3D_ZONLY = 97, LAT_2D_IN_3D_XYONLY = 162
} LatticeDimensionality;

typedef enum PGenPointFlags {                      /* lattices.h:132-169 */
	PGEN_NOTDONE = 2, PGEN_OLD_R = 1, PGEN_SINGLE_R = 16, PGEN_AT_Z = 4, PGEN_AT_XY = 8,
	PGEN_METHOD_UNAVAILABLE = 2048, PGEN_DONE = 0,
	PGEN_COORDS_CART1 = 64, PGEN_COORDS_POL = 128, PGEN_COORDS_SPH = 256,
	PGEN_COORDS_CART2 = 512, PGEN_COORDS_CART3 = 1024, PGEN_COORDS_BITRANGE = 1984
} PGenPointFlags;

typedef struct PGen { const struct PGenClassInfo *c; void *stateData; } PGen; /* lattices.h:127-130 */
typedef struct PGenReturnDataBulk { PGenPointFlags flags; size_t generated; } PGenReturnDataBulk;
typedef struct PGenReturnData { PGenPointFlags flags; anycoord_point_t point; } PGenReturnData;
typedef struct PGenZReturnData { PGenPointFlags flags; double point_z; } PGenZReturnData;
typedef struct PGenPolReturnData { PGenPointFlags flags; pol_t point_pol; } PGenPolReturnData;
typedef struct PGenSphReturnData { PGenPointFlags flags; sph_t point_sph; } PGenSphReturnData;
typedef struct PGenCart2ReturnData { PGenPointFlags flags; cart2_t point_cart2; } PGenCart2ReturnData;
typedef struct PGenCart3ReturnData { PGenPointFlags flags; cart3_t point_cart3; } PGenCart3ReturnData;

typedef struct PGenClassInfo {                     /* lattices.h:215-236 */
	char *const name;
	int dimensionality;
	PGenPointFlags native_point_flags;
	PGenReturnData (*next)(struct PGen *);
	PGenZReturnData (*next_z)(struct PGen *);
	PGenPolReturnData (*next_pol)(struct PGen *);
	PGenSphReturnData (*next_sph)(struct PGen *);
	PGenCart2ReturnData (*next_cart2)(struct PGen *);
	PGenCart3ReturnData (*next_cart3)(struct PGen *);
	PGenReturnDataBulk (*fetch)(struct PGen *, size_t, anycoord_point_t *);
	PGenReturnDataBulk (*fetch_z)(struct PGen *, size_t, double *);
	PGenReturnDataBulk (*fetch_pol)(struct PGen *, size_t, pol_t *);
	PGenReturnDataBulk (*fetch_sph)(struct PGen *, size_t, sph_t *);
	PGenReturnDataBulk (*fetch_cart2)(struct PGen *, size_t, cart2_t *);
	PGenReturnDataBulk (*fetch_cart3)(struct PGen *, size_t, cart3_t *);
	void (*destructor)(struct PGen *);
} PGenClassInfo;

/* latticegens.c:449-468: hexagonal/rectangular "spiderweb" ring generator. */
PGen PGen_xyWeb_new(cart2_t b1, cart2_t b2, double rtol, cart2_t offset,
		double minR, bool inc_minR, double maxR, bool inc_maxR);
PGenSphReturnData PGen_next_sph(PGen *g);   /* lattices.h: dispatch through g->c */
PGenCart2ReturnData PGen_next_cart2(PGen *g);
void PGen_destroy(PGen *g);
#define BASIS_RTOL 1e-13                   /* lattices.h:650 */

/* lattices2d.c:635, 816, 850: 2D lattice helpers used by the Ewald entry points. */
void l2d_reduceBasis(cart2_t in1, cart2_t in2, cart2_t *out1, cart2_t *out2);
int l2d_reciprocalBasis2pi(cart2_t b1, cart2_t b2, cart2_t *rb1, cart2_t *rb2);
double l2d_unitcell_area(cart2_t b1, cart2_t b2);
double l2d_hexWebInCircleRadius(cart2_t b1, cart2_t b2);

/* ------------------------------------------------------------------------ */
/* Ewald constants and σ sums — ewald.h:39-340                                */
/* ------------------------------------------------------------------------ */
typedef enum {                                     /* ewald.h:39-44 */
	QPMS_EWALD_LONG_RANGE = 1,
	QPMS_EWALD_SHORT_RANGE = 2,
	QPMS_EWALD_0TERM = 4,
	QPMS_EWALD_FULL = 7
} qpms_ewald_part;

/* ewald.h:56-131. `legendre_normconv` holds a gsl_sf_legendre_t (an int-sized
 * enum; 3 = GSL_SF_LEGENDRE_NONE). */
typedef struct qpms_ewald3_constants_t {
	qpms_l_t lMax;
	qpms_y_t nelem_sc;
	qpms_l_t *s1_jMaxes;
	_Complex double **s1_constfacs;
	_Complex double *s1_constfacs_base;
	_Complex double **S1_constfacs;
	_Complex double *S1_constfacs_base;
	_Complex double **s1_constfacs_1Dz;
	_Complex double *s1_constfacs_1Dz_base;
	double *legendre0;
	double *legendre_plus1;
	double *legendre_minus1;
	int legendre_normconv;
	int legendre_csphase;
} qpms_ewald3_constants_t;

typedef struct qpms_csf_result { _Complex double val; double err; } qpms_csf_result; /* ewald.h:142-145 */

qpms_ewald3_constants_t *qpms_ewald3_constants_init(qpms_l_t lMax, int csphase); /* ewald.c:104-243 */
void qpms_ewald3_constants_free(qpms_ewald3_constants_t *c);                      /* ewald.c:246-257 */

/* ewald.c:260-273 */
int ewald3_sigma0(_Complex double *result, double *err, const qpms_ewald3_constants_t *c,
		double eta, _Complex double wavenumber);
/* ewald.c:742-920 — runs k_ewald_short on the GPU over the points the generator yields. */
int ewald3_sigma_short(_Complex double *target_sigmasr_y, double *target_sigmasr_y_err,
		const qpms_ewald3_constants_t *c, double eta, _Complex double wavenumber,
		LatticeDimensionality latdim, PGen *pgen_R, bool pgen_generates_shifted_points,
		cart3_t k, cart3_t particle_shift);
/* ewald.c:623-649 → 286-503 — runs k_ewald_long on the GPU. */
int ewald3_sigma_long(_Complex double *target_sigmalr_y, double *target_sigmalr_y_err,
		const qpms_ewald3_constants_t *c, double eta, _Complex double wavenumber,
		double unitcell_volume, LatticeDimensionality latdim, PGen *pgen_K,
		bool pgen_generates_shifted_points, cart3_t k, cart3_t particle_shift);

/* ------------------------------------------------------------------------ */
/* Translation calculator — translations.h:64-207                            */
/* ------------------------------------------------------------------------ */
/* translations.h:64-84 (built with LATTICESUMS32, so e3c is present). */
typedef struct qpms_trans_calculator {
	qpms_normalisation_t normalisation;
	qpms_l_t lMax;
	qpms_y_t nelem;
	_Complex double **A_multipliers;
	_Complex double **B_multipliers;
	qpms_ewald3_constants_t *e3c;
	double *legendre0;
} qpms_trans_calculator;

qpms_trans_calculator *qpms_trans_calculator_init(qpms_l_t lMax, qpms_normalisation_t nt); /* translations.c:439-508 */
void qpms_trans_calculator_free(qpms_trans_calculator *c);                                  /* translations.c:288-298 */

/* translations.c:688-698 */
int qpms_trans_calculator_get_AB_arrays(const qpms_trans_calculator *c,
		_Complex double *Adest, _Complex double *Bdest,
		size_t deststride, size_t srcstride,
		csph_t kdlj, bool r_ge_d, qpms_bessel_t J);
/* translations.c:1124-1132 */
int qpms_trans_calculator_get_AB_arrays_ext(const qpms_trans_calculator *c,
		_Complex double *Adest, _Complex double *Bdest,
		size_t deststride, size_t srcstride,
		_Complex double kdlj_r, double kdlj_theta, double kdlj_phi,
		int r_ge_d, int J);
/* translations.c:701-722 */
qpms_errno_t qpms_trans_calculator_get_trans_array(const qpms_trans_calculator *c,
		_Complex double *target,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		csph_t kdlj, bool r_ge_d, qpms_bessel_t J);
/* translations.c:796-812 */
qpms_errno_t qpms_trans_calculator_get_trans_array_lc3p(const qpms_trans_calculator *c,
		_Complex double *target,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		_Complex double k, cart3_t destpos, cart3_t srcpos, qpms_bessel_t J);
/* translations.c:926-1094 */
int qpms_trans_calculator_get_AB_arrays_e32(const qpms_trans_calculator *c,
		_Complex double *Adest, double *Aerr, _Complex double *Bdest, double *Berr,
		ptrdiff_t deststride, ptrdiff_t srcstride,
		double eta, _Complex double wavenumber, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK);
int qpms_trans_calculator_get_AB_arrays_e32_e(const qpms_trans_calculator *c,
		_Complex double *Adest, double *Aerr, _Complex double *Bdest, double *Berr,
		ptrdiff_t deststride, ptrdiff_t srcstride,
		double eta, _Complex double wavenumber, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK,
		qpms_ewald_part parts);
/* translations.c:724-793 */
qpms_errno_t qpms_trans_calculator_get_trans_array_e32(const qpms_trans_calculator *c,
		_Complex double *target, double *err,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		double eta, _Complex double wavenumber, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK);
qpms_errno_t qpms_trans_calculator_get_trans_array_e32_e(const qpms_trans_calculator *c,
		_Complex double *target, double *err,
		const qpms_vswf_set_spec_t *destspec, size_t deststride,
		const qpms_vswf_set_spec_t *srcspec, size_t srcstride,
		double eta, _Complex double wavenumber, cart2_t b1, cart2_t b2,
		cart2_t beta, cart3_t particle_shift, double maxR, double maxK,
		qpms_ewald_part parts);

/* ------------------------------------------------------------------------ */
/* Scattering system — scatsystem.h:36-549                                   */
/* ------------------------------------------------------------------------ */
typedef struct qpms_particle_tid_t { cart3_t pos; qpms_ss_tmi_t tmatrix_id; } qpms_particle_tid_t; /* scatsystem.h:36-40 */

/* scatsystem.h:128-131 */
typedef struct qpms_ss_derived_tmatrix_t {
	qpms_ss_tmgi_t tmgi;
	qpms_tmatrix_operation_t op;
} qpms_ss_derived_tmatrix_t;

/* scatsystem.h:133-159 */
typedef struct qpms_scatsys_periodic_info_t {
	cart3_t lattice_basis[3];
	cart3_t reciprocal_basis[3];
	double unitcell_volume;
	double eta;
} qpms_scatsys_periodic_info_t;

struct qpms_finite_group_t;
struct qpms_tolerance_spec_t;
struct qpms_ss_orbit_type_t;
struct qpms_ss_particle_orbitinfo;

/* scatsystem.h:169-227. The symmetry-packing fields are kept for layout;
 * this library fills them for the trivial group only. */
typedef struct qpms_scatsys_t {
	int lattice_dimension;
	qpms_epsmu_generator_t medium;
	qpms_tmatrix_function_t *tmg;
	qpms_ss_tmgi_t tmg_count;
	qpms_ss_derived_tmatrix_t *tm;
	qpms_ss_tmi_t tm_count;
	qpms_ss_tmi_t tm_capacity;
	qpms_particle_tid_t *p;
	qpms_ss_pi_t p_count;
	qpms_ss_pi_t p_capacity;
	const struct qpms_finite_group_t *sym;
	qpms_ss_pi_t *p_sym_map;
	qpms_ss_tmi_t *tm_sym_map;
	qpms_ss_tmi_t orbit_type_count;
	struct qpms_ss_orbit_type_t *orbit_types;
	struct qpms_ss_particle_orbitinfo *p_orbitinfo;
	size_t fecv_size;
	size_t *saecv_sizes;
	size_t *fecv_pstarts;
	size_t *saecv_ot_offsets;
	size_t max_bspecn;
	qpms_ss_pi_t *p_by_orbit;
	char *otspace;
	char *otspace_end;
	double lenscale;
	struct qpms_trans_calculator *c;
	qpms_scatsys_periodic_info_t per;
} qpms_scatsys_t;

/* scatsystem.h:239-249 */
typedef struct qpms_scatsys_at_omega_t {
	const qpms_scatsys_t *ss;
	qpms_tmatrix_t **tm;
	_Complex double omega;
	qpms_epsmu_t medium;
	_Complex double wavenumber;
} qpms_scatsys_at_omega_t;

/* scatsystem.h:488-492 */
typedef struct qpms_scatsys_at_omega_k_t {
	const qpms_scatsys_at_omega_t *ssw;
	double k[3];
	double eta;
} qpms_scatsys_at_omega_k_t;

/* scatsystem.h:432-441 */
typedef struct qpms_ss_LU {
	const qpms_scatsys_at_omega_t *ssw;
	const struct qpms_scatsys_at_omega_k_t *sswk;
	bool full;
	qpms_iri_t iri;
	_Complex double *a;
	int *ipiv;
} qpms_ss_LU;

/* scatsystem.h:292-296 / scatsystem.c:233-563. Only the trivial symmetry
 * group (sym == NULL or a group of order 1) is supported; the irrep-packed
 * paths are not on this library's path. */
qpms_scatsys_at_omega_t *qpms_scatsys_apply_symmetry(const qpms_scatsys_t *orig,
		const struct qpms_finite_group_t *sym, _Complex double omega,
		const struct qpms_tolerance_spec_t *tol);
void qpms_scatsys_free(qpms_scatsys_t *s);                                  /* scatsystem.c:569-587 */
qpms_scatsys_at_omega_t *qpms_scatsys_at_omega(const qpms_scatsys_t *ss, _Complex double omega); /* :606-630 */
void qpms_scatsys_at_omega_free(qpms_scatsys_at_omega_t *ssw);              /* :632-640 */
double qpms_ss_adjusted_eta(const qpms_scatsys_t *ss, _Complex double wavenumber, const double k[3]); /* :68-80 */

/* scatsystem.c:1155-1164: W with J = HANKEL_PLUS; diagonal blocks zero. */
_Complex double *qpms_scatsys_build_translation_matrix_full(_Complex double *target,
		const qpms_scatsys_t *ss, _Complex double k);
/* scatsystem.c:1180-1217 */
_Complex double *qpms_scatsys_build_translation_matrix_e_full(_Complex double *target,
		const qpms_scatsys_t *ss, _Complex double k, qpms_bessel_t J);
/* scatsystem.c:1166-1178 */
_Complex double *qpms_scatsyswk_build_translation_matrix_full(_Complex double *target,
		const qpms_scatsys_at_omega_k_t *sswk);
/* scatsystem.c:1252-1281 */
_Complex double *qpms_scatsys_periodic_build_translation_matrix_full(_Complex double *target,
		const qpms_scatsys_t *ss, _Complex double wavenumber, const cart3_t *wavevector, double eta);
/* scatsystem.c:1345-1350 — intended semantics (the reference aborts here,
 * scatsystem.c:1293): M = I − T_R S(R←C), no self block. */
_Complex double *qpms_scatsysw_build_modeproblem_matrix_full(_Complex double *target,
		const qpms_scatsys_at_omega_t *ssw);
/* Spelling of the same entry used by some callers; identical behaviour. */
_Complex double *qpms_scatsys_build_modeproblem_matrix_full(_Complex double *target,
		const qpms_scatsys_at_omega_t *ssw);
/* scatsystem.c:1352-1357 */
_Complex double *qpms_scatsyswk_build_modeproblem_matrix_full(_Complex double *target,
		const qpms_scatsys_at_omega_k_t *sswk);

void qpms_ss_LU_free(qpms_ss_LU lu);                                        /* scatsystem.c:2197-2200 */
/* scatsystem.c:2202-2224 — FP64 LU on the GPU, ipiv 1-based like zgetrf. */
qpms_ss_LU qpms_scatsysw_modeproblem_matrix_full_factorise(_Complex double *modeproblem_matrix_full,
		int *target_piv, const qpms_scatsys_at_omega_t *ssw, const qpms_scatsys_at_omega_k_t *sswk);
/* scatsystem.c:2243-2249 */
qpms_ss_LU qpms_scatsysw_build_modeproblem_matrix_full_LU(_Complex double *target, int *target_piv,
		const qpms_scatsys_at_omega_t *ssw);
/* scatsystem.c:2251-2256 */
qpms_ss_LU qpms_scatsyswk_build_modeproblem_matrix_full_LU(_Complex double *target, int *target_piv,
		const qpms_scatsys_at_omega_k_t *sswk);
/* scatsystem.c:2265-2273 */
_Complex double *qpms_scatsys_scatter_solve(_Complex double *f, const _Complex double *a_inc, qpms_ss_LU lu);

/* ------------------------------------------------------------------------ */
/* Incident fields — vswf.h:17-124, scatsystem.h:618-636                     */
/* ------------------------------------------------------------------------ */
/* vswf.h:17-24 */
typedef qpms_errno_t (*qpms_incfield_t)(_Complex double *target, const qpms_vswf_set_spec_t *bspec,
		const cart3_t evalpoint, const void *args, bool add);

/* vswf.h:93-103 */
typedef struct qpms_incfield_planewave_params_t {
	bool use_cartesian;
	union { ccart3_t cart; csph_t sph; } k;
	union { ccart3_t cart; csphvec_t sph; } E;
} qpms_incfield_planewave_params_t;

/* vswf.c:480-530: regular-VSWF coefficients of a plane wave at evalpoint,
 * computed on the GPU (default normalisation; real wave vector, as the
 * reference requires). */
qpms_errno_t qpms_incfield_planewave(_Complex double *target, const qpms_vswf_set_spec_t *bspec,
		const cart3_t evalpoint, const void *args, bool add);
/* scatsystem.c:1970-1982 (one launch for all particles when f is
 * qpms_incfield_planewave and every particle has the default basis). */
_Complex double *qpms_scatsys_incident_field_vector_full(_Complex double *target_full, const qpms_scatsys_t *ss,
		qpms_incfield_t field_at_point, const void *args, bool add);
/* scatsystem.c:1996-2011 */
_Complex double *qpms_scatsysw_apply_Tmatrices_full(_Complex double *target_full, const _Complex double *inc_full,
		const qpms_scatsys_at_omega_t *ssw);

/* ------------------------------------------------------------------------ */
/* Scattered field at points — scatsystem.h:704-792                          */
/* ------------------------------------------------------------------------ */
/* scatsystem.h:715-721, scatsystem.c:2013-2037: the field of the coefficient
 * vector cvf (fecv layout) at `where`, waves of kind btyp (QPMS_HANKEL_PLUS
 * for a scattered field) at wavenumber k, evaluated on the GPU. Default VSWF
 * normalisation; electric and magnetic waves up to the calculator's lMax (1..4). */
ccart3_t qpms_scatsys_scattered_E(const qpms_scatsys_t *ss, qpms_bessel_t btyp, _Complex double k,
		const _Complex double *cvf, const cart3_t where);
/* scatsystem.h:734-740, scatsystem.c:2039-2045 (finite systems only, aborts otherwise) */
ccart3_t qpms_scatsysw_scattered_E(const qpms_scatsys_at_omega_t *ssw, qpms_bessel_t btyp,
		const _Complex double *cvf, const cart3_t where);
/* scatsystem.h:750-756, scatsystem.c:2048-2109: the reference's translation-based
 * variant. It translates with the vector particle - point, which gives the direct
 * sum with electric coefficients times (-1)^(l+1) and magnetic ones times (-1)^l;
 * that value is returned (QPMS_HIP_SCATTERED_E_ALT=direct returns the direct sum). */
ccart3_t qpms_scatsys_scattered_E__alt(const qpms_scatsys_t *ss, qpms_bessel_t btyp, _Complex double k,
		const _Complex double *cvf, const cart3_t where);
/* scatsystem.h:767-773, scatsystem.c:2105-2109 */
ccart3_t qpms_scatsysw_scattered_E__alt(const qpms_scatsys_at_omega_t *ssw, qpms_bessel_t btyp,
		const _Complex double *cvf, const cart3_t where);
/* scatsystem.h:786-792, scatsystem.c:2113-2188: 2D-periodic systems, from the
 * lattice-summed translation arrays (QPMS_HANKEL_PLUS only; points in the plane
 * of the lattice, as the e32 arrays take in-plane shifts only). */
ccart3_t qpms_scatsyswk_scattered_E(const qpms_scatsys_at_omega_k_t *sswk, qpms_bessel_t btyp,
		const _Complex double *cvf, const cart3_t where);
/* Extension: qpms_scatsys_scattered_E at npts points in one launch (host arrays). */
void qpms_hip_scatsys_scattered_E_points(const qpms_scatsys_t *ss, qpms_bessel_t btyp, _Complex double k,
		const _Complex double *cvf, size_t npts, const cart3_t *where, ccart3_t *out);
/* Extension, device pointers on `stream`: particles at pos_dev[3 npart], coefficients
 * cvf_dev[npart][2 nelem(lMax)] (complex; electric block, then magnetic, default
 * order), points where_dev[3 npts]; out_dev[3 npts] complex (x, y, z). parity != 0
 * multiplies electric coefficients by (-1)^(l+1), magnetic ones by (-1)^l. */
int qpms_hip_scattered_E_dev(int lMax, size_t npart, const double *pos_dev, _Complex double k, qpms_bessel_t btyp,
		const double *cvf_dev, size_t npts, const double *where_dev, double *out_dev, int parity, void *stream);

/* ------------------------------------------------------------------------ */
/* Beyn contour eigensolver — beyn.h:11-126                                  */
/* ------------------------------------------------------------------------ */
typedef int (*beyn_function_M_t)(_Complex double *target_M, size_t m, _Complex double z, void *params);
typedef int (*beyn_function_M_inv_Vhat_t)(_Complex double *target_M_inv_Vhat, size_t m, size_t l,
		const _Complex double *Vhat, _Complex double z, void *params);

/* beyn.h:30-43 */
typedef struct beyn_contour_t {
	size_t n;
	_Complex double centre;
	bool (*inside_test)(struct beyn_contour_t *, _Complex double z);
	_Complex double z_dz[][2];
} beyn_contour_t;

typedef enum {                                     /* beyn.h:49-54 */
	BEYN_CONTOUR_HALFELLIPSE_RE_PLUS = 3,
	BEYN_CONTOUR_HALFELLIPSE_RE_MINUS = 1,
	BEYN_CONTOUR_HALFELLIPSE_IM_PLUS = 0,
	BEYN_CONTOUR_HALFELLIPSE_IM_MINUS = 2
} beyn_contour_halfellipse_orientation;

beyn_contour_t *beyn_contour_ellipse(_Complex double centre, double halfax_re, double halfax_im, size_t npoints); /* beyn.c:69-86 */
/* beyn.c:120-163: half-ellipse; n/2 nodes on the straight cut, the rest on the arc. */
beyn_contour_t *beyn_contour_halfellipse(_Complex double centre, double halfax_re, double halfax_im, size_t npoints,
		beyn_contour_halfellipse_orientation orientation);
/* beyn.c:165-248: half-ellipse with rounded corners, rounding in [0, 0.5). */
beyn_contour_t *beyn_contour_kidney(_Complex double centre, double halfax_re, double halfax_im,
		double rounding, size_t n, beyn_contour_halfellipse_orientation orientation);

/* beyn.h:82-94; `gsl` is this library's private bookkeeping pointer. */
typedef struct beyn_result_t {
	size_t neig;
	size_t vlen;
	_Complex double *eigval;
	_Complex double *eigval_err;
	double *residuals;
	_Complex double *eigvec;
	double *ranktest_SV;
	void *gsl;
} beyn_result_t;
void beyn_result_free(beyn_result_t *result);                               /* beyn.c:662-666 */

/* beyn.c:638-646. V̂ is drawn from a generator seeded with the
 * QPMS_HIP_BEYN_SEED environment variable (default 1) instead of time(NULL). */
beyn_result_t *beyn_solve(size_t m, size_t l, beyn_function_M_t M, beyn_function_M_inv_Vhat_t M_inv_Vhat,
		void *params, const beyn_contour_t *contour, double rank_tol, size_t rank_min_sel, double res_tol);

/* scatsystem.c:2301-2325 (full system only: iri must be QPMS_NO_IRREP). */
struct beyn_result_t *qpms_scatsys_finite_find_eigenmodes(const qpms_scatsys_t *ss, qpms_iri_t iri,
		_Complex double omega_centre, double omega_rr, double omega_ri, size_t contour_npoints,
		double rank_tol, size_t rank_sel_min, double res_tol);
/* scatsystem.c:2351-2370. Like the finite wrapper, uses l = min(m, 64) probe
 * columns (QPMS_HIP_BEYN_L) instead of l = m. */
struct beyn_result_t *qpms_scatsys_periodic_find_eigenmodes(const qpms_scatsys_t *ss, const double k[3],
		_Complex double omega_centre, double omega_rr, double omega_ri, size_t contour_npoints,
		double rank_tol, size_t rank_sel_min, double res_tol);

/* ======================================================================== */
/* Extensions (no counterpart in the reference)                              */
/* ======================================================================== */

/* Library status. Returns 0 when a HIP device is usable; otherwise prints the
 * HIP error and returns it. Every compute entry aborts when no device is
 * present: there is no CPU fallback. */
int qpms_hip_init(void);
const char *qpms_hip_version(void);
/* Synchronise the library's stream. */
int qpms_hip_sync(void);
/* The library's own hipStream_t (the one NULL stream arguments select). */
void *qpms_hip_stream(void);
/* Device memory helpers for the *_dev entry points (bytes). */
void *qpms_hip_malloc(size_t bytes);
void qpms_hip_free(void *dptr);
int qpms_hip_memcpy_h2d(void *dst, const void *src, size_t bytes);
int qpms_hip_memcpy_d2h(void *dst, const void *src, size_t bytes);

/* Batched AB arrays: for each point i, A and B (nelem × nelem, row = dest
 * y(m,n), column = source y(μ,ν)) are written at A_dev + i*nelem*nelem.
 * kr_dev holds npts complex kr values, theta_dev/phi_dev npts doubles.
 * Same semantics per point as qpms_trans_calculator_get_AB_arrays with
 * deststride = nelem, srcstride = 1. */
int qpms_trans_calculator_get_AB_arrays_batch_dev(const qpms_trans_calculator *c,
		size_t npts, const _Complex double *kr_dev, const double *theta_dev, const double *phi_dev,
		int r_ge_d, qpms_bessel_t J, _Complex double *A_dev, _Complex double *B_dev, void *stream);

/* Full translation matrix of a finite cluster on the GPU: positions (npart ×
 * cart3_t, device), all particles with the default bspec of lMax (E block
 * then M block), W written row-major (ld = 2*nelem*npart) to W_dev. */
int qpms_hip_build_translation_matrix_dev(const qpms_trans_calculator *c, size_t npart,
		const cart3_t *pos_dev, _Complex double k, qpms_bessel_t J, _Complex double *W_dev, void *stream);

/* Rows [row_p0, row_p0 + row_np) of particles of the same matrix, written to
 * W_dev with leading dimension ld (≥ 2*nelem*npart). */
int qpms_hip_build_translation_matrix_rows_dev(const qpms_trans_calculator *c, size_t npart,
		const cart3_t *pos_dev, _Complex double k, qpms_bessel_t J,
		size_t row_p0, size_t row_np, _Complex double *W_dev, size_t ld, void *stream);

/* Mode-problem matrix M = I − T W for a finite cluster of identical
 * particles with diagonal T (tdiag_dev: nb complex per T type, type index per
 * particle in ttype_dev), all on the device. */
int qpms_hip_build_modeproblem_matrix_dev(const qpms_trans_calculator *c, size_t npart,
		const cart3_t *pos_dev, const int *ttype_dev, const _Complex double *tdiag_dev,
		_Complex double k, _Complex double *M_dev, void *stream);

/* Dense FP64 LU with partial pivoting, row-major n × n, in place on the
 * device; ipiv_dev receives 1-based pivots (int). Returns 0 or the first
 * zero pivot (1-based), like zgetrf's info. The pivots are zgetrf's: pivot j
 * is the first row of largest |re| + |im| in column j at or below the
 * diagonal (replaces LAPACKE_zgetrf at scatsystem.c:2214). */
int qpms_hip_zgetrf_dev(size_t n, _Complex double *A_dev, int *ipiv_dev, void *stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under graph capture, serving loops that forbid hipMalloc):
 * qpms_hip_zgetrf_dev without any hipMalloc/hipFree or allocating library
 * call. Its scratch (info word and strip workspace) is a static
 * 16 MiB device array of the library, so the order is limited to
 * qpms_hip_zgetrf_noalloc_max_n() (125,952); a larger n returns -1
 * with a message on stderr and touches nothing. Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other host threads, or
 * other streams before this call's synchronisation) race. Same results as
 * qpms_hip_zgetrf_dev up to the rounding of the trailing GEMM (hand-written
 * MFMA kernel here). */
int qpms_hip_zgetrf_dev_noalloc(size_t n, _Complex double *A_dev, int *ipiv_dev, void *stream);
size_t qpms_hip_zgetrf_noalloc_max_n(void);
/* Solve with nrhs right-hand sides stored row-major (n × nrhs). */
int qpms_hip_zgetrs_dev(size_t n, size_t nrhs, const _Complex double *LU_dev, const int *ipiv_dev,
		_Complex double *B_dev, void *stream);

/* Autotuner of the FP32 LU's trailing update (no reference counterpart):
 * times the library GEMM and the hand-written kernels (3M MFMA kernel in
 * three tile shapes, four-product kernel) per size class on scratch operands
 * (allocated and freed here) and records the fastest for every later LU in
 * the process; the single-matrix classes of order 2048 to 16384 are then
 * decided by timing whole FP32 factorisations of that order (on a stream of
 * the tuner's own), with the class's updates switched between the library
 * and the two fastest 3M tiles. A hand-written variant within 3 % of the
 * library is preferred. Returns the number of GEMM-timed classes; `report`
 * (optional, 2 x 6 x 5 floats) receives the milliseconds
 * [batched][class][variant]. The serving loop runs it at creation when
 * QPMS_HIP_AUTOTUNE is set. */
int qpms_hip_autotune(float *report);
/* The tuned table as text ("batched class-dim variant" per line). */
int qpms_hip_autotune_table(char *buf, size_t len);

/* Batched Ewald sums: ntrip (beta, k, s) triples on one 2D lattice in the xy
 * plane. triples_dev holds 8 doubles per triple {beta_x, beta_y, beta_z,
 * Re k, Im k, s_x, s_y, s_z}; R_dev nR direct-lattice points (x, y, z), the
 * triple's s subtracted in-kernel (a point with R = s contributes nothing, so
 * one list including the origin serves s = 0 and s != 0 alike); K_dev nK reciprocal points (x, y).
 * Outputs (either may be NULL) hold c->nelem_sc complex values per triple,
 * the same quantities as ewald3_sigma_short / ewald3_sigma_long. */
int qpms_hip_ewald3_sigma_batch_dev(const qpms_ewald3_constants_t *c, double eta, double unitcell_area,
		size_t ntrip, const double *triples_dev, size_t nR, const double *R_dev, size_t nK, const double *K_dev,
		_Complex double *sigma_short_dev, _Complex double *sigma_long_dev, void *stream);

/* Mixed-precision solve: X = A⁻¹ B with an FP32-complex LU on MFMA and FP64
 * iterative refinement until ‖B − A X‖ / (‖A‖‖X‖) ≤ tol or maxit steps.
 * A_dev is left intact. Returns the number of refinement steps taken, or −1
 * if refinement stalled (then X holds the FP64-LU solution). */
int qpms_hip_solve_mixed_dev(size_t n, size_t nrhs, const _Complex double *A_dev,
		_Complex double *B_dev, _Complex double *X_dev, double tol, int maxit, void *stream);

/* Beyn with an injected probe matrix V̂ (m × l row-major, host). */
beyn_result_t *beyn_solve_v(size_t m, size_t l, beyn_function_M_t M, beyn_function_M_inv_Vhat_t M_inv_Vhat,
		void *params, const beyn_contour_t *contour, double rank_tol, size_t rank_min_sel,
		double res_tol, const _Complex double *Vhat);

/* ------------------------------------------------------------------------ */
/* Measurement probes (no reference counterpart): the measured HBM bandwidth */
/* and FP64 vector throughput that roofline fractions are compared with.     */
/* ------------------------------------------------------------------------ */
/* STREAM triad a = b + s c over n doubles (n even; 24 n bytes), device pointers. */
int qpms_hip_probe_triad(size_t n, double *a, const double *b, const double *c, double s, void *stream);
/* Write-only stream a = s over n doubles (n even; 8 n bytes), device pointer. */
int qpms_hip_probe_fill(size_t n, double *a, double s, void *stream);
/* blocks x 256 lanes of 8 independent FP64 FMA chains for iters steps
 * (16 iters flops per lane); out_dev holds blocks * 256 doubles. */
int qpms_hip_probe_fma64(int blocks, int iters, double *out_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* QPMS_HIP_H */
