/* Serving loop for multiple-scattering requests on one MI355X.
 *
 * No counterpart in the reference: the reference answers one request at a
 * time through the scattering-system API (scatsystem.h). This interface
 * takes a stream of independent requests of differing cluster size, lMax,
 * frequency and Bloch vector, keeps them resident in HBM from build to
 * result, and works on several at once:
 *
 *  - admission by HBM footprint: a request enters when its matrices fit in
 *    the arena that the server allocated at creation, and leaves it when its
 *    result is copied out;
 *  - continuous batching: small plane-wave solves admitted together are
 *    padded to a common size and factorised by one batched launch sequence;
 *    larger requests run on their own streams, overlapping each other;
 *  - precision: plane-wave solves and Beyn node solves factorise in
 *    FP32-complex on MFMA and are refined to FP64 accuracy (FP64 residuals);
 *    determinants factorise in FP64. Matrix builds, Ewald sums and the Beyn
 *    SVD / eigenproblem are FP64 throughout.
 *
 * Threading: one server is driven by one host thread (submit / progress /
 * poll are not re-entrant on the same server).
 */
#ifndef QPMS_HIP_SERVE_H
#define QPMS_HIP_SERVE_H
#include "qpms_hip.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
	QPMS_HIP_REQ_SCATTER = 0,      /* plane-wave solve on a finite cluster -> extinction */
	QPMS_HIP_REQ_PERIODIC_DET = 1, /* log det of the 2D-periodic mode problem at (omega, k) */
	QPMS_HIP_REQ_BEYN = 2          /* finite-cluster eigenfrequencies inside an ellipse */
} qpms_hip_req_kind;

#define QPMS_HIP_SERVE_MAX_EIG 64

/* One request: identical spheres (radius, eps_in) in a constant background
 * eps_bg (mu = 1 everywhere), default VSWF basis of lMax (1..4). */
typedef struct qpms_hip_request_t {
	qpms_hip_req_kind kind;
	int lMax;
	size_t npart;
	const double *pos;          /* npart x (x, y, z), metres; copied at submit */
	double radius;
	_Complex double eps_in;
	double eps_bg;
	double omega;               /* rad/s; BEYN: real part of the contour centre */
	/* SCATTER: plane wave e^{i k.r} E with k = k_bg * kdir (unit vector) */
	double kdir[3];
	_Complex double E[3];
	/* PERIODIC_DET: the particles form the unit cell of the lattice
	 * {lattice[0], lattice[1]} in the xy plane; Bloch vector beta (1/m) */
	double lattice[2][2];
	double beta[2];
	/* BEYN: ellipse half-axes (rad/s), nodes, probe columns, probe seed */
	double beyn_rRe, beyn_rIm;
	int beyn_nodes;
	int beyn_l;
	unsigned long long beyn_seed;
	double rank_tol;            /* BEYN: absolute singular-value cut (beyn.c:382); no forced minimum rank */
	double beyn_centre_im;      /* BEYN: imaginary part of the contour centre (rad/s) */
} qpms_hip_request_t;

typedef struct qpms_hip_response_t {
	unsigned long long id;
	int status;                 /* 0 ok; < 0 error (see message) */
	qpms_hip_req_kind kind;
	size_t dim;                 /* mode-problem dimension L */
	double sigma_ext;           /* SCATTER: -Re(a^H f) / k^2 */
	double rel_residual;        /* SCATTER: ||T a - M f|| / (||M|| ||f||) after refinement */
	int refine_steps;
	double logabsdet;           /* PERIODIC_DET: log |det M| */
	double argdet;              /* PERIODIC_DET: arg det M in (-pi, pi] */
	int neig;                   /* BEYN: eigenvalues whose singular value exceeds rank_tol */
	_Complex double eigval[QPMS_HIP_SERVE_MAX_EIG];
	double sv[QPMS_HIP_SERVE_MAX_EIG];  /* BEYN: singular values of A0 (beyn_l of them) */
	double t_submit, t_start, t_done;   /* seconds on the server's clock */
	char message[96];
	int batch;                  /* requests factorised together in this request's unit */
	int padded_dim;             /* order the request was padded to (batched solves) */
} qpms_hip_response_t;

typedef struct qpms_hip_server_opts_t {
	size_t arena_bytes;         /* device arena for resident requests (0: the device's free memory at
	                               creation less max(8 GiB, 1/16 of the device), rounded down to GiB) */
	int nstreams;               /* concurrent streams, at most 8 (0: 8) */
	int mixed;                  /* 1: FP32 LU + FP64 refinement for solves; 0: FP64 LU */
	size_t batch_max_dim;       /* plane-wave solves with L <= this are batched (0: 8192) */
	int batch_max;              /* most requests per batch (0: 32) */
	int refine_steps;           /* fixed refinement steps of a mixed solve (0: 3) */
} qpms_hip_server_opts_t;

typedef struct qpms_hip_server_t qpms_hip_server_t;

/* Returns NULL (after a message on stderr) when there is no device or the
 * arena or the pinned staging pool cannot be allocated; nothing is left
 * allocated then. */
qpms_hip_server_t *qpms_hip_server_create(const qpms_hip_server_opts_t *opts);
/* Waits for the work in flight, then releases every device resource the
 * server holds: arena, staging pool, lanes and their per-stream LU state. */
void qpms_hip_server_destroy(qpms_hip_server_t *srv);
/* Size of the server's device arena in bytes. */
size_t qpms_hip_server_arena_bytes(const qpms_hip_server_t *srv);
/* Seconds since the server was created (the clock of t_submit/t_start/t_done). */
double qpms_hip_server_clock(const qpms_hip_server_t *srv);
/* Queue a request (host data copied); returns its id (> 0) or 0 if it can
 * never fit in the arena. Does not touch the GPU. */
unsigned long long qpms_hip_server_submit(qpms_hip_server_t *srv, const qpms_hip_request_t *req);
/* Admit queued requests that fit, launch their work, retire finished work.
 * Returns the number of responses ready for qpms_hip_server_poll. With
 * wait != 0, blocks until at least one in-flight request has finished (if
 * any is in flight). */
int qpms_hip_server_progress(qpms_hip_server_t *srv, int wait);
/* Move up to max finished responses to out; returns how many. */
int qpms_hip_server_poll(qpms_hip_server_t *srv, qpms_hip_response_t *out, int max);
/* Requests queued or in flight. */
size_t qpms_hip_server_pending(const qpms_hip_server_t *srv);
/* Stop: queued requests are answered with status -2 at once; Beyn contours
 * in flight stop after their current chunk of nodes (status -2); other work
 * in flight completes normally. Keep calling progress until pending is 0. */
void qpms_hip_server_cancel(qpms_hip_server_t *srv);
/* Start no request after time t (server clock, seconds) unless its estimated
 * device time fits before t (0: no deadline). Requests already running are
 * not affected; those that do not fit stay queued. */
void qpms_hip_server_set_deadline(qpms_hip_server_t *srv, double t);

/* The same request answered through the public scattering-system API one
 * call after another, with host buffers and FP64 LU (the plain drop-in path
 * the serving loop is measured against). */
int qpms_hip_request_plain(const qpms_hip_request_t *req, qpms_hip_response_t *out);

#ifdef __cplusplus
}
#endif
#endif /* QPMS_HIP_SERVE_H */
