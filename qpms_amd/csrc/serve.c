/* Serving loop: admission by HBM footprint, continuous batching of small
 * plane-wave solves, per-request streams for the rest, FP32-LU + FP64
 * refinement for solves. Declarations and semantics: qpms_hip_serve.h.
 *
 * Everything a request needs on the device comes from one arena allocated at
 * creation (first-fit, coalescing), so the loop itself never calls hipMalloc
 * or hipFree; inputs and results pass through pinned staging slots.
 *
 * A request's life: queued (host copy) -> admitted into a work unit (arena
 * block reserved, inputs staged, kernels enqueued on one of the server's
 * streams, an event recorded) -> retired when the event has fired (results
 * read from pinned memory, arena block released). Plane-wave solves with
 * L <= batch_max_dim that are queued together are padded to a common order
 * (identity rows / columns, zero right-hand side) and go through one batched
 * factorisation; the padded system's solution restricted to the first L
 * entries is the unpadded one.
 */
#define _GNU_SOURCE   /* qsort_r */
#include "qpms_internal.h"
#include "../../include/qpms_hip_serve.h"
#include <hip/hip_runtime_api.h>
#include <complex.h>
#include <math.h>
#include <string.h>
#include <time.h>

#define SRV_HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
	QH_DIE("HIP error %d (%s) in %s", (int)e_, hipGetErrorString(e_), #x); } while (0)

#define EV_RAD_PER_S (1.602176634e-19 / 1.054571817e-34)
#define MAX_LANES 8
#define MAX_BATCH 64
#define MAX_UNITS 64

/* launchers from the HIP sources */
int qh_launch_planewave(int lMax, size_t npart, const double *pos_dev, const double kvec[3],
		const double *E_dev, double *out_dev, long ld, void *stream);
int qh_launch_neg_re_cdot(int batch, long n, const double *a_dev, long sa, const double *f_dev, long sf,
		double *out_dev, void *stream);
int qh_launch_expand_tdiag(int lMax, int nsets, const double *R, double *td, void *stream);
int qh_launch_tdiag_mul(int batch, int L, int nb, int npad, const double *td, long std_, const double *a,
		long sa, double *out, long so, void *stream);
int qh_launch_pad_identity(int batch, int L, int npad, double *M, long sM, void *stream);
int qh_launch_sumsq_batched(int batch, long n, const double *x, long sx, double *part, double *out, void *stream);
#define QH_SUMSQ_PARTS 512
int qh_launch_logdet(int n, const double *LU, const int *ipiv, double *out, void *stream);
int qh_launch_probe(long n, unsigned long long seed, double *V, void *stream);
void qh_stream_release(void *stream);
void qh_probe_host(long n, unsigned long long seed, double *V);
int qh_launch_pair_triples(int np, const double *pos, double bx, double by, double kre, double kim,
		double *tr, void *stream);
int qh_launch_pair_sigma_total(int np, int nsc, const double *ss, const double *sl, const double *s0,
		double *tot, void *stream);
int qh_launch_e32_assemble_M(const qh_trans_tables *t, int np, int nsc, const double *sig, int bq_mode,
		const int *ttype, const double *tdiag, double *M, long ld, void *stream);
int qh_launch_sigma0(int n, const double *k_dev, double eta, double leg00, double *out_dev, void *stream);
size_t qh_getrf_ws_bytes(int prec, int n, int batch);
int qh_getrf_batched_ws_f64(int n, int batch, double *A, long sA, int *ipiv, int *perm, int *info, void *ws,
		void *stream);
int qh_getrf_batched_ws_f32(int n, int batch, float *A, long sA, int *ipiv, int *perm, int *info, void *ws,
		void *stream);
int qh_getrf_batched_wsla_f32(int n, int batch, float *A, long sA, int *ipiv, int *perm, int *info, void *ws,
		int lookahead, void *stream);
int qh_getrf_batched_wsla_f64(int n, int batch, double *A, long sA, int *ipiv, int *perm, int *info, void *ws,
		int lookahead, void *stream);
/* Orders from which a served single-matrix LU uses the lookahead (the bulk
 * of each trailing update on the device's side stream beside the next
 * block's strips), and up to which Beyn chunks are replayed from a captured
 * graph (launch-bound orders; a capture excludes the library GEMM and the
 * lookahead, which cost more than the launches from order ~4096 on: a
 * 96-node contour of order 9216 took 4.7 s captured against 4.0 s eager). */
#define SRV_LOOKAHEAD_MIN 4096
#define SRV_GRAPH_MAX 4096
/* Beyn contours of this order and above take one node per chunk: a single
 * LU with the lookahead beats a batch without it there (two LUs of order
 * 38100: 2.61 s singly against 2.70 s as a batch of two; at order 9210 the
 * batch wins, 101 against 146 ms). */
#define SRV_BEYN_SINGLE_MIN 16384
int qh_beyn_accum(long n, const double *X, double dzr, double dzi, double wr, double wi, int coarse,
		double *A0, double *A1, double *A0c, double *A1c, void *stream);
int qh_beyn_process_s(size_t m, size_t l, const complex double *dA0, const complex double *dA1, complex double z0,
		double rank_tol, size_t rank_sel_min, complex double *eigval, double *sigma, void *stream, void *ws);
size_t qh_beyn_ws_bytes(size_t m, size_t l);
int qh_e32_bq_mode(void);
double l2d_unitcell_area(cart2_t b1, cart2_t b2);
double *qh_collect(PGen *g, size_t *n, int dims);   /* ewald.c: PGen points as (x, y[, z]) */
int l2d_reciprocalBasis2pi(cart2_t b1, cart2_t b2, cart2_t *rb1, cart2_t *rb2);

/* ---------------------------------------------------------------------- */
/* Arena: first-fit over a sorted list of blocks                            */
/* ---------------------------------------------------------------------- */
typedef struct { size_t off, size; int used; } qh_blk;
typedef struct {
	char *base;
	size_t size;
	qh_blk *b;
	int nb, cap;
	size_t in_use;
} qh_arena;

/* Returns 0, or the HIP error of the allocation (nothing is left allocated). */
static int arena_init(qh_arena *a, size_t size, int pinned_host) {
	memset(a, 0, sizeof(*a));
	const hipError_t e = pinned_host ? hipHostMalloc((void **)&a->base, size, 0) : hipMalloc((void **)&a->base, size);
	if (e != hipSuccess) {
		(void)hipGetLastError();   /* clear the sticky error of the failed allocation */
		a->base = NULL;
		return (int)e;
	}
	a->size = size;
	a->cap = 64;
	QH_MALLOC(a->b, a->cap * sizeof(qh_blk));
	a->b[0] = (qh_blk){0, size, 0};
	a->nb = 1;
	a->in_use = 0;
	return 0;
}

static long arena_alloc(qh_arena *a, size_t bytes) {
	bytes = (bytes + 4095) & ~(size_t)4095;
	for (int i = 0; i < a->nb; ++i) {
		if (a->b[i].used || a->b[i].size < bytes) continue;
		const size_t off = a->b[i].off;
		if (a->b[i].size > bytes) {
			if (a->nb == a->cap) { a->cap *= 2; a->b = realloc(a->b, a->cap * sizeof(qh_blk)); QH_ENSURE(a->b, "oom"); }
			memmove(&a->b[i + 2], &a->b[i + 1], (a->nb - i - 1) * sizeof(qh_blk));
			a->b[i + 1] = (qh_blk){off + bytes, a->b[i].size - bytes, 0};
			a->b[i].size = bytes;
			++a->nb;
		}
		a->b[i].used = 1;
		a->in_use += bytes;
		return (long)off;
	}
	return -1;
}

static void arena_free(qh_arena *a, long off) {
	for (int i = 0; i < a->nb; ++i) {
		if ((long)a->b[i].off != off) continue;
		QH_ENSURE(a->b[i].used, "arena: double free");
		a->b[i].used = 0;
		a->in_use -= a->b[i].size;
		if (i + 1 < a->nb && !a->b[i + 1].used) {
			a->b[i].size += a->b[i + 1].size;
			memmove(&a->b[i + 1], &a->b[i + 2], (a->nb - i - 2) * sizeof(qh_blk));
			--a->nb;
		}
		if (i > 0 && !a->b[i - 1].used) {
			a->b[i - 1].size += a->b[i].size;
			memmove(&a->b[i], &a->b[i + 1], (a->nb - i - 1) * sizeof(qh_blk));
			--a->nb;
		}
		return;
	}
	QH_DIE("arena: unknown block %ld", off);
}

/* Bump allocator over one arena block. */
typedef struct { char *p; size_t left; } qh_bump;
static void *bump(qh_bump *b, size_t bytes) {
	bytes = (bytes + 255) & ~(size_t)255;
	QH_ENSURE(bytes <= b->left, "work-unit layout exceeds its reservation");
	void *r = b->p;
	b->p += bytes;
	b->left -= bytes;
	return r;
}

/* ---------------------------------------------------------------------- */
/* Jobs and work units                                                      */
/* ---------------------------------------------------------------------- */
typedef struct qh_job {
	unsigned long long id;
	qpms_hip_request_t req;
	double *pos;           /* owned copy */
	size_t L, nb;
	double t_submit;
	double est;            /* estimated device seconds (scheduling only) */
	struct qh_job *next;
} qh_job;

enum { U_SCATTER = 0, U_PERIODIC = 1, U_BEYN = 2 };

typedef struct {
	int kind, njobs, lane;
	qh_job *jobs[MAX_BATCH];
	long arena_off;
	size_t npad;
	hipEvent_t done;
	char *pin;             /* pinned results (and staged inputs), from the pinned pool */
	long pin_off;
	size_t pin_bytes;
	double t_start;
	/* Beyn: device moment matrices, processed after the last chunk; the
	 * node chunks are enqueued one at a time from the progress loop */
	complex double *dA0, *dA1, *dA0c, *dA1c;
	int chunk, next_node, cancelled;
	double *bV, *bM, *bX, *bR, *bW, *btd, *bpos;
	float *bM32, *bR32, *bW32;
	int *bipiv, *bperm, *binfo, *btt;
	void *bws, *btail;
	beyn_contour_t *bcont;
	double *btab;          /* node table: k, dz, z - z0, coarse flag (8 doubles per node) */
	int *bcur;             /* device cursor: first node of the chunk being run */
	hipGraphExec_t gexec;  /* one chunk's launches, captured on the first chunk */
	int active;
} qh_unit;

struct qpms_hip_server_t {
	qpms_hip_server_opts_t o;
	hipStream_t lane[MAX_LANES];
	int lane_busy[MAX_LANES];
	qh_arena arena;
	qh_arena pinned;       /* host staging pool (hipHostMalloc once) */
	qpms_trans_calculator *calc[5];
	qh_job *qhead, *qtail;
	size_t nqueued;
	qh_unit unit[MAX_UNITS];
	int ninflight;
	qpms_hip_response_t *ready;
	int nready, capready;
	qh_job **order;        /* launch-order scratch */
	double deadline;       /* 0, or: start only what is estimated to finish by then */
	int graphs;            /* Beyn chunks replayed from a captured hipGraph (QPMS_HIP_GRAPHS != 0) */
	int beyn_chunk_max;    /* most nodes per chunk (QPMS_HIP_BEYN_CHUNK, 0: no limit) */
	size_t order_cap;
	unsigned long long next_id;
	struct timespec t0;
};

double qpms_hip_server_clock(const qpms_hip_server_t *s) {
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (t.tv_sec - s->t0.tv_sec) + 1e-9 * (t.tv_nsec - s->t0.tv_nsec);
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

/* Dimension buckets of the batched path (order of the padded matrices). */
/* Padded orders of batched solves. Coarse on purpose: up to a few thousand
 * the LU is bound by its chain of small launches, not by flops, so padding
 * costs little and a coarser bucket gathers a larger batch. */
static size_t pad_bucket(size_t L) {
	static const size_t b[] = {256, 512, 1024, 1536, 2048, 3072, 4096, 6144, 8192};
	for (size_t i = 0; i < sizeof b / sizeof b[0]; ++i) if (L <= b[i]) return b[i];
	return (L + 255) & ~(size_t)255;
}

/* Device bytes of a plane-wave-solve unit of `batch` systems of order npad. */
static size_t scatter_bytes(const qpms_hip_server_t *s, size_t npad, int batch, size_t max_np) {
	const size_t nn = npad * npad;
	size_t b = 0;
	b += align256(batch * nn * 16);                    /* M (FP64) */
	if (s->o.mixed) b += align256(batch * nn * 8);     /* M (FP32) */
	b += 5 * align256(batch * npad * 16);              /* a, b, x, r, work */
	b += 2 * align256(batch * npad * 8);               /* r32, w32 */
	b += 2 * align256(batch * npad * 4) + align256(batch * 4);
	b += align256(qh_getrf_ws_bytes(s->o.mixed ? 32 : 64, (int)npad, batch));
	b += align256(batch * QH_SUMSQ_PARTS * 8);         /* norm partial sums */
	/* per job: staged inputs, Mie output, T diagonal, type indices */
	b += batch * (align256(max_np * 24 + 128) + align256(10 * 16) + align256(48 * 16) + align256(max_np * 4));
	b += align256(batch * 8 * 8);
	return b + 65536;
}

static size_t periodic_bytes(size_t L, size_t np, int nsc, size_t nR, size_t nK) {
	const size_t npairs = np * np;
	size_t b = align256(L * L * 16) + align256(L * 8) + align256(qh_getrf_ws_bytes(64, (int)L, 1));
	b += align256(npairs * 64) + 3 * align256(npairs * nsc * 16);
	b += align256(nR * 24) + align256(nK * 16) + align256(np * 24) + align256(np * 4) + 4 * align256(48 * 16);
	b += align256(nK * 55 * 16);   /* per-point long-range terms, up to (LE+1)(LE+2)/2 = 55 each */
	return b + 65536;
}

static size_t beyn_node_bytes(const qpms_hip_server_t *s, size_t L, size_t l) {
	return align256(L * L * 16) + (s->o.mixed ? align256(L * L * 8) : 0) + 4 * align256(L * l * 16)
		+ 2 * align256(L * l * 8) + 2 * align256(L * 4) + 1024;
}
/* A contour of N nodes handled `chunk` nodes at a time: per-chunk matrices and
 * workspaces, the probe and moment blocks, and per-node Mie parameters and
 * T diagonals for all N nodes (lMax <= 4: nb <= 48). */
static size_t beyn_bytes(const qpms_hip_server_t *s, size_t L, size_t l, int chunk, int N, int lMax) {
	const size_t nb = 2 * (size_t)lMax * (lMax + 2);
	return chunk * beyn_node_bytes(s, L, l) + align256(qh_getrf_ws_bytes(s->o.mixed ? 32 : 64, (int)L, chunk))
		+ 5 * align256(L * l * 16) + align256(L * 24) + align256(qh_beyn_ws_bytes(L, l))
		+ align256(N * 9 * 8) + align256(N * 2 * (lMax + 1) * 16) + align256(N * nb * 16) + align256(N * 64) + 256
		+ 65536 + chunk * 4096;
}

static const qpms_trans_calculator *srv_calc(qpms_hip_server_t *s, int lMax) {
	QH_ENSURE(lMax >= 1 && lMax <= 4, "lMax %d outside 1..4", lMax);
	return s->calc[lMax];
}

/* Device memory a server leaves to the rest of the process when it sizes its
 * arena itself: the larger of 8 GiB and 1/16 of the device. */
static size_t auto_headroom(size_t total) {
	const size_t h = total / 16, floor_ = (size_t)8 << 30;
	return h > floor_ ? h : floor_;
}

static void server_release(qpms_hip_server_t *s) {
	for (int i = 0; i < s->o.nstreams; ++i)
		if (s->lane[i]) {
			SRV_HIPCHK(hipStreamSynchronize(s->lane[i]));
			qh_stream_release(s->lane[i]);
			SRV_HIPCHK(hipStreamDestroy(s->lane[i]));
		}
	if (s->arena.base) SRV_HIPCHK(hipFree(s->arena.base));
	free(s->arena.b);
	if (s->pinned.base) SRV_HIPCHK(hipHostFree(s->pinned.base));
	free(s->pinned.b);
	for (int l = 1; l <= 4; ++l) if (s->calc[l]) qpms_trans_calculator_free(s->calc[l]);
	free(s->ready);
	free(s->order);
	free(s);
}

qpms_hip_server_t *qpms_hip_server_create(const qpms_hip_server_opts_t *opts) {
	if (qpms_hip_init() != 0) {
		fprintf(stderr, "qpms_hip_server_create: no HIP device\n");
		return NULL;
	}
	static int tuned = 0;   /* the GEMM autotuner runs once per process, before the arena takes the memory */
	const char *et = getenv("QPMS_HIP_AUTOTUNE");
	if (!tuned && et && strcmp(et, "0")) {
		qpms_hip_autotune(NULL);
		tuned = 1;
	}
	qpms_hip_server_t *s;
	QH_CALLOC(s, 1, sizeof(*s));
	if (opts) s->o = *opts;
	size_t dfree = 0, dtotal = 0;
	SRV_HIPCHK(hipMemGetInfo(&dfree, &dtotal));
	if (!s->o.arena_bytes) {
		const size_t head = auto_headroom(dtotal);
		if (dfree <= head + ((size_t)1 << 30)) {
			fprintf(stderr, "qpms_hip_server_create: %zu MiB free on the device, below the %zu MiB headroom\n",
				dfree >> 20, head >> 20);
			free(s);
			return NULL;
		}
		s->o.arena_bytes = (dfree - head) & ~(((size_t)1 << 30) - 1);
	}
	if (s->o.nstreams <= 0) s->o.nstreams = 8;
	if (s->o.nstreams > MAX_LANES) s->o.nstreams = MAX_LANES;
	if (!s->o.batch_max_dim) s->o.batch_max_dim = 8192;
	if (s->o.batch_max <= 0) s->o.batch_max = 32;
	if (s->o.batch_max > MAX_BATCH) s->o.batch_max = MAX_BATCH;
	if (s->o.refine_steps <= 0) s->o.refine_steps = 3;
	for (int i = 0; i < s->o.nstreams; ++i) SRV_HIPCHK(hipStreamCreateWithFlags(&s->lane[i], hipStreamNonBlocking));
	int e = arena_init(&s->arena, s->o.arena_bytes, 0);
	if (e) {
		fprintf(stderr, "qpms_hip_server_create: device arena of %zu MiB: HIP error %d (%s); %zu of %zu MiB free\n",
			s->o.arena_bytes >> 20, e, hipGetErrorString((hipError_t)e), dfree >> 20, dtotal >> 20);
		server_release(s);
		return NULL;
	}
	e = arena_init(&s->pinned, (size_t)64 << 20, 1);
	if (e) {
		fprintf(stderr, "qpms_hip_server_create: pinned staging pool: HIP error %d (%s)\n", e,
			hipGetErrorString((hipError_t)e));
		server_release(s);
		return NULL;
	}
	for (int l = 1; l <= 4; ++l) {
		s->calc[l] = qpms_trans_calculator_init(l, QPMS_NORMALISATION_DEFAULT);
		qh_tables_get(s->calc[l]);   /* device tables now, not inside the loop */
	}
	const char *eg = getenv("QPMS_HIP_GRAPHS"), *ec = getenv("QPMS_HIP_BEYN_CHUNK");
	s->graphs = !(eg && !strcmp(eg, "0"));
	s->beyn_chunk_max = ec ? atoi(ec) : 0;
	qh_lu_prepare();   /* kernel attributes set before any launch sequence is captured */
	s->next_id = 1;
	clock_gettime(CLOCK_MONOTONIC, &s->t0);
	return s;
}

size_t qpms_hip_server_arena_bytes(const qpms_hip_server_t *s) { return s->arena.size; }

void qpms_hip_server_destroy(qpms_hip_server_t *s) {
	if (!s) return;
	for (int i = 0; i < s->o.nstreams; ++i) SRV_HIPCHK(hipStreamSynchronize(s->lane[i]));
	for (int u = 0; u < MAX_UNITS; ++u)
		if (s->unit[u].active) {
			SRV_HIPCHK(hipEventDestroy(s->unit[u].done));
			if (s->unit[u].gexec) SRV_HIPCHK(hipGraphExecDestroy(s->unit[u].gexec));
			free(s->unit[u].bcont);
			for (int j = 0; j < s->unit[u].njobs; ++j) { free(s->unit[u].jobs[j]->pos); free(s->unit[u].jobs[j]); }
		}
	while (s->qhead) { qh_job *j = s->qhead; s->qhead = j->next; free(j->pos); free(j); }
	server_release(s);
}

static size_t job_L(const qpms_hip_request_t *r) { return r->npart * 2 * (size_t)qpms_lMax2nelem(r->lMax); }

static void periodic_counts(const qpms_hip_request_t *r, size_t *nR, size_t *nK);
size_t qh_ewald3_long_shifts_scratch(const qpms_ewald3_constants_t *c, size_t nK);
int qh_ewald3_sigma_long_shifts_dev(const qpms_ewald3_constants_t *c, double eta, double unitcell_area, size_t ntrip,
		const double *triples_dev, size_t nK, const double *K_dev, void *F_dev, _Complex double *sigma_long_dev,
		void *stream);
static size_t job_bytes(const qpms_hip_server_t *s, const qh_job *j, int *beyn_chunk) {
	const size_t L = j->L;
	switch (j->req.kind) {
		case QPMS_HIP_REQ_SCATTER:
			if (L <= s->o.batch_max_dim) return scatter_bytes(s, pad_bucket(L), 1, j->req.npart);
			return scatter_bytes(s, L, 1, j->req.npart);
		case QPMS_HIP_REQ_PERIODIC_DET: {
			const int LE = 2 * j->req.lMax + 1;
			size_t nR, nK;
			periodic_counts(&j->req, &nR, &nK);
			return periodic_bytes(L, j->req.npart, (LE + 1) * (LE + 1), nR, nK);
		}
		case QPMS_HIP_REQ_BEYN: {
			const size_t l = (size_t)j->req.beyn_l;
			int c = j->req.beyn_nodes;
			const int N = j->req.beyn_nodes, lm = j->req.lMax;
			/* the largest divisor of N that fits half the arena (one node from
			 * SRV_BEYN_SINGLE_MIN on): every chunk then has the same launch
			 * sequence (one captured graph serves them all) */
			while (c > 1 && (N % c || beyn_bytes(s, L, l, c, N, lm) > s->arena.size / 2 ||
					(s->beyn_chunk_max > 0 && c > s->beyn_chunk_max) || L >= SRV_BEYN_SINGLE_MIN)) --c;
			if (beyn_chunk) *beyn_chunk = c;
			return beyn_bytes(s, L, l, c, N, lm);
		}
	}
	return 0;
}

/* Rough device seconds of a job, for the launch order only: LU flops at
 * 50 (FP32) / 35 (FP64) TFLOP/s, the refinement or node solves, the
 * periodic lattice sums per particle pair, and a fixed launch cost. */
static double job_est(const qpms_hip_server_t *s, const qh_job *j) {
	const double L = (double)j->L, lu = 8.0 / 3.0 * L * L * L / (s->o.mixed ? 5e13 : 3.5e13);
	const double np = (double)j->req.npart;
	switch (j->req.kind) {
		case QPMS_HIP_REQ_SCATTER: return 2e-3 + lu + 8.0 * L * L * s->o.refine_steps / 5e12;
		case QPMS_HIP_REQ_PERIODIC_DET: {
			const double ne = (2.0 * j->req.lMax + 1) * (2.0 * j->req.lMax + 1);
			size_t nR, nK;
			periodic_counts(&j->req, &nR, &nK);
			return 5e-3 + 8.0 / 3.0 * L * L * L / 3.5e13 + np * np * (double)(nR + nK) * ne * 2e-12;
		}
		case QPMS_HIP_REQ_BEYN:
			return 5e-3 + j->req.beyn_nodes * (lu + 16.0 * L * L * j->req.beyn_l * s->o.refine_steps / 5e13);
	}
	return 0;
}

unsigned long long qpms_hip_server_submit(qpms_hip_server_t *s, const qpms_hip_request_t *r) {
	QH_ENSURE(r->lMax >= 1 && r->lMax <= 4 && r->npart >= 1, "invalid request (lMax %d, npart %zu)", r->lMax, r->npart);
	qh_job *j;
	QH_CALLOC(j, 1, sizeof(*j));
	j->req = *r;
	QH_MALLOC(j->pos, r->npart * 3 * sizeof(double));
	memcpy(j->pos, r->pos, r->npart * 3 * sizeof(double));
	j->req.pos = j->pos;
	if (j->req.kind == QPMS_HIP_REQ_BEYN) {
		if (j->req.beyn_l <= 0) j->req.beyn_l = 32;
		if (j->req.beyn_nodes <= 0) j->req.beyn_nodes = 32;
		if (j->req.rank_tol <= 0) j->req.rank_tol = 1e-4;
		if ((size_t)j->req.beyn_l > job_L(r)) j->req.beyn_l = (int)job_L(r);
		QH_ENSURE(j->req.beyn_l <= QPMS_HIP_SERVE_MAX_EIG, "beyn_l %d above %d", j->req.beyn_l, QPMS_HIP_SERVE_MAX_EIG);
	}
	j->L = job_L(r);
	j->nb = 2 * (size_t)qpms_lMax2nelem(r->lMax);
	j->t_submit = qpms_hip_server_clock(s);
	j->est = job_est(s, j);
	if (((job_bytes(s, j, NULL) + 4095) & ~(size_t)4095) > s->arena.size) { free(j->pos); free(j); return 0; }
	j->id = s->next_id++;
	if (s->qtail) s->qtail->next = j; else s->qhead = j;
	s->qtail = j;
	++s->nqueued;
	return j->id;
}

size_t qpms_hip_server_pending(const qpms_hip_server_t *s) { return s->nqueued + (size_t)s->ninflight; }

static void q_remove(qpms_hip_server_t *s, qh_job *prev, qh_job *j) {
	if (prev) prev->next = j->next; else s->qhead = j->next;
	if (s->qtail == j) s->qtail = prev;
	j->next = NULL;
	--s->nqueued;
}

static qh_unit *unit_new(qpms_hip_server_t *s) {
	for (int u = 0; u < MAX_UNITS; ++u) if (!s->unit[u].active) {
		qh_unit *x = &s->unit[u];
		memset(x, 0, sizeof(*x));
		x->active = 1;
		return x;
	}
	return NULL;
}

/* Pinned staging for a unit from the pool (waits for retirements if full). */
static void pin_take(qpms_hip_server_t *s, qh_unit *u) {
	long off = arena_alloc(&s->pinned, u->pin_bytes);
	QH_ENSURE(off >= 0, "pinned staging pool exhausted (%zu bytes asked)", u->pin_bytes);
	u->pin_off = off;
	u->pin = s->pinned.base + off;
}

static double wavenumber_re(double eps_bg, double omega) { return sqrt(eps_bg) * omega / SPEED_OF_LIGHT; }

/* Mie parameters of one sphere at complex omega into 9 doubles (k_mie layout). */
static void mie_params(const qpms_hip_request_t *r, complex double omega, double *p) {
	const complex double ki = csqrt(r->eps_in) * omega / SPEED_OF_LIGHT;
	const complex double ke = csqrt((complex double)r->eps_bg) * omega / SPEED_OF_LIGHT;
	p[0] = r->radius;
	p[1] = creal(ki); p[2] = cimag(ki);
	p[3] = creal(ke); p[4] = cimag(ke);
	p[5] = 1; p[6] = 0; p[7] = 1; p[8] = 0;
}

/* ---------------------------------------------------------------------- */
/* Plane-wave solves (batched when small)                                   */
/* ---------------------------------------------------------------------- */
static void launch_scatter(qpms_hip_server_t *s, qh_unit *u, hipStream_t st) {
	const int B = u->njobs;
	const size_t npad = u->npad, nn = npad * npad;
	size_t max_np = 0;
	for (int j = 0; j < B; ++j) if (u->jobs[j]->req.npart > max_np) max_np = u->jobs[j]->req.npart;
	qh_bump bp = {s->arena.base + u->arena_off, scatter_bytes(s, npad, B, max_np)};
	double *M = bump(&bp, B * nn * 16);
	float *M32 = s->o.mixed ? bump(&bp, B * nn * 8) : NULL;
	double *a = bump(&bp, B * npad * 16), *b = bump(&bp, B * npad * 16), *x = bump(&bp, B * npad * 16);
	double *r = bump(&bp, B * npad * 16), *work = bump(&bp, B * npad * 16);
	float *r32 = bump(&bp, B * npad * 8), *w32 = bump(&bp, B * npad * 8);
	int *ipiv = bump(&bp, B * npad * 4), *perm = bump(&bp, B * npad * 4), *info = bump(&bp, B * 4);
	void *ws = bump(&bp, qh_getrf_ws_bytes(s->o.mixed ? 32 : 64, (int)npad, B));
	double *outs = bump(&bp, B * 8 * 8);   /* per job: dot, |r|^2, |x|^2, |M|^2, info */
	double *part = bump(&bp, B * QH_SUMSQ_PARTS * 8);   /* partial sums of the norms */
	/* staged inputs: per job pos | mie params | E (6 doubles) */
	size_t in_off[MAX_BATCH], in_bytes = 0;
	for (int j = 0; j < B; ++j) { in_off[j] = in_bytes; in_bytes += align256(u->jobs[j]->req.npart * 24 + 9 * 8 + 6 * 8); }
	u->pin_bytes = in_bytes + B * 8 * 8;
	pin_take(s, u);
	char *din = bump(&bp, in_bytes);
	for (int j = 0; j < B; ++j) {
		const qh_job *jb = u->jobs[j];
		double *h = (double *)(u->pin + in_off[j]);
		memcpy(h, jb->pos, jb->req.npart * 24);
		mie_params(&jb->req, jb->req.omega, h + 3 * jb->req.npart);
		double *E = h + 3 * jb->req.npart + 9;
		for (int c = 0; c < 3; ++c) { E[2 * c] = creal(jb->req.E[c]); E[2 * c + 1] = cimag(jb->req.E[c]); }
	}
	SRV_HIPCHK(hipMemcpyAsync(din, u->pin, in_bytes, hipMemcpyHostToDevice, st));
	qh_zero(a, B * npad * 16, st);
	qh_zero(info, B * 4, st);
	qh_zero(x, B * npad * 16, st);
	for (int j = 0; j < B; ++j) {
		const qh_job *jb = u->jobs[j];
		const int lMax = jb->req.lMax;
		const size_t np = jb->req.npart, L = jb->L, nb = jb->nb;
		const double *dpos = (const double *)(din + in_off[j]);
		const double *dprm = dpos + 3 * np, *dE = dprm + 9;
		double *mie = bump(&bp, 2 * (lMax + 1) * 16), *td = bump(&bp, nb * 16);
		qh_launch_mie(lMax, 1, dprm, mie, st);
		qh_launch_expand_tdiag(lMax, 1, mie, td, st);
		/* k_w_build reads ttype[p] for every particle: all zeros */
		int *tt = bump(&bp, np * 4);
		qh_zero(tt, np * 4, st);
		const double k = wavenumber_re(jb->req.eps_bg, jb->req.omega);
		qh_launch_modeproblem(qh_tables_get(srv_calc(s, lMax)), np, dpos, tt, td, k, 0.0, M + 2 * j * nn, npad, st);
		qh_launch_pad_identity(1, (int)L, (int)npad, M + 2 * j * nn, (long)nn, st);
		const double kv[3] = {k * jb->req.kdir[0], k * jb->req.kdir[1], k * jb->req.kdir[2]};
		qh_launch_planewave(lMax, np, dpos, kv, dE, a + j * 2 * npad, (long)nb, st);
		qh_launch_tdiag_mul(1, (int)L, (int)nb, (int)npad, td, 0, a + j * 2 * npad, 0, b + j * 2 * npad, 0, st);
	}
	const long sM = (long)nn, sV = (long)npad;
	qh_launch_sumsq_batched(B, (long)nn, M, sM, part, outs + 3 * B, st);
	if (s->o.mixed) {
		qh_d2f((long)(B * nn), M, M32, st);
		qh_getrf_batched_wsla_f32((int)npad, B, M32, sM, ipiv, perm, info, ws, npad >= SRV_LOOKAHEAD_MIN, st);
		qh_copy_c64((long)(B * npad), b, r, st);
		for (int it = 0; it < s->o.refine_steps; ++it) {
			qh_d2f((long)(B * npad), r, r32, st);
			qh_getrs_batched_f32((int)npad, 1, B, M32, sM, perm, r32, sV, w32, st);
			qh_f2d_add((long)(B * npad), r32, x, st);
			qh_copy_c64((long)(B * npad), b, r, st);
			qh_gemm_sub_f64((int)npad, 1, (int)npad, B, M, (long)npad, sM, x, 1, sV, r, 1, sV, st);
		}
	} else {
		/* FP64 LU in place on a copy would need another L^2: factor M itself,
		 * so the residual uses the factors' reconstruction bound instead. */
		qh_getrf_batched_wsla_f64((int)npad, B, M, sM, ipiv, perm, info, ws, npad >= SRV_LOOKAHEAD_MIN, st);
		qh_copy_c64((long)(B * npad), b, x, st);
		qh_getrs_batched_f64((int)npad, 1, B, M, sM, perm, x, sV, work, st);
		qh_zero(r, B * npad * 16, st);
	}
	qh_launch_sumsq_batched(B, (long)npad, r, sV, part, outs + B, st);
	qh_launch_sumsq_batched(B, (long)npad, x, sV, part, outs + 2 * B, st);
	qh_launch_neg_re_cdot(B, (long)npad, a, sV, x, sV, outs, st);
	SRV_HIPCHK(hipMemcpyAsync(outs + 4 * B, info, B * 4, hipMemcpyDeviceToDevice, st));
	SRV_HIPCHK(hipMemcpyAsync(u->pin + in_bytes, outs, B * 8 * 8, hipMemcpyDeviceToHost, st));
}

static void finish_scatter(qpms_hip_server_t *s, qh_unit *u, qpms_hip_response_t *out) {
	const int B = u->njobs;
	size_t in_bytes = 0;
	for (int j = 0; j < B; ++j) in_bytes += align256(u->jobs[j]->req.npart * 24 + 9 * 8 + 6 * 8);
	const double *o = (const double *)(u->pin + in_bytes);
	const int *info = (const int *)(o + 4 * B);
	for (int j = 0; j < B; ++j) {
		const qh_job *jb = u->jobs[j];
		const double k = wavenumber_re(jb->req.eps_bg, jb->req.omega);
		out[j].sigma_ext = o[j] / (k * k);
		out[j].rel_residual = s->o.mixed ? sqrt(o[B + j]) / (sqrt(o[3 * B + j]) * sqrt(o[2 * B + j])) : 0;
		out[j].refine_steps = s->o.mixed ? s->o.refine_steps : 0;
		if (info[j]) { out[j].status = -1; snprintf(out[j].message, sizeof out[j].message, "singular (info %d)", info[j]); }
	}
}

/* ---------------------------------------------------------------------- */
/* Periodic determinants                                                     */
/* ---------------------------------------------------------------------- */
typedef struct { double *R, *K; size_t nR, nK; } qh_points;

/* Reference cutoffs of the periodic builders (scatsystem.c:1226-1227). */
/* Ewald parameters of a periodic request: eta is the reference's adjusted
 * value (scatsystem.c:68-80, default sqrt(pi / A)); the reciprocal cutoff
 * is the reference's, maxK = 2048 2 pi / (64 sqrt(A)) (its truncation is
 * part of the value: with a raised eta the K sum is not converged there).
 * The direct-space cutoff is the reference's 64 sqrt(A) or, if smaller, the
 * radius beyond which every term has |R - s|^2 eta^2 - k^2 / (4 eta^2) >=
 * EW_X: such terms carry e^{-EW_X} = 1.6e-28 times polynomial prefactors of
 * at most (2 |R - s| / k)^9, far below FP64 rounding of the retained sum,
 * where the reference sums them down to underflow. */
#define EW_X 64.0
static void periodic_params(const qpms_hip_request_t *r, double *area, double *eta, double *maxR, double *maxK,
		cart2_t *b1, cart2_t *b2) {
	cart2_t a1 = {r->lattice[0][0], r->lattice[0][1]}, a2 = {r->lattice[1][0], r->lattice[1][1]};
	l2d_reduceBasis(a1, a2, b1, b2);
	*area = l2d_unitcell_area(*b1, *b2);
	const double eta0 = 2. / M_2_SQRTPI / sqrt(*area);
	const double k = wavenumber_re(r->eps_bg, r->omega);
	const double k2 = r->beta[0] * r->beta[0] + r->beta[1] * r->beta[1], kappa2 = k * k;
	*eta = eta0;
	if (kappa2 >= k2) {
		const double mina = 0.5 * (2 - 1) - r->lMax;
		const double eta_min = sqrt(fabs((kappa2 - k2) * (mina - 1.) / 16.));   /* scatsystem.c:42 */
		*eta = fmax(eta0, eta_min);
	}
	/* largest in-cell shift |r_i - r_j|: the bounding-box diagonal */
	double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
	for (size_t i = 0; i < r->npart; ++i)
		for (int c = 0; c < 2; ++c) { lo[c] = fmin(lo[c], r->pos[3 * i + c]); hi[c] = fmax(hi[c], r->pos[3 * i + c]); }
	const double diam = hypot(hi[0] - lo[0], hi[1] - lo[1]);
	const double e = *eta, refR = 64 * sqrt(*area);
	*maxR = fmin(refR, sqrt(EW_X + kappa2 / (4 * e * e)) / e + diam);
	*maxK = 2048 * 2 * M_PI / refR;
}

static void periodic_points(const qpms_hip_request_t *r, qh_points *pt, double *area, double *eta) {
	cart2_t b1, b2, rb1, rb2, z = {0, 0};
	double maxR, maxK;
	periodic_params(r, area, eta, &maxR, &maxK, &b1, &b2);
	QH_ENSURE(l2d_reciprocalBasis2pi(b1, b2, &rb1, &rb2) == 0, "degenerate lattice");
	PGen g = PGen_xyWeb_new(b1, b2, 1e-13, z, 0, true, maxR, false);
	pt->R = qh_collect(&g, &pt->nR, 3);
	for (size_t i = 0; i < pt->nR; ++i) pt->R[3 * i + 2] = 0;
	PGen_destroy(&g);
	g = PGen_xyWeb_new(rb1, rb2, 1e-13, z, 0, true, maxK, false);
	pt->K = qh_collect(&g, &pt->nK, 2);
	PGen_destroy(&g);
}

/* Point counts of periodic_points (the generator's layers are polygons
 * around the disc, so the points are counted, not estimated). */
static void periodic_counts(const qpms_hip_request_t *r, size_t *nR, size_t *nK) {
	qh_points pt;
	double area, eta;
	periodic_points(r, &pt, &area, &eta);
	*nR = pt.nR;
	*nK = pt.nK;
	free(pt.R);
	free(pt.K);
}

static void launch_periodic(qpms_hip_server_t *s, qh_unit *u, hipStream_t st) {
	qh_job *jb = u->jobs[0];
	const qpms_hip_request_t *r = &jb->req;
	const int lMax = r->lMax, LE = 2 * lMax + 1, nsc = (LE + 1) * (LE + 1);
	const size_t np = r->npart, L = jb->L, nb = jb->nb;
	qh_points pt;
	double area, eta;
	periodic_points(r, &pt, &area, &eta);
	size_t cR, cK;
	periodic_counts(r, &cR, &cK);
	QH_ENSURE(pt.nR == cR && pt.nK == cK, "lattice point counts differ between admission and launch");
	qh_bump bp = {s->arena.base + u->arena_off, periodic_bytes(L, np, nsc, cR, cK)};
	double *M = bump(&bp, L * L * 16);
	int *ipiv = bump(&bp, L * 4), *perm = bump(&bp, L * 4);
	void *ws = bump(&bp, qh_getrf_ws_bytes(64, (int)L, 1));
	double *tr = bump(&bp, np * np * 64);
	double *ss = bump(&bp, np * np * nsc * 16), *sl = bump(&bp, np * np * nsc * 16), *tot = bump(&bp, np * np * nsc * 16);
	double *dR = bump(&bp, pt.nR * 24), *dK = bump(&bp, pt.nK * 16), *dpos = bump(&bp, np * 24);
	void *lsF = bump(&bp, qh_ewald3_long_shifts_scratch(srv_calc(s, lMax)->e3c, pt.nK));
	int *tt = bump(&bp, np * 4);
	double *mie = bump(&bp, 2 * (lMax + 1) * 16), *td = bump(&bp, nb * 16), *dk = bump(&bp, 64), *s0 = bump(&bp, 64);
	double *outs = bump(&bp, 64), *dprm = bump(&bp, 128);
	const size_t in_bytes = align256(pt.nR * 24) + align256(pt.nK * 16) + align256(np * 24) + 256;
	u->pin_bytes = in_bytes + 64;
	pin_take(s, u);
	char *h = u->pin;
	memcpy(h, pt.R, pt.nR * 24);
	memcpy(h + align256(pt.nR * 24), pt.K, pt.nK * 16);
	memcpy(h + align256(pt.nR * 24) + align256(pt.nK * 16), jb->pos, np * 24);
	double *hp = (double *)(h + align256(pt.nR * 24) + align256(pt.nK * 16) + align256(np * 24));
	mie_params(r, r->omega, hp);
	const double k = wavenumber_re(r->eps_bg, r->omega);
	hp[9] = k; hp[10] = 0;
	SRV_HIPCHK(hipMemcpyAsync(dR, h, pt.nR * 24, hipMemcpyHostToDevice, st));
	SRV_HIPCHK(hipMemcpyAsync(dK, h + align256(pt.nR * 24), pt.nK * 16, hipMemcpyHostToDevice, st));
	SRV_HIPCHK(hipMemcpyAsync(dpos, h + align256(pt.nR * 24) + align256(pt.nK * 16), np * 24, hipMemcpyHostToDevice, st));
	SRV_HIPCHK(hipMemcpyAsync(dprm, hp, 11 * 8, hipMemcpyHostToDevice, st));
	free(pt.R); free(pt.K);
	qh_zero(tt, np * 4, st);
	qh_launch_mie(lMax, 1, dprm, mie, st);
	qh_launch_expand_tdiag(lMax, 1, mie, td, st);
	SRV_HIPCHK(hipMemcpyAsync(dk, dprm + 9, 16, hipMemcpyDeviceToDevice, st));
	const qpms_trans_calculator *c = srv_calc(s, lMax);
	qh_launch_pair_triples((int)np, dpos, r->beta[0], r->beta[1], k, 0.0, tr, st);
	qpms_hip_ewald3_sigma_batch_dev(c->e3c, eta, area, np * np, tr, pt.nR, dR, 0, NULL, (complex double *)ss, NULL, st);
	qh_ewald3_sigma_long_shifts_dev(c->e3c, eta, area, np * np, tr, pt.nK, dK, lsF, (complex double *)sl, st);
	qh_launch_sigma0(1, dk, eta, c->e3c->legendre0[0], s0, st);
	qh_launch_pair_sigma_total((int)np, nsc, ss, sl, s0, tot, st);
	qh_launch_e32_assemble_M(qh_tables_get(c), (int)np, nsc, tot, qh_e32_bq_mode(), tt, td, M, (long)L, st);
	int *info = tt;   /* reuse: np >= 1 ints */
	qh_zero(info, 4, st);
	qh_getrf_batched_wsla_f64((int)L, 1, M, 0, ipiv, perm, info, ws, L >= SRV_LOOKAHEAD_MIN, st);
	qh_launch_logdet((int)L, M, ipiv, outs, st);
	SRV_HIPCHK(hipMemcpyAsync(outs + 2, info, 4, hipMemcpyDeviceToDevice, st));
	SRV_HIPCHK(hipMemcpyAsync(u->pin + in_bytes, outs, 24, hipMemcpyDeviceToHost, st));
}

static void finish_periodic(qh_unit *u, qpms_hip_response_t *out) {
	const qh_job *jb = u->jobs[0];
	const int LE = 2 * jb->req.lMax + 1;
	(void)LE;
	const double *o = (const double *)(u->pin + (u->pin_bytes - 64));
	out->logabsdet = o[0];
	out->argdet = o[1];
	const int info = *(const int *)(o + 2);
	if (info) { out->status = -1; snprintf(out->message, sizeof out->message, "singular (info %d)", info); }
}

/* ---------------------------------------------------------------------- */
/* Beyn contours                                                             */
/* ---------------------------------------------------------------------- */
static void launch_beyn(qpms_hip_server_t *s, qh_unit *u, hipStream_t st, int chunk) {
	qh_job *jb = u->jobs[0];
	const qpms_hip_request_t *r = &jb->req;
	const int lMax = r->lMax;
	const size_t np = r->npart, L = jb->L, nb = jb->nb, l = (size_t)r->beyn_l;
	const int N = r->beyn_nodes;
	qh_bump bp = {s->arena.base + u->arena_off, beyn_bytes(s, L, l, chunk, N, lMax)};
	u->chunk = chunk;
	u->next_node = 0;
	u->bV = bump(&bp, L * l * 16);
	u->dA0 = bump(&bp, L * l * 16); u->dA1 = bump(&bp, L * l * 16);
	u->dA0c = bump(&bp, L * l * 16); u->dA1c = bump(&bp, L * l * 16);
	u->bpos = bump(&bp, np * 24);
	u->bws = bump(&bp, qh_getrf_ws_bytes(s->o.mixed ? 32 : 64, (int)L, chunk));
	u->bM = bump(&bp, chunk * L * L * 16);
	u->bM32 = s->o.mixed ? bump(&bp, chunk * L * L * 8) : NULL;
	u->bX = bump(&bp, chunk * L * l * 16); u->bR = bump(&bp, chunk * L * l * 16); u->bW = bump(&bp, chunk * L * l * 16);
	u->bR32 = bump(&bp, chunk * L * l * 8); u->bW32 = bump(&bp, chunk * L * l * 8);
	u->bipiv = bump(&bp, chunk * L * 4); u->bperm = bump(&bp, chunk * L * 4); u->binfo = bump(&bp, chunk * 4);
	u->btt = bump(&bp, np * 4);
	double *dprm = bump(&bp, N * 9 * 8), *mie = bump(&bp, N * 2 * (lMax + 1) * 16);
	u->btd = bump(&bp, N * nb * 16);
	u->btail = bump(&bp, qh_beyn_ws_bytes(L, l));
	u->btab = bump(&bp, N * 64);
	u->bcur = bump(&bp, 256);
	/* staged: positions, the Mie parameters and the table entry of every node */
	const size_t in_bytes = align256(np * 24) + align256(N * 9 * 8) + align256(N * 64);
	u->pin_bytes = in_bytes + 64;
	pin_take(s, u);
	memcpy(u->pin, jb->pos, np * 24);
	u->bcont = beyn_contour_ellipse(CMPLX(r->omega, r->beyn_centre_im), r->beyn_rRe, r->beyn_rIm, (size_t)N);
	double *hp = (double *)(u->pin + align256(np * 24));
	double *ht = (double *)(u->pin + align256(np * 24) + align256(N * 9 * 8));
	const complex double z0 = u->bcont->centre;
	for (int n = 0; n < N; ++n) {
		const complex double om = u->bcont->z_dz[n][0], dz = u->bcont->z_dz[n][1];
		const complex double k = csqrt((complex double)r->eps_bg) * om / SPEED_OF_LIGHT;
		mie_params(r, om, hp + 9 * n);
		double *t = ht + 8 * n;
		t[0] = creal(k); t[1] = cimag(k);
		t[2] = creal(dz); t[3] = cimag(dz);
		t[4] = creal(om - z0); t[5] = cimag(om - z0);
		t[6] = (n % 2) == 0; t[7] = 0;
	}
	SRV_HIPCHK(hipMemcpyAsync(u->bpos, u->pin, np * 24, hipMemcpyHostToDevice, st));
	SRV_HIPCHK(hipMemcpyAsync(dprm, hp, N * 9 * 8, hipMemcpyHostToDevice, st));
	SRV_HIPCHK(hipMemcpyAsync(u->btab, ht, N * 64, hipMemcpyHostToDevice, st));
	qh_zero(u->bcur, 256, st);
	qh_zero(u->btt, np * 4, st);
	qh_zero(u->dA0, L * l * 16, st);
	qh_zero(u->dA1, L * l * 16, st);
	qh_zero(u->dA0c, L * l * 16, st);
	qh_zero(u->dA1c, L * l * 16, st);
	qh_launch_probe((long)(L * l), r->beyn_seed, u->bV, st);
	qh_launch_mie(lMax, N, dprm, mie, st);
	qh_launch_expand_tdiag(lMax, N, mie, u->btd, st);
}

/* The launches of one chunk of C quadrature nodes: M(z_i), X_i = M(z_i)^{-1} V
 * (batched LU over the chunk, refined), moments A0, A1 (+ coarse). The
 * node-dependent launches read their node (k, dz, z - z0, coarse flag) from
 * the node table at the device cursor, which the chunk advances by C at its
 * end: the sequence is the same for every chunk of the contour, whether it
 * is launched eagerly or replayed from a captured graph. */
static void beyn_chunk_body(qpms_hip_server_t *s, qh_unit *u, hipStream_t st, int C) {
	qh_job *jb = u->jobs[0];
	const qpms_hip_request_t *r = &jb->req;
	const size_t np = r->npart, L = jb->L, l = (size_t)r->beyn_l;
	const qh_trans_tables *tb = qh_tables_get(srv_calc(s, r->lMax));
	const long sM = (long)(L * L), sV = (long)(L * l);
	double *M = u->bM, *X = u->bX, *R = u->bR;
	for (int i = 0; i < C; ++i)
		qh_launch_modeproblem_node(tb, np, u->bpos, u->btt, u->btd, u->btab, u->bcur, i, M + 2 * i * sM, L, st);
	qh_zero(u->binfo, C * 4, st);
	qh_zero(X, C * L * l * 16, st);
	for (int i = 0; i < C; ++i) qh_copy_c64((long)(L * l), u->bV, R + i * sV * 2, st);
	if (s->o.mixed) {
		qh_d2f((long)(C * L * L), M, u->bM32, st);
		qh_getrf_batched_wsla_f32((int)L, C, u->bM32, sM, u->bipiv, u->bperm, u->binfo, u->bws,
			L >= SRV_LOOKAHEAD_MIN, st);
		/* nodes near an eigenvalue make M ill-conditioned: twice the steps of a solve */
		for (int it = 0; it < 2 * s->o.refine_steps; ++it) {
			qh_d2f((long)(C * L * l), R, u->bR32, st);
			qh_getrs_batched_f32((int)L, (int)l, C, u->bM32, sM, u->bperm, u->bR32, sV, u->bW32, st);
			qh_f2d_add((long)(C * L * l), u->bR32, X, st);
			for (int i = 0; i < C; ++i) qh_copy_c64((long)(L * l), u->bV, R + i * sV * 2, st);
			qh_gemm_sub_f64((int)L, (int)l, (int)L, C, M, (long)L, sM, X, (long)l, sV, R, (long)l, sV, st);
		}
	} else {
		qh_getrf_batched_wsla_f64((int)L, C, M, sM, u->bipiv, u->bperm, u->binfo, u->bws, L >= SRV_LOOKAHEAD_MIN, st);
		for (int i = 0; i < C; ++i) qh_copy_c64((long)(L * l), u->bV, X + i * sV * 2, st);
		qh_getrs_batched_f64((int)L, (int)l, C, M, sM, u->bperm, X, sV, u->bW, st);
	}
	for (int i = 0; i < C; ++i)
		qh_beyn_accum_node((long)(L * l), X + i * sV * 2, u->btab, u->bcur, i, (double *)u->dA0, (double *)u->dA1,
			(double *)u->dA0c, (double *)u->dA1c, st);
	qh_cursor_advance(u->bcur, C, st);
}

/* Enqueue the next chunk. With graphs on, the first chunk's launches are
 * captured into a graph (launched at once) and every later chunk is one
 * launch of that graph; the chunk size divides the node count, so all chunks
 * share the sequence. */
static void beyn_enqueue_chunk(qpms_hip_server_t *s, qh_unit *u, hipStream_t st) {
	const int N = u->jobs[0]->req.beyn_nodes, n0 = u->next_node;
	const int C = (N - n0 < u->chunk) ? N - n0 : u->chunk;
	if (s->graphs && C == u->chunk && N / C > 1 && u->jobs[0]->L <= SRV_GRAPH_MAX) {
		if (!u->gexec) {
			hipGraph_t g;
			SRV_HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
			beyn_chunk_body(s, u, st, C);
			SRV_HIPCHK(hipStreamEndCapture(st, &g));
			SRV_HIPCHK(hipGraphInstantiate(&u->gexec, g, NULL, NULL, 0));
			SRV_HIPCHK(hipGraphDestroy(g));
		}
		SRV_HIPCHK(hipGraphLaunch(u->gexec, st));
	} else {
		beyn_chunk_body(s, u, st, C);
	}
	u->next_node = n0 + C;
}

static void finish_beyn(qpms_hip_server_t *s, qh_unit *u, hipStream_t st, qpms_hip_response_t *out) {
	(void)s;
	const qh_job *jb = u->jobs[0];
	const qpms_hip_request_t *r = &jb->req;
	const size_t l = (size_t)r->beyn_l;
	free(u->bcont);
	u->bcont = NULL;
	if (u->cancelled) {
		out->status = -2;
		snprintf(out->message, sizeof out->message, "cancelled after %d of %d nodes", u->next_node, r->beyn_nodes);
		return;
	}
	const int K = qh_beyn_process_s(jb->L, l, u->dA0, u->dA1, CMPLX(r->omega, r->beyn_centre_im), r->rank_tol, 0,
		out->eigval, out->sv, st,
		u->btail);
	out->neig = K;
}

/* ---------------------------------------------------------------------- */
/* Progress                                                                  */
/* ---------------------------------------------------------------------- */
static int free_lane(const qpms_hip_server_t *s) {
	for (int i = 0; i < s->o.nstreams; ++i) if (!s->lane_busy[i]) return i;
	return -1;
}

static void push_ready(qpms_hip_server_t *s, const qpms_hip_response_t *r) {
	if (s->nready == s->capready) {
		s->capready = s->capready ? 2 * s->capready : 64;
		s->ready = realloc(s->ready, s->capready * sizeof(*s->ready));
		QH_ENSURE(s->ready, "oom");
	}
	s->ready[s->nready++] = *r;
}

static void retire(qpms_hip_server_t *s, qh_unit *u) {
	const double now = qpms_hip_server_clock(s);
	qpms_hip_response_t out[MAX_BATCH];
	memset(out, 0, u->njobs * sizeof(out[0]));
	for (int j = 0; j < u->njobs; ++j) {
		out[j].id = u->jobs[j]->id;
		out[j].kind = u->jobs[j]->req.kind;
		out[j].dim = u->jobs[j]->L;
		out[j].t_submit = u->jobs[j]->t_submit;
		out[j].t_start = u->t_start;
		out[j].batch = u->njobs;
		out[j].padded_dim = (int)u->npad;
	}
	if (u->kind == U_SCATTER) finish_scatter(s, u, out);
	else if (u->kind == U_PERIODIC) finish_periodic(u, out);
	else finish_beyn(s, u, s->lane[u->lane], out);
	const double t_done = (u->kind == U_BEYN) ? qpms_hip_server_clock(s) : now;
	for (int j = 0; j < u->njobs; ++j) {
		out[j].t_done = t_done;
		push_ready(s, &out[j]);
		free(u->jobs[j]->pos);
		free(u->jobs[j]);
	}
	SRV_HIPCHK(hipEventDestroy(u->done));
	if (u->gexec) SRV_HIPCHK(hipGraphExecDestroy(u->gexec));
	u->gexec = NULL;
	arena_free(&s->pinned, u->pin_off);
	arena_free(&s->arena, u->arena_off);
	s->lane_busy[u->lane] = 0;
	u->active = 0;
	--s->ninflight;
}

/* Launch order: shortest estimated work first, with aging (a job's key
 * drops by SRV_AGING seconds per second waited), so that short requests
 * are not held behind long contours while long ones still start
 * eventually. Ties keep arrival order. */
#define SRV_AGING 0.1
static int key_cmp(const void *a, const void *b, void *now_) {
	const double now = *(const double *)now_;
	const qh_job *x = *(qh_job *const *)a, *y = *(qh_job *const *)b;
	const double kx = x->est - SRV_AGING * (now - x->t_submit), ky = y->est - SRV_AGING * (now - y->t_submit);
	if (kx != ky) return kx < ky ? -1 : 1;
	return x->id < y->id ? -1 : (x->id > y->id);
}

static qh_job *q_prev(qpms_hip_server_t *s, qh_job *j) {
	qh_job *pv = NULL;
	for (qh_job *c = s->qhead; c && c != j; c = c->next) pv = c;
	return pv;
}

static int try_launch(qpms_hip_server_t *s) {
	const int lane = free_lane(s);
	if (lane < 0 || !s->qhead) return 0;
	if (s->nqueued > s->order_cap) {
		s->order_cap = s->nqueued * 2;
		s->order = realloc(s->order, s->order_cap * sizeof(*s->order));
		QH_ENSURE(s->order, "oom");
	}
	size_t nq = 0;
	for (qh_job *c = s->qhead; c; c = c->next) s->order[nq++] = c;
	double now = qpms_hip_server_clock(s);
	qsort_r(s->order, nq, sizeof(*s->order), key_cmp, &now);
	/* first job in that order that fits (backfilling past jobs that do not) */
	for (size_t oi = 0; oi < nq; ++oi) {
		qh_job *j = s->order[oi];
		/* 1.5: margin on the estimate, which ignores contention with other lanes */
		if (s->deadline > 0 && now + 1.5 * j->est > s->deadline) continue;
		int chunk = 1;
		const int small = j->req.kind == QPMS_HIP_REQ_SCATTER && j->L <= s->o.batch_max_dim;
		qh_unit *u = unit_new(s);
		if (!u) return 0;
		if (small) {
			/* batch: this job and the next queued small solves (launch order) of the same padded order */
			const size_t npad = pad_bucket(j->L);
			qh_job *members[MAX_BATCH];
			int B = 0;
			for (size_t q = oi; q < nq && B < s->o.batch_max; ++q) {
				qh_job *c = s->order[q];
				if (c->req.kind == QPMS_HIP_REQ_SCATTER && c->L <= s->o.batch_max_dim && pad_bucket(c->L) == npad)
					members[B++] = c;
			}
			long off = -1;
			while (B > 0) {
				size_t mnp = 0;
				for (int q = 0; q < B; ++q) if (members[q]->req.npart > mnp) mnp = members[q]->req.npart;
				off = arena_alloc(&s->arena, scatter_bytes(s, npad, B, mnp));
				if (off >= 0) break;
				--B;
			}
			if (off < 0) { u->active = 0; continue; }
			for (int q = 0; q < B; ++q) q_remove(s, q_prev(s, members[q]), members[q]);
			u->kind = U_SCATTER;
			u->njobs = B;
			for (int q = 0; q < B; ++q) u->jobs[q] = members[q];
			u->npad = npad;
			u->arena_off = off;
		} else {
			const size_t need = job_bytes(s, j, &chunk);
			const long off = arena_alloc(&s->arena, need);
			if (off < 0) { u->active = 0; continue; }
			q_remove(s, q_prev(s, j), j);
			u->kind = (j->req.kind == QPMS_HIP_REQ_SCATTER) ? U_SCATTER : (j->req.kind == QPMS_HIP_REQ_PERIODIC_DET) ? U_PERIODIC : U_BEYN;
			u->njobs = 1;
			u->jobs[0] = j;
			u->npad = j->L;
			u->arena_off = off;
		}
		u->lane = lane;
		s->lane_busy[lane] = 1;
		++s->ninflight;
		u->t_start = qpms_hip_server_clock(s);
		hipStream_t st = s->lane[lane];
		if (u->kind == U_SCATTER) launch_scatter(s, u, st);
		else if (u->kind == U_PERIODIC) launch_periodic(s, u, st);
		else { launch_beyn(s, u, st, chunk); beyn_enqueue_chunk(s, u, st); }
		SRV_HIPCHK(hipEventCreateWithFlags(&u->done, hipEventDisableTiming));
		SRV_HIPCHK(hipEventRecord(u->done, st));
		return 1;
	}
	return 0;
}

int qpms_hip_server_progress(qpms_hip_server_t *s, int wait) {
	for (;;) {
		/* retire finished units */
		for (int i = 0; i < MAX_UNITS; ++i) {
			qh_unit *u = &s->unit[i];
			if (!u->active || !u->done) continue;
			const hipError_t q = hipEventQuery(u->done);
			if (q == hipSuccess) {
				if (u->kind == U_BEYN && !u->cancelled && u->next_node < u->jobs[0]->req.beyn_nodes) {
					beyn_enqueue_chunk(s, u, s->lane[u->lane]);
					SRV_HIPCHK(hipEventRecord(u->done, s->lane[u->lane]));
				} else retire(s, u);
			} else if (q != hipErrorNotReady) SRV_HIPCHK(q);
		}
		while (try_launch(s)) {}
		if (!wait || s->nready > 0 || s->ninflight == 0) return s->nready;
		/* wait for whichever unit finishes first: poll every event (a blocking
		 * wait on one of them would hold up the others' retirement) */
		const struct timespec nap = {0, 50000};
		nanosleep(&nap, NULL);
	}
}

int qpms_hip_server_poll(qpms_hip_server_t *s, qpms_hip_response_t *out, int max) {
	const int n = s->nready < max ? s->nready : max;
	memcpy(out, s->ready, n * sizeof(*out));
	memmove(s->ready, s->ready + n, (s->nready - n) * sizeof(*out));
	s->nready -= n;
	return n;
}

void qpms_hip_server_set_deadline(qpms_hip_server_t *s, double t) { s->deadline = t; }

void qpms_hip_server_cancel(qpms_hip_server_t *s) {
	const double now = qpms_hip_server_clock(s);
	while (s->qhead) {
		qh_job *j = s->qhead;
		q_remove(s, NULL, j);
		qpms_hip_response_t r;
		memset(&r, 0, sizeof r);
		r.id = j->id; r.kind = j->req.kind; r.dim = j->L; r.status = -2;
		r.t_submit = j->t_submit; r.t_start = r.t_done = now;
		snprintf(r.message, sizeof r.message, "cancelled before admission");
		push_ready(s, &r);
		free(j->pos);
		free(j);
	}
	for (int i = 0; i < MAX_UNITS; ++i)
		if (s->unit[i].active && s->unit[i].kind == U_BEYN) s->unit[i].cancelled = 1;
}
