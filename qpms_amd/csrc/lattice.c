/* 2D lattice helpers and the "spiderweb" ring point generator.
 *
 * The lattice point sets decide which terms enter the Ewald sums, so the
 * reduction, reciprocal basis and ring rules follow the reference step for
 * step: lattices2d.c:635-653 (Lagrange–Gauss reduction), 663-686 (shortest
 * base triple, obtuse test), 746-756 (right-angle detection), 801-823
 * (reciprocal bases), 841-853 (ring in-radius, cell area) and
 * latticegens.c:449-567 (ring order: origin, then layer l = 1, 2, ... walked
 * along 6 (hexagonal) or 4 (rectangular) edges; layer l is generated iff
 * l*h >= minR (strictly > if !inc_minR) and l*h <= maxR (strictly < if
 * !inc_maxR), h = in-radius of the layer-1 hexagon/rectangle).
 * Enumeration is host plumbing: the points are uploaded and summed on the GPU.
 */
#include "qpms_internal.h"
#include <math.h>
#include <string.h>

static inline double dot2(cart2_t a, cart2_t b) { return a.x * b.x + a.y * b.y; }
static inline cart2_t sub2(cart2_t a, cart2_t b) { cart2_t r = {a.x - b.x, a.y - b.y}; return r; }
static inline cart2_t add2(cart2_t a, cart2_t b) { cart2_t r = {a.x + b.x, a.y + b.y}; return r; }
static inline cart2_t scale2(double s, cart2_t a) { cart2_t r = {s * a.x, s * a.y}; return r; }

void l2d_reduceBasis(cart2_t b1, cart2_t b2, cart2_t *out1, cart2_t *out2) {
	double B1 = dot2(b1, b1);
	b2 = sub2(b2, scale2(round(dot2(b1, b2) / B1), b1));
	double B2 = dot2(b2, b2);
	while (B2 < B1) {
		cart2_t t = b1; b1 = b2; b2 = t;
		B1 = B2;
		b2 = sub2(b2, scale2(round(dot2(b1, b2) / B1), b1));
		B2 = dot2(b2, b2);
	}
	*out1 = b1;
	*out2 = b2;
}

static int qh_is_obtuse(cart2_t b1, cart2_t b2) {
	cart2_t b3 = sub2(b2, b1);
	return dot2(b3, b3) - dot2(b2, b2) - dot2(b1, b1) > 0;
}

static void qh_shortest_base3(cart2_t b1, cart2_t b2, cart2_t *o1, cart2_t *o2, cart2_t *o3) {
	l2d_reduceBasis(b1, b2, &b1, &b2);
	*o1 = b1;
	if (qh_is_obtuse(b1, b2)) { *o3 = b2; *o2 = add2(b2, b1); }
	else { *o2 = b2; *o3 = sub2(b2, b1); }
}

static int qh_right_angle(cart2_t b1, cart2_t b2, double rtol) {
	l2d_reduceBasis(b1, b2, &b1, &b2);
	cart2_t h = sub2(b2, b1);
	double b1s = dot2(b1, b1), b2s = dot2(b2, b2), hs = dot2(h, h);
	return (hs - b2s - b1s <= rtol * (b1s + b2s));
}

static int qh_reciprocal1(cart2_t b1, cart2_t b2, cart2_t *rb1, cart2_t *rb2) {
	l2d_reduceBasis(b1, b2, &b1, &b2);
	const double det = b1.x * b2.y - b1.y * b2.x;
	if (!det) { rb1->x = rb1->y = rb2->x = rb2->y = NAN; return QPMS_ERROR; }
	rb1->x = b2.y / det; rb1->y = -b2.x / det;
	rb2->x = -b1.y / det; rb2->y = b1.x / det;
	return QPMS_SUCCESS;
}

int l2d_reciprocalBasis2pi(cart2_t b1, cart2_t b2, cart2_t *rb1, cart2_t *rb2) {
	int r = qh_reciprocal1(b1, b2, rb1, rb2);
	if (r == QPMS_SUCCESS) { *rb1 = scale2(2 * M_PI, *rb1); *rb2 = scale2(2 * M_PI, *rb2); }
	return r;
}

int qh_reciprocalBasis1(cart2_t b1, cart2_t b2, cart2_t *rb1, cart2_t *rb2) {
	return qh_reciprocal1(b1, b2, rb1, rb2);
}

double l2d_hexWebInCircleRadius(cart2_t i1, cart2_t i2) {
	cart2_t b1, b2, b3;
	qh_shortest_base3(i1, i2, &b1, &b2, &b3);
	const double r1 = sqrt(dot2(b1, b1)), r2 = sqrt(dot2(b2, b2)), r3 = sqrt(dot2(b3, b3));
	const double p = (r1 + r2 + r3) * 0.5;
	return 2 * sqrt(p * (p - r1) * (p - r2) * (p - r3)) / r3;
}

double l2d_unitcell_area(cart2_t b1, cart2_t b2) {
	l2d_reduceBasis(b1, b2, &b1, &b2);
	return fabs(b1.x * b2.y - b1.y * b2.x);
}

/* ------------------------------------------------------------------------ */
/* PGen plumbing                                                              */
/* ------------------------------------------------------------------------ */
typedef struct qh_xyweb_state {
	long i, j;
	int phase;
	long layer, last_layer;
	double h;
	int rect;
	cart2_t b1, b2, offset;
} qh_xyweb_state;

static PGenCart2ReturnData qh_xyweb_next_cart2(PGen *g);
static PGenSphReturnData qh_xyweb_next_sph(PGen *g);
static void qh_xyweb_destroy(PGen *g) { free(g->stateData); g->stateData = NULL; }

static const PGenClassInfo qh_PGen_xyWeb = {
	"PGen_xyWeb", 2, PGEN_COORDS_CART2,
	NULL, NULL, NULL, qh_xyweb_next_sph, qh_xyweb_next_cart2, NULL,
	NULL, NULL, NULL, NULL, NULL, NULL, qh_xyweb_destroy
};

PGen PGen_xyWeb_new(cart2_t b1, cart2_t b2, double rtol, cart2_t offset,
		double minR, bool inc_minR, double maxR, bool inc_maxR) {
	qh_xyweb_state *s;
	QH_MALLOC(s, sizeof(*s));
	l2d_reduceBasis(b1, b2, &s->b1, &s->b2);
	s->offset = offset;
	s->rect = qh_right_angle(s->b1, s->b2, rtol);
	s->h = l2d_hexWebInCircleRadius(s->b1, s->b2);
	s->layer = (long)ceil(minR / s->h);
	if (!inc_minR && s->layer * s->h <= minR) ++s->layer;
	s->i = s->layer; s->j = 0; s->phase = 0;
	s->last_layer = (long)floor(maxR / s->h);
	if (!inc_maxR && s->last_layer * s->h >= maxR) --s->last_layer;
	PGen g = {&qh_PGen_xyWeb, s};
	return g;
}

static PGenCart2ReturnData qh_xyweb_next_cart2(PGen *g) {
	PGenCart2ReturnData done = {PGEN_DONE, {0, 0}};
	qh_xyweb_state *s = g->stateData;
	if (!s) return done;
	if (s->layer > s->last_layer) { PGen_destroy(g); return done; }
	cart2_t pt = add2(s->offset, add2(scale2((double)s->i, s->b1), scale2((double)s->j, s->b2)));
	if (s->layer == 0) {
		++s->layer; s->phase = 0; s->i = s->layer; s->j = 0;
	} else if (s->rect) {
		switch (s->phase) {
			case 0: --s->i; ++s->j; if (s->i <= 0) ++s->phase; break;
			case 1: --s->i; --s->j; if (s->j <= 0) ++s->phase; break;
			case 2: ++s->i; --s->j; if (s->i >= 0) ++s->phase; break;
			default: ++s->i; ++s->j; if (s->j >= 0) ++s->phase; break;
		}
		if (s->phase == 4) { ++s->layer; s->phase = 0; s->i = s->layer; s->j = 0; }
	} else {
		switch (s->phase) {
			case 0: --s->i; ++s->j; if (s->i <= 0) ++s->phase; break;
			case 1: --s->i; if (s->i + s->j <= 0) ++s->phase; break;
			case 2: --s->j; if (s->j <= 0) ++s->phase; break;
			case 3: ++s->i; --s->j; if (s->i >= 0) ++s->phase; break;
			case 4: ++s->i; if (s->i + s->j >= 0) ++s->phase; break;
			default: ++s->j; if (s->j >= 0) ++s->phase; break;
		}
		if (s->phase == 6) { ++s->layer; s->phase = 0; s->i = s->layer; s->j = 0; }
	}
	PGenCart2ReturnData r = {PGEN_NOTDONE | PGEN_AT_XY | PGEN_COORDS_CART2, pt};
	return r;
}

static PGenSphReturnData qh_xyweb_next_sph(PGen *g) {
	PGenCart2ReturnData c = qh_xyweb_next_cart2(g);
	PGenSphReturnData r = {PGEN_DONE, {0, 0, 0}};
	if (!(c.flags & PGEN_NOTDONE)) return r;
	r.flags = (PGenPointFlags)((c.flags & ~PGEN_COORDS_BITRANGE) | PGEN_COORDS_SPH);
	r.point_sph.r = hypot(c.point_cart2.x, c.point_cart2.y);
	r.point_sph.theta = M_PI_2;
	r.point_sph.phi = atan2(c.point_cart2.y, c.point_cart2.x);
	return r;
}

PGenSphReturnData PGen_next_sph(PGen *g) {
	PGenSphReturnData done = {PGEN_DONE, {0, 0, 0}};
	if (!g->stateData) return done;
	QH_ENSURE(g->c->next_sph, "generator %s has no spherical-coordinate method", g->c->name);
	return g->c->next_sph(g);
}

PGenCart2ReturnData PGen_next_cart2(PGen *g) {
	PGenCart2ReturnData done = {PGEN_DONE, {0, 0}};
	if (!g->stateData) return done;
	QH_ENSURE(g->c->next_cart2, "generator %s has no 2D-cartesian method", g->c->name);
	return g->c->next_cart2(g);
}

void PGen_destroy(PGen *g) {
	if (g->stateData && g->c->destructor) g->c->destructor(g);
	g->stateData = NULL;
}
