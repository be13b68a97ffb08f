/*
 * CPU oracle for psim's phonon path -- TEST INFRASTRUCTURE ONLY (see
 * psim_oracle.h).  Plain C99 restatement of the reference; file:line
 * citations refer to GwGibson/Psim.
 *
 * Deliberate differences from the reference, all statistically neutral:
 *  - random numbers come from the Philox schedule of psim_rng.h;
 *  - the transport loop uses psim_log / psim_cos2pi / psim_sinh / psim_pow4
 *    (psim_math.h) where the reference calls glibc log / cos / sinh / pow;
 *    the table builder keeps glibc like the reference;
 *  - the reference evaluates getRandPoint(urand(), urand()) with unspecified
 *    argument order; here r1 is drawn first;
 *  - reads the reference would make out of bounds (a birth step >= the number
 *    of steps, a steady-state tally index >= the recorded length) are
 *    skipped and counted;
 *  - a phonon does not stop at every measurement boundary: it is counted and
 *    tallied at each boundary its flight passes (see simulate_phonon), so
 *    positions and times to scatter round differently, nothing else changes.
 */
#define _POSIX_C_SOURCE 200809L
#include "psim_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "psim_gpu.h" /* psim_flight_cap */
#include "psim_math.h"
#include "psim_rng.h"

/* ------------------------------------------------------------------------ */
/* Constants (psim/include/psim/utils.h:10-11, material.h:14,
 * modelSimulator.cpp:17-25, model.cpp:11-25, material.cpp:12-13,
 * sensorInterpreter.cpp:9-10). */
#define GEOEPS (DBL_EPSILON * 1E9)
#define NBINS 1000
#define SCALING_FACTOR 1e9
#define VELOCITY_EPS 0.01
#define MAX_COLLISIONS 100
#define TEMP_BOUND_EPS 10.0
#define PHASOR_TEMP_BOUND_EPS (TEMP_BOUND_EPS * 100.0)
#define SS_STEPS_PERCENT 0.1
#define TEMP_INTERVAL 0.1f
#define HBAR 1.054517e-34
#define BOLTZ 1.38065e-23
#define INV_EPS 0.0001
#define INV_MAX_ITERS 40
#define MAX_SUBS 16
/* Loop iterations allowed inside one measurement step before the phonon is removed; the
 * reference has no bound, the GPU engine applies the same one so every wave terminates. */
#define MAX_EVENTS_PER_STEP 1000000

enum { SIM_SS = 0, SIM_PERIODIC = 1, SIM_TRANSIENT = 2 };
enum { POL_LA = 0, POL_TA = 1 };

/* ------------------------------------------------------------------------ */
/* Geometry (psim/src/geometry.cpp) */

typedef struct {
    double x, y;
} pt;

typedef struct {
    pt p1, p2;
    double slope, intercept;
    pt bl, tr;
    double length;
} line_t;

/* Point equality within GEOEPS, geometry.cpp:34-39 */
static int pt_eq(pt a, pt b) {
    const double dx = a.x - b.x;
    const double dy = a.y - b.y;
    return dx * dx + dy * dy < GEOEPS * GEOEPS;
}

/* Utils::approxEqual, utils.h:68-70 */
static int approx_equal(double a, double b) { return fabs(a - b) < GEOEPS; }

/* getSlope (vertical lines get slope 0), geometry.cpp:325-327 */
static double get_slope(pt p1, pt p2) { return approx_equal(p1.x, p2.x) ? 0. : (p1.y - p2.y) / (p1.x - p2.x); }

/* getIntercept, geometry.cpp:329-331 */
static double get_intercept(double x, double y, double slope) { return y - slope * x; }

/* Line::Line, geometry.cpp:45-53; findBoundingBox :267-280; getLineLength :319-323 */
static int line_make(pt p1, pt p2, line_t* l) {
    l->p1 = p1;
    l->p2 = p2;
    l->slope = get_slope(p1, p2);
    l->intercept = get_intercept(p1.x, p1.y, l->slope);
    double bl_x = p1.x, bl_y = p1.y, tr_x = p2.x, tr_y = p2.y, t;
    if (p2.x <= p1.x) { t = bl_x; bl_x = tr_x; tr_x = t; }
    if (p2.y <= p1.y) { t = bl_y; bl_y = tr_y; tr_y = t; }
    l->bl.x = bl_x;
    l->bl.y = bl_y;
    l->tr.x = tr_x;
    l->tr.y = tr_y;
    l->length = sqrt((p2.x - p1.x) * (p2.x - p1.x) + (p2.y - p1.y) * (p2.y - p1.y));
    return pt_eq(p1, p2) ? -1 : 0; /* LineError */
}

/* crossProduct / dotProduct, geometry.cpp:311-317 */
static double cross(pt a, pt b) { return a.x * b.y - b.x * a.y; }
static double dot(pt a, pt b) { return a.x * b.x + a.y * b.y; }
static pt sub(pt a, pt b) {
    pt r = { a.x - b.x, a.y - b.y };
    return r;
}

/* isPointOnLine (infinite line), geometry.cpp:291-295 */
static int point_on_line(const line_t* l, pt p) {
    pt d = { l->p2.x - l->p1.x, l->p2.y - l->p1.y };
    pt q = { p.x - l->p1.x, p.y - l->p1.y };
    return fabs(cross(d, q)) < GEOEPS;
}

/* isPointRightOfLine, geometry.cpp:297-300 */
static int point_right_of_line(const line_t* l, pt p) {
    pt d = { l->p2.x - l->p1.x, l->p2.y - l->p1.y };
    pt q = { p.x - l->p1.x, p.y - l->p1.y };
    return cross(d, q) < 0.;
}

/* doesSegmentCrossLine, geometry.cpp:302-306 */
static int segment_crosses_line(const line_t* l1, const line_t* l2) {
    return point_on_line(l1, l2->p1) || point_on_line(l1, l2->p2)
           || ((point_right_of_line(l1, l2->p1) ^ point_right_of_line(l1, l2->p2)) != 0);
}

/* boxesIntersect, geometry.cpp:282-289 */
static int boxes_intersect(const line_t* a, const line_t* b) {
    return a->bl.x <= b->tr.x - GEOEPS && a->tr.x >= b->bl.x + GEOEPS && a->bl.y <= b->tr.y - GEOEPS
           && a->tr.y >= b->bl.y + GEOEPS;
}

/* Line::intersects, geometry.cpp:92-95 */
static int line_intersects(const line_t* a, const line_t* b) {
    return boxes_intersect(a, b) && segment_crosses_line(a, b) && segment_crosses_line(b, a);
}

/* equals lambda of getIntersection, geometry.cpp:108-110 */
static int rel_equals(double a, double b) {
    return fabs(a - b) <= ((fabs(a) < fabs(b) ? fabs(b) : fabs(a)) * GEOEPS);
}

/* handle_vert lambda, geometry.cpp:112-121 */
static int handle_vert(pt pt1, pt pt3, pt pt4, pt* poi) {
    if (rel_equals(pt3.x, pt4.x)) { return 0; }
    const double m = get_slope(pt3, pt4);
    const double b = get_intercept(pt3.x, pt3.y, m);
    poi->x = pt1.x;
    poi->y = m * pt1.x + b;
    return 1;
}

/* Line::getIntersection, geometry.cpp:104-138 */
static int line_get_intersection(const line_t* self, const line_t* other, pt* poi) {
    if (!line_intersects(self, other)) { return 0; }
    if (rel_equals(self->p1.x, self->p2.x)) { return handle_vert(self->p1, other->p1, other->p2, poi); }
    if (rel_equals(other->p1.x, other->p2.x)) { return handle_vert(other->p1, self->p1, self->p2, poi); }
    if (rel_equals(self->slope, other->slope)) { return 0; }
    const double x = (other->intercept - self->intercept) / (self->slope - other->slope);
    poi->x = x;
    poi->y = self->slope * x + self->intercept;
    return 1;
}

/* Line::normal, geometry.cpp:97-100 */
static pt line_normal(const line_t* l, int norm_sign) {
    norm_sign = (norm_sign >= 0) ? 1 : -1;
    pt n = { (double)norm_sign * (l->p2.y - l->p1.y) / l->length,
        (double)(-norm_sign) * (l->p2.x - l->p1.x) / l->length };
    return n;
}

/* Line::contains(Line), geometry.cpp:74-86 */
static int line_contains_line(const line_t* self, const line_t* other) {
    if (point_on_line(self, other->p1) && point_on_line(self, other->p2) && self->length >= other->length) {
        const double l1_x_min = fmin(self->p1.x, self->p2.x), l1_x_max = fmax(self->p1.x, self->p2.x);
        const double l2_x_min = fmin(other->p1.x, other->p2.x), l2_x_max = fmax(other->p1.x, other->p2.x);
        const double l1_y_min = fmin(self->p1.y, self->p2.y), l1_y_max = fmax(self->p1.y, self->p2.y);
        const double l2_y_min = fmin(other->p1.y, other->p2.y), l2_y_max = fmax(other->p1.y, other->p2.y);
        return (l1_x_max >= l2_x_max - GEOEPS && l1_x_min <= l2_x_min + GEOEPS && l1_y_max >= l2_y_max - GEOEPS
                && l1_y_min <= l2_y_min + GEOEPS);
    }
    return 0;
}

/* Line::operator==, geometry.h:61-63 */
static int line_eq(const line_t* a, const line_t* b) { return pt_eq(a->p1, b->p1) && pt_eq(a->p2, b->p2); }

/* Line::overlaps, geometry.cpp:56-71 */
static int line_overlaps(const line_t* self, const line_t* other) {
    if (line_eq(self, other)) { return 1; }
    pt pa, pb, pc, pd;
    if (self->p1.x + self->p1.y - (self->p2.x + self->p2.y) >= 0.) {
        pa = self->p2;
        pb = self->p1;
    } else {
        pa = self->p1;
        pb = self->p2;
    }
    if (other->p1.x + other->p1.y - (other->p2.x + other->p2.y) >= 0.) {
        pc = other->p2;
        pd = other->p1;
    } else {
        pc = other->p1;
        pd = other->p2;
    }
    int overlaps = 0;
    if (!(pt_eq(pd, pa) || pt_eq(pb, pc))) {
        if ((pd.x + pd.y - (pa.x + pa.y) >= 0.) && (pb.x + pb.y - (pc.x + pc.y) >= 0.)) { overlaps = 1; }
    }
    return overlaps;
}

/* Line::getRandPoint, geometry.cpp:140-143 */
static pt line_rand_point(const line_t* l, double r1) {
    const double r2 = 1. - r1;
    pt p = { l->p1.x * r1 + l->p2.x * r2, l->p1.y * r1 + l->p2.y * r2 };
    return p;
}

typedef struct {
    pt p1, p2, p3;
} tri_t;

/* Triangle::Triangle validity, geometry.cpp:150-158 */
static int tri_valid(pt p1, pt p2, pt p3) {
    if ((fabs(get_slope(p1, p2)) == fabs(get_slope(p2, p3)) && fabs(get_slope(p2, p3)) == fabs(get_slope(p3, p1)))
        || pt_eq(p1, p2) || pt_eq(p2, p3) || pt_eq(p3, p1)) {
        return 0;
    }
    return 1;
}

/* Triangle::area (Heron on the three line lengths), geometry.cpp:252-259 */
static double tri_area(const tri_t* t) {
    line_t l1, l2, l3;
    line_make(t->p1, t->p2, &l1);
    line_make(t->p2, t->p3, &l2);
    line_make(t->p3, t->p1, &l3);
    const double a = l1.length, b = l2.length, c = l3.length;
    const double p = (a + b + c) / 2.;
    return sqrt(p * (p - a) * (p - b) * (p - c));
}

/* getArea lambda of Triangle::contains(Point): area, or 0 for an invalid triangle, geometry.cpp:190-195 */
static double tri_area_or_zero(pt p1, pt p2, pt p3) {
    if (!tri_valid(p1, p2, p3)) { return 0.; }
    tri_t t = { p1, p2, p3 };
    return tri_area(&t);
}

/* Triangle::contains(Point), geometry.cpp:189-215 */
static int tri_contains_point(const tri_t* t, pt p) {
    const pt bp1 = sub(t->p2, t->p1);
    const pt bp2 = sub(t->p3, t->p1);
    const double d00 = dot(bp1, bp1);
    const double d01 = dot(bp1, bp2);
    const double d11 = dot(bp2, bp2);
    const double denom = d00 * d11 - d01 * d01;
    const pt bp3 = sub(p, t->p1);
    const double d20 = dot(bp3, bp1);
    const double d21 = dot(bp3, bp2);
    const double u = (d11 * d20 - d01 * d21) / denom;
    const double v = (d00 * d21 - d01 * d20) / denom;
    if (u >= GEOEPS && u <= 1. + GEOEPS && v >= GEOEPS && v <= 1. + GEOEPS && u + v <= 1. + GEOEPS) {
        if ((tri_area(t) - tri_area_or_zero(p, t->p2, t->p3) + tri_area_or_zero(t->p1, p, t->p3)
                + tri_area_or_zero(t->p1, t->p2, p))
            < GEOEPS) {
            return 1;
        }
    }
    return 0;
}

/* Triangle::contains(Triangle), geometry.cpp:264-266 */
static int tri_contains_tri(const tri_t* t, const tri_t* o) {
    return tri_contains_point(t, o->p1) || tri_contains_point(t, o->p2) || tri_contains_point(t, o->p3);
}

static void tri_lines(const tri_t* t, line_t* l) {
    line_make(t->p1, t->p2, &l[0]);
    line_make(t->p2, t->p3, &l[1]);
    line_make(t->p3, t->p1, &l[2]);
}

/* Triangle::intersects, geometry.cpp:163-185 */
static int tri_intersects(const tri_t* t, const tri_t* o) {
    line_t a[3], b[3];
    tri_lines(t, a);
    tri_lines(o, b);
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            pt poi;
            if (line_get_intersection(&a[i], &b[j], &poi)) {
                if (!(pt_eq(poi, a[i].p1) || pt_eq(poi, a[i].p2) || pt_eq(poi, b[j].p1) || pt_eq(poi, b[j].p2))) {
                    return 1;
                }
            }
        }
    }
    return 0;
}

/* Triangle::isClockwise, geometry.cpp:217-222 */
static int tri_is_clockwise(const tri_t* t) {
    const double p1_x = t->p2.x - t->p1.x, p1_y = t->p2.y + t->p1.y;
    const double p2_x = t->p3.x - t->p2.x, p2_y = t->p3.y + t->p2.y;
    const double p3_x = t->p1.x - t->p3.x, p3_y = t->p1.y + t->p3.y;
    return p1_x * p1_y + p2_x * p2_y + p3_x * p3_y >= 0.;
}

/* Triangle::getRandPoint (reflection method), geometry.cpp:234-242 */
static pt tri_rand_point(const tri_t* t, double r1, double r2) {
    if (r1 + r2 > 1.) {
        r1 = 1 - r1;
        r2 = 1 - r2;
    }
    pt p = { t->p1.x + (t->p2.x - t->p1.x) * r1 + (t->p3.x - t->p1.x) * r2,
        t->p1.y + (t->p2.y - t->p1.y) * r1 + (t->p3.y - t->p1.y) * r2 };
    return p;
}

/* Triangle::operator==, geometry.cpp:244-249 */
static int tri_eq(const tri_t* a, const tri_t* b) {
#define TRI_HAS(t, p) (pt_eq((p), (t)->p1) || pt_eq((p), (t)->p2) || pt_eq((p), (t)->p3))
    return TRI_HAS(b, a->p1) && TRI_HAS(b, a->p2) && TRI_HAS(b, a->p3);
#undef TRI_HAS
}

/* ------------------------------------------------------------------------ */
/* Model objects */

typedef struct {
    line_t line;
    pt normal;
    int cell; /* the neighbour this surface hands the phonon to */
} trans_t;

typedef struct {
    line_t line;
    pt normal;
    double spec;
    double temp, duration, start;
    int table; /* emit table slot */
} emit_t;

typedef struct {
    line_t line; /* main surface */
    pt normal;
    double spec;
    int n_trans, n_emit;
    trans_t trans[MAX_SUBS];
    emit_t emit[MAX_SUBS];
} csurf_t;

typedef struct {
    tri_t tri;
    int sensor; /* index into sensors */
    csurf_t b[3];
} cell_t;

typedef struct {
    double cdf[NBINS];
    double frac[NBINS];
    double sum;
} table_t;

typedef struct {
    char name[64];
    double la[3], ta[3];
    double w_max_la, w_max_ta, b_l, b_tn, b_tu, b_i, w;
    int full;
    double fw;
    double freq[NBINS], dens_la[NBINS], dens_ta[NBINS], vel_la[NBINS], vel_ta[NBINS];
    int64_t n_temps;
    double* temps;
    double* base_sums;
    /* cache of built tables: key = index * 3 + kind */
    int n_cached;
    int64_t cached_key[64];
    table_t* cached[64];
} material_t;

typedef struct {
    int64_t id;
    int mat;
    double t_init;
    double area;
    double heat_capacity; /* baseEnergy(t_init) */
    const table_t* base;
    const table_t* scatter;
    double t3, t4; /* pow(t_init, 3), pow(t_init, 4) */
    double t_steady; /* SensorController::t_steady_ */
} sensor_t;

typedef struct {
    int cell;
    int kind; /* 0 cell origin, 1 emit surface */
    int b, e; /* boundary and emit sub-surface for kind 1 */
    double energy;
    int64_t count;
    int64_t offset;
} source_t;

struct or_model {
    int64_t num_runs, num_measurements, num_phonons, step_interval;
    double sim_time, t_eq;
    int sim_type, phasor;
    double step_time;
    int64_t step_adjustment; /* ModelSimulator::step_adjustment_ */
    int step_by_step; /* the reference's loop as written: one drift per measurement step */
    uint32_t stream; /* random-stream index: model m of a batch draws the numbers of stream m */
    int planted_fault; /* test hook (or_plant_flight_fault): 0, or a known mistake in the flights */
    int64_t start_step; /* Model::start_step_ */
    int64_t n_record; /* Sensor tally length */
    int n_mat, cap_mat;
    material_t* mats;
    int n_sensors, cap_sensors;
    sensor_t* sensors;
    int n_cells, cap_cells;
    cell_t* cells;
    /* prepared state */
    int prepared;
    double lb, ub, eff_energy;
    int n_sources;
    source_t* sources;
    int64_t total_phonons;
    char err[512];
};

static int set_err(or_model* m, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(m->err, sizeof m->err, fmt, ap);
    va_end(ap);
    return -1;
}

const char* or_error(const or_model* m) { return m ? m->err : "null model"; }

/* Model::Model and Model::setSimulationType, model.cpp:31-70 */
or_model* or_model_new(const double* s, char* err, int errlen) {
    or_model* m = (or_model*)calloc(1, sizeof(or_model));
    m->num_runs = (int64_t)s[0];
    m->num_measurements = (int64_t)s[1];
    m->num_phonons = (int64_t)s[2];
    m->sim_time = s[3];
    m->t_eq = s[4];
    m->sim_type = (int)s[5];
    m->step_interval = (int64_t)s[6];
    m->phasor = s[7] != 0.;
    /* ModelSimulator::ModelSimulator, modelSimulator.cpp:29-37 */
    m->step_time = m->sim_time / (double)m->num_measurements;
    const char* msg = NULL;
    const double n_meas = (double)m->num_measurements;
    if (m->sim_type == SIM_TRANSIENT || m->sim_type == SIM_PERIODIC) {
        m->start_step = (int64_t)(n_meas - n_meas * SS_STEPS_PERCENT);
        if (m->step_interval == 0) { msg = "Step interval of 0 is invalid for transient and periodic simulations."; }
    }
    if (!msg && m->sim_type == SIM_TRANSIENT && m->t_eq == 0) {
        msg = "Transient simulations must be run using the deviational approach.";
    }
    if (!msg && m->sim_type == SIM_SS) {
        m->step_adjustment = (int64_t)(n_meas - n_meas * SS_STEPS_PERCENT);
        if (m->step_interval > 0) { msg = "Step interval > 0 is invalid for steady-state simulations"; }
    }
    /* Model::addSensor recorded length, model.cpp:87-90 */
    m->n_record = (m->sim_type == SIM_SS) ? (int64_t)(n_meas * SS_STEPS_PERCENT) : m->num_measurements;
    if (msg) {
        if (err && errlen > 0) { snprintf(err, (size_t)errlen, "%s", msg); }
        free(m);
        return NULL;
    }
    return m;
}

void or_model_free(or_model* m) {
    if (!m) { return; }
    for (int i = 0; i < m->n_mat; ++i) {
        free(m->mats[i].temps);
        free(m->mats[i].base_sums);
        for (int j = 0; j < m->mats[i].n_cached; ++j) { free(m->mats[i].cached[j]); }
    }
    free(m->mats);
    free(m->sensors);
    free(m->cells);
    free(m->sources);
    free(m);
}

/* Material::getK / getGv, material.cpp:86-95 (pow(x, 2) is x*x) */
static double mat_get_k(double freq, const double* c) {
    const double d = c[1] * c[1] - 4. * c[0] * (c[2] - freq);
    const double a = (-c[1] - sqrt(d)) / (2. * c[0]);
    const double b = (-c[1] + sqrt(d)) / (2. * c[0]);
    return (a < b) ? a : b;
}
static double mat_get_gv(double freq, const double* c) { return 2. * c[0] * mat_get_k(freq, c) + c[1]; }

/* Material::Material, material.cpp:20-51 */
int or_add_material(or_model* m, const char* name, const double* p) {
    for (int i = 0; i < m->n_mat; ++i) {
        if (strcmp(m->mats[i].name, name) == 0) { return set_err(m, "A duplicate material name was detected."); }
    }
    if (m->n_mat == m->cap_mat) {
        m->cap_mat = m->cap_mat ? 2 * m->cap_mat : 4;
        m->mats = (material_t*)realloc(m->mats, sizeof(material_t) * (size_t)m->cap_mat);
    }
    material_t* mt = &m->mats[m->n_mat++];
    memset(mt, 0, sizeof *mt);
    snprintf(mt->name, sizeof mt->name, "%s", name);
    for (int i = 0; i < 3; ++i) {
        mt->la[i] = p[i];
        mt->ta[i] = p[3 + i];
    }
    mt->w_max_la = p[6];
    mt->w_max_ta = p[7];
    mt->b_l = p[8];
    mt->b_tn = p[9];
    mt->b_tu = p[10];
    mt->b_i = p[11];
    mt->w = p[12];
    mt->full = (m->t_eq == 0.); /* inputManager.cpp:44 setFullSimulation */
    mt->fw = fmax(mt->w_max_la, mt->w_max_ta) / (double)NBINS;
    const double pi2 = PSIM_PI * PSIM_PI;
    for (int i = 0; i < NBINS; ++i) {
        const double freq = (double)(2 * i + 1) * mt->fw / 2.;
        mt->freq[i] = freq;
        const double la_gv = mat_get_gv(freq, mt->la);
        mt->vel_la[i] = la_gv;
        const double kl = mat_get_k(freq, mt->la);
        mt->dens_la[i] = kl * kl / 2. / pi2 / la_gv;
        const double ta_gv = mat_get_gv(freq, mt->ta);
        if (!isnan(ta_gv)) {
            mt->vel_ta[i] = ta_gv;
            const double kt = mat_get_k(freq, mt->ta);
            mt->dens_ta[i] = kt * kt / pi2 / ta_gv;
        }
    }
    return 0;
}

/* Material::tauNInv / tauUInv / tauIInv with glibc, material.cpp:207-239 (table builder) */
static void mat_rates_libm(const material_t* mt, double temp, double freq, int pol, double* r) {
    if (pol == POL_LA) {
        r[0] = mt->b_l * freq * freq * pow(temp, 3);
        r[1] = mt->b_l * freq * freq * pow(temp, 3);
    } else {
        r[0] = (freq < mt->w) ? mt->b_tn * freq * pow(temp, 4) : 0.;
        r[1] = (freq >= mt->w) ? mt->b_tu * freq * freq / sinh(HBAR * freq / (temp * BOLTZ)) : 0.;
    }
    r[2] = mt->b_i * pow(freq, 4);
}

/* Material::phononDist, material.cpp:184-204 */
static void mat_phonon_dist(const material_t* mt, double temp, int pol, double* out) {
    const double* dens = (pol == POL_LA) ? mt->dens_la : mt->dens_ta;
    const double cc = HBAR / (BOLTZ * temp);
    for (int i = 0; i < NBINS; ++i) {
        const double freq = mt->freq[i];
        double dist = freq * HBAR / expm1(cc * freq) * mt->fw * dens[i];
        if (!mt->full) { dist *= cc * freq * exp(cc * freq) / (expm1(cc * freq) * temp); }
        out[i] = dist;
    }
}

static double accumulate(const double* a, int n) {
    double s = 0.;
    for (int i = 0; i < n; ++i) { s = s + a[i]; }
    return s;
}

/* Material::buildCumulDist, material.cpp:170-179 */
static void build_cumul(const double* t1, const double* t2, table_t* t) {
    const double cumul_sum = accumulate(t1, NBINS) + accumulate(t2, NBINS);
    t->cdf[0] = (t1[0] + t2[0]) / cumul_sum;
    t->frac[0] = t1[0] / (t1[0] + t2[0]);
    for (int i = 1; i < NBINS; ++i) {
        t->cdf[i] = t->cdf[i - 1] + (t1[i] + t2[i]) / cumul_sum;
        t->frac[i] = t1[i] / (t1[i] + t2[i]);
    }
}

/* Material::initializeTables grid, material.cpp:101-109 */
static void mat_init_grid(material_t* mt, double low, double high) {
    const double interval = (double)TEMP_INTERVAL;
    const int64_t num_steps = (int64_t)((high - low) / interval);
    free(mt->temps);
    free(mt->base_sums);
    mt->n_temps = num_steps + 1;
    mt->temps = (double*)malloc(sizeof(double) * (size_t)mt->n_temps);
    mt->base_sums = (double*)malloc(sizeof(double) * (size_t)mt->n_temps);
    double n = -1.;
    for (int64_t i = 0; i < num_steps; ++i) {
        n += 1;
        mt->temps[i] = low + interval * n;
    }
    mt->temps[num_steps] = high;
    double la[NBINS], ta[NBINS];
    for (int64_t i = 0; i < mt->n_temps; ++i) {
        mat_phonon_dist(mt, mt->temps[i], POL_LA, la);
        mat_phonon_dist(mt, mt->temps[i], POL_TA, ta);
        mt->base_sums[i] = accumulate(la, NBINS) + accumulate(ta, NBINS);
    }
    for (int j = 0; j < mt->n_cached; ++j) { free(mt->cached[j]); }
    mt->n_cached = 0;
}

/* Material::getTempIndex (lower_bound, capped), material.cpp:241-246 */
static int64_t mat_temp_index(const material_t* mt, double temp) {
    int64_t lo = 0, hi = mt->n_temps;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (mt->temps[mid] < temp) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return (lo > mt->n_temps - 1) ? mt->n_temps - 1 : lo;
}

/* Material::initializeTables body for one grid temperature, material.cpp:111-122, with
 * cumulDistEmit :140-150 and cumulDistScatter :152-168 */
static const table_t* mat_table(material_t* mt, int kind, int64_t index) {
    const int64_t key = index * 3 + kind;
    for (int j = 0; j < mt->n_cached; ++j) {
        if (mt->cached_key[j] == key) { return mt->cached[j]; }
    }
    const double temp = mt->temps[index];
    double la[NBINS], ta[NBINS];
    mat_phonon_dist(mt, temp, POL_LA, la);
    mat_phonon_dist(mt, temp, POL_TA, ta);
    table_t* t = (table_t*)malloc(sizeof(table_t));
    if (kind == 1) {
        for (int i = 0; i < NBINS; ++i) {
            la[i] = la[i] * mt->vel_la[i];
            ta[i] = ta[i] * mt->vel_ta[i];
        }
    } else if (kind == 2) {
        double r[3];
        for (int i = 0; i < NBINS; ++i) {
            mat_rates_libm(mt, temp, mt->freq[i], POL_LA, r);
            la[i] = la[i] * (0. + r[0] + r[1] + r[2]);
        }
        for (int i = 0; i < NBINS; ++i) {
            mat_rates_libm(mt, temp, mt->freq[i], POL_TA, r);
            ta[i] = ta[i] * (0. + r[0] + r[1] + r[2]);
        }
    }
    build_cumul(la, ta, t);
    t->sum = accumulate(la, NBINS) + accumulate(ta, NBINS);
    if (mt->n_cached < 64) {
        mt->cached_key[mt->n_cached] = key;
        mt->cached[mt->n_cached++] = t;
    }
    return t;
}

static int find_material(const or_model* m, const char* name) {
    for (int i = 0; i < m->n_mat; ++i) {
        if (strcmp(m->mats[i].name, name) == 0) { return i; }
    }
    return -1;
}

/* Model::addSensor, model.cpp:81-95 */
int or_add_sensor(or_model* m, int64_t id, const char* material, double t_init) {
    for (int i = 0; i < m->n_sensors; ++i) {
        if (m->sensors[i].id == id) { return set_err(m, "Sensor with this ID already exists"); }
    }
    const int mat = find_material(m, material);
    if (mat < 0) { return set_err(m, "unknown material %s", material); }
    if (m->n_sensors == m->cap_sensors) {
        m->cap_sensors = m->cap_sensors ? 2 * m->cap_sensors : 16;
        m->sensors = (sensor_t*)realloc(m->sensors, sizeof(sensor_t) * (size_t)m->cap_sensors);
    }
    sensor_t* s = &m->sensors[m->n_sensors++];
    memset(s, 0, sizeof *s);
    s->id = id;
    s->mat = mat;
    s->t_init = t_init;
    s->t_steady = t_init; /* SensorController ctor, sensorController.cpp:15-20 */
    return 0;
}

/* Surface::Surface normal, surface.cpp:11-16 */
static void csurf_init(csurf_t* c, const line_t* l, double spec, int norm_sign) {
    memset(c, 0, sizeof *c);
    c->line = *l;
    c->normal = line_normal(l, norm_sign);
    c->spec = spec;
}

/* CompositeSurface::verifySurfaceLine, compositeSurface.cpp:68-87; returns 1 if the line lies in
 * the main surface, 0 if not, -1 (CompositeSurfaceError) on an overlap with a sub-surface. */
static int csurf_verify(const csurf_t* c, const line_t* l) {
    if (line_contains_line(&c->line, l)) {
        for (int i = 0; i < c->n_trans; ++i) {
            if (line_overlaps(l, &c->trans[i].line)) { return -1; }
        }
        for (int i = 0; i < c->n_emit; ++i) {
            if (line_overlaps(l, &c->emit[i].line)) { return -1; }
        }
        return 1;
    }
    return 0;
}

/* Cell::setTransitionSurface + CompositeSurface::addTransitionSurface, cell.cpp:126-132,
 * compositeSurface.cpp:37-45 */
static int cell_set_transition(or_model* m, int cell, const line_t* l, int target) {
    cell_t* c = &m->cells[cell];
    const int norm_sign = tri_is_clockwise(&c->tri) ? 1 : -1;
    for (int b = 0; b < 3; ++b) {
        const int v = csurf_verify(&c->b[b], l);
        if (v < 0) { return -1; }
        if (v == 1) {
            if (c->b[b].n_trans >= MAX_SUBS) { return -2; }
            trans_t* t = &c->b[b].trans[c->b[b].n_trans++];
            t->line = *l;
            t->normal = line_normal(l, norm_sign);
            t->cell = target;
            return 1;
        }
    }
    return 0;
}

/* Cell::findTransitionSurface, cell.cpp:81-98 */
static int cell_find_transition(or_model* m, int inc, int other) {
    line_t lines[3], other_lines[3];
    for (int i = 0; i < 3; ++i) {
        lines[i] = m->cells[inc].b[i].line;
        other_lines[i] = m->cells[other].b[i].line;
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            const line_t* l1 = &lines[i];
            const line_t* l2 = &other_lines[j];
            const line_t* l = NULL;
            if (line_contains_line(l1, l2)) {
                l = l2;
            } else if (line_contains_line(l2, l1)) {
                l = l1;
            }
            if (l) {
                const int a = cell_set_transition(m, inc, l, other);
                if (a == -2) { return set_err(m, "too many sub-surfaces on one edge"); }
                const int b = (a == 1) ? cell_set_transition(m, other, l, inc) : 0;
                if (b == -2) { return set_err(m, "too many sub-surfaces on one edge"); }
                if (a != 1 || b != 1) {
                    return set_err(m, "An existing surface conflicts with the location of the incoming surface.");
                }
            }
        }
    }
    return 0;
}

static int find_sensor(const or_model* m, int64_t id) {
    for (int i = 0; i < m->n_sensors; ++i) {
        if (m->sensors[i].id == id) { return i; }
    }
    return -1;
}

/* Model::addCell, model.cpp:98-113, with Cell::Cell cell.cpp:13-19 and Cell::validate :21-25 */
int or_add_cell(or_model* m, const double* t6, int64_t sensor_id, double spec) {
    const int sidx = find_sensor(m, sensor_id);
    if (sidx < 0) { return set_err(m, "Sensor does not exist"); }
    pt p1 = { t6[0], t6[1] }, p2 = { t6[2], t6[3] }, p3 = { t6[4], t6[5] };
    if (!tri_valid(p1, p2, p3)) { return set_err(m, "These 3 points do not allow for a valid triangle"); }
    if (m->n_cells == m->cap_cells) {
        m->cap_cells = m->cap_cells ? 2 * m->cap_cells : 16;
        m->cells = (cell_t*)realloc(m->cells, sizeof(cell_t) * (size_t)m->cap_cells);
    }
    const int inc = m->n_cells++;
    cell_t* c = &m->cells[inc];
    memset(c, 0, sizeof *c);
    c->tri.p1 = p1;
    c->tri.p2 = p2;
    c->tri.p3 = p3;
    c->sensor = sidx;
    /* Cell::buildCompositeSurfaces, cell.cpp:114-124 */
    if (spec < 0.) {
        spec = 0.;
    } else if (spec > 1.) {
        spec = 1.;
    }
    const int norm_sign = tri_is_clockwise(&c->tri) ? 1 : -1;
    line_t l;
    line_make(p1, p2, &l);
    csurf_init(&c->b[0], &l, spec, norm_sign);
    line_make(p2, p3, &l);
    csurf_init(&c->b[1], &l, spec, norm_sign);
    line_make(p3, p1, &l);
    csurf_init(&c->b[2], &l, spec, norm_sign);
    m->sensors[sidx].area += tri_area(&c->tri); /* Sensor::addToArea */
    int identical = 0;
    for (int i = 0; i < m->n_cells; ++i) {
        if (!tri_eq(&m->cells[inc].tri, &m->cells[i].tri)) {
            if (tri_intersects(&m->cells[inc].tri, &m->cells[i].tri)) {
                return set_err(m, "Incoming triangle intersects an existing triangle");
            }
            if (tri_contains_tri(&m->cells[inc].tri, &m->cells[i].tri)
                || tri_contains_tri(&m->cells[i].tri, &m->cells[inc].tri)) {
                return set_err(m, "A triangle is contained within another triangle");
            }
            if (cell_find_transition(m, inc, i) != 0) { return -1; }
        } else if (++identical == 2) {
            return set_err(m, "Duplicate cell detected.");
        }
    }
    return 0;
}

/* Model::setEmitSurface, model.cpp:125-138; Cell::setEmitSurface cell.cpp:29-35;
 * CompositeSurface::addEmitSurface compositeSurface.cpp:20-35 */
int or_add_emit(or_model* m, const double* e) {
    const double temp = e[4], duration = e[5], start = e[6];
    if ((start < 0. || start >= m->sim_time) || (duration < 0. || duration > m->sim_time - start)) {
        return set_err(m, "Transient Surface start_time or duration specifications are invalid.");
    }
    if ((start > 0. || duration < m->sim_time) && m->sim_type != SIM_TRANSIENT) {
        return set_err(m, "Cannot add a transient surface to a non transient simulation.");
    }
    pt p1 = { e[0], e[1] }, p2 = { e[2], e[3] };
    line_t l;
    if (line_make(p1, p2, &l) != 0) { return set_err(m, "Cannot create a line using 2 identical points"); }
    for (int ci = 0; ci < m->n_cells; ++ci) {
        cell_t* c = &m->cells[ci];
        for (int b = 0; b < 3; ++b) {
            const int v = csurf_verify(&c->b[b], &l);
            if (v < 0) {
                return set_err(m, "An existing surface conflicts with the location of the incoming surface.");
            }
            if (v == 1) {
                if (c->b[b].n_emit >= MAX_SUBS) { return set_err(m, "too many sub-surfaces on one edge"); }
                emit_t* es = &c->b[b].emit[c->b[b].n_emit++];
                es->line = l;
                es->normal = c->b[b].normal; /* setNormal(main normal), compositeSurface.cpp:31 */
                es->spec = c->b[b].spec;
                es->temp = temp;
                es->duration = duration;
                es->start = start;
                es->table = -1;
                return 0;
            }
        }
    }
    return set_err(m, "Unable to add emitting surface.");
}

/* libstdc++'s std::reduce / std::transform_reduce order for random-access ranges
 * (bits/numeric: blocks of four summed pairwise, then the tail), which the
 * reference reaches through std::execution::seq. */
static double reduce4(const double* a, int64_t n) {
    double init = 0.;
    int64_t i = 0;
    for (; n - i >= 4; i += 4) {
        const double v1 = a[i] + a[i + 1];
        const double v2 = a[i + 2] + a[i + 3];
        init = init + (v1 + v2);
    }
    for (; i < n; ++i) { init = init + a[i]; }
    return init;
}

/* Cell::getInitEnergy, cell.cpp:38-41 */
static double cell_init_energy(const or_model* m, int ci, double t_eq) {
    const cell_t* c = &m->cells[ci];
    const sensor_t* s = &m->sensors[c->sensor];
    const double init_energy = tri_area(&c->tri) * s->heat_capacity;
    const double t_init = (m->sim_type == SIM_SS) ? s->t_steady : s->t_init; /* getInitTemp */
    return (t_eq == 0.) ? init_energy : init_energy * fabs(t_init - t_eq);
}

static double emit_energy_value(material_t* mt, double temp) {
    return mat_table(mt, 1, mat_temp_index(mt, temp))->sum; /* Material::emitEnergy */
}

/* Cell::getEmitEnergy, cell.cpp:45-63 */
static double cell_emit_energy(or_model* m, int ci, double t_eq) {
    cell_t* c = &m->cells[ci];
    material_t* mt = &m->mats[m->sensors[c->sensor].mat];
    double emit_energy = 0.;
    for (int b = 0; b < 3; ++b) {
        double vals[MAX_SUBS];
        for (int e = 0; e < c->b[b].n_emit; ++e) {
            const emit_t* es = &c->b[b].emit[e];
            const double energy = es->line.length * es->duration * emit_energy_value(mt, es->temp) / 4.;
            vals[e] = (t_eq == 0.) ? energy : energy * fabs(es->temp - t_eq);
        }
        emit_energy += reduce4(vals, c->b[b].n_emit);
    }
    return emit_energy;
}

/* Model::runSimulation preamble (model.cpp:141-153) and ModelSimulator::initPhononBuilders
 * (modelSimulator.cpp:43-85), with the Bernoulli draw of source s taken from the budget stream. */
int or_set_step_by_step(or_model* m, int on) {
    m->step_by_step = on != 0;
    return 0;
}

/* Test hook: plants a known mistake in the tallies of the boundaries a flight passes, so that a
 * test can show that its comparison detects it.  1: tallied one index late; 2: tallied into the
 * next sensor.  0 removes it. */
int or_plant_flight_fault(or_model* m, int kind) {
    if (kind < 0 || kind > 2) { return set_err(m, "unknown planted fault"); }
    m->planted_fault = kind;
    return 0;
}

/* The random-stream index of every draw (budget, emission, transport), as a batch of models gives
 * model m stream m with its phonon ids counted from 0; the budget must be prepared again. */
int or_set_stream(or_model* m, int stream) {
    if (stream < 0 || stream > 0xffff) { return set_err(m, "stream index out of range"); }
    m->stream = (uint32_t)stream;
    m->prepared = 0;
    return 0;
}

int or_prepare(or_model* m, uint64_t seed) {
    if (m->n_cells == 0 || m->n_sensors == 0) { return set_err(m, "model has no cells or sensors"); }
    /* Model::setTemperatureBounds, model.cpp:203-218 */
    double tmin = 0, tmax = 0;
    int first = 1;
    for (int ci = 0; ci < m->n_cells; ++ci) {
        const sensor_t* s = &m->sensors[m->cells[ci].sensor];
        const double t = (m->sim_type == SIM_SS) ? s->t_steady : s->t_init;
        if (first || t < tmin) { tmin = t; }
        if (first || t > tmax) { tmax = t; }
        first = 0;
        for (int b = 0; b < 3; ++b) {
            for (int e = 0; e < m->cells[ci].b[b].n_emit; ++e) {
                const double te = m->cells[ci].b[b].emit[e].temp;
                if (te < tmin) { tmin = te; }
                if (te > tmax) { tmax = te; }
            }
        }
    }
    const double bound = m->phasor ? PHASOR_TEMP_BOUND_EPS : TEMP_BOUND_EPS;
    m->lb = fmax(tmin - bound, 0.);
    m->ub = tmax + bound;
    /* Model::initializeMaterialTables, model.cpp:221-227; SensorController::updateTables
     * sensorController.cpp:38-50 */
    for (int i = 0; i < m->n_mat; ++i) { mat_init_grid(&m->mats[i], tmin, tmax); }
    for (int i = 0; i < m->n_sensors; ++i) {
        sensor_t* s = &m->sensors[i];
        material_t* mt = &m->mats[s->mat];
        const int64_t idx = mat_temp_index(mt, s->t_init);
        s->base = mat_table(mt, 0, idx);
        s->heat_capacity = mt->base_sums[idx];
        s->scatter = mat_table(mt, 2, idx);
        s->t3 = pow(s->t_init, 3);
        s->t4 = pow(s->t_init, 4);
    }
    /* Model::getTotalInitialEnergy (transform_reduce, seq), model.cpp:196-201 */
    double* per_cell = (double*)calloc((size_t)m->n_cells, sizeof(double));
    for (int ci = 0; ci < m->n_cells; ++ci) {
        per_cell[ci] = cell_init_energy(m, ci, m->t_eq) + cell_emit_energy(m, ci, m->t_eq);
    }
    const double total_energy = reduce4(per_cell, m->n_cells);
    free(per_cell);
    m->eff_energy = total_energy / (double)m->num_phonons;

    /* Sources in the reference's builder order: per cell, the cell-origin builder, then every
     * emit sub-surface of its three boundaries. */
    free(m->sources);
    size_t cap = (size_t)m->n_cells;
    for (int ci = 0; ci < m->n_cells; ++ci) {
        for (int b = 0; b < 3; ++b) { cap += (size_t)m->cells[ci].b[b].n_emit; }
    }
    m->sources = (source_t*)calloc(cap, sizeof(source_t));
    m->n_sources = 0;
    int64_t offset = 0;
    uint32_t sidx = 0;
    for (int ci = 0; ci < m->n_cells; ++ci) {
        cell_t* c = &m->cells[ci];
        material_t* mt = &m->mats[m->sensors[c->sensor].mat];
        for (int k = -1; k < 3 * MAX_SUBS; ++k) {
            double energy;
            int b = -1, e = -1;
            if (k < 0) {
                energy = cell_init_energy(m, ci, m->t_eq);
            } else {
                b = k / MAX_SUBS;
                e = k % MAX_SUBS;
                if (e >= c->b[b].n_emit) { continue; }
                const emit_t* es = &c->b[b].emit[e];
                const double factor = emit_energy_value(mt, es->temp) * es->duration * es->line.length / 4.;
                energy = (m->t_eq == 0.) ? factor : factor * fabs(m->t_eq - es->temp);
            }
            /* getPhonons lambda, modelSimulator.cpp:44-49 */
            double ipart = 0;
            const double frac = modf(energy / m->eff_energy, &ipart);
            int64_t n = (int64_t)ipart;
            psim_stream st;
            psim_stream_init(&st, seed, psim_stream_tag(PSIM_TAG_BUDGET, m->stream), sidx++, 0);
            if (psim_urand(&st) < frac) { ++n; }
            source_t* src = &m->sources[m->n_sources++];
            src->cell = ci;
            src->kind = (k < 0) ? 0 : 1;
            src->b = b;
            src->e = e;
            src->energy = energy;
            src->count = n;
            src->offset = offset;
            offset += n;
            if (k >= 0) { c->b[b].emit[e].table = 0; }
        }
    }
    m->total_phonons = offset;
    m->prepared = 1;
    return 0;
}

int or_num_sensors(const or_model* m) { return m->n_sensors; }
int64_t or_num_record(const or_model* m) { return m->n_record; }
int64_t or_total_phonons(const or_model* m) { return m->total_phonons; }
double or_eff_energy(const or_model* m) { return m->eff_energy; }
int or_num_sources(const or_model* m) { return m->n_sources; }

int or_sources(const or_model* m, int64_t* cell, int32_t* kind, double* energy, int64_t* count) {
    for (int i = 0; i < m->n_sources; ++i) {
        cell[i] = m->sources[i].cell;
        kind[i] = m->sources[i].kind;
        energy[i] = m->sources[i].energy;
        count[i] = m->sources[i].count;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Phonon and the event loop */

typedef struct {
    int sign;
    double lifetime;
    int64_t lifestep;
    double px, py, dx, dy;
    int64_t freq_index;
    double freq, velocity;
    int polar;
    int cell; /* -1 = outside (absorbed) */
} phonon_t;

typedef struct {
    int64_t* counts;
    double* flux;
    int64_t counters[OR_NUM_COUNTERS];
} tally_t;

/* Material::freqIndex (bisection returning `high`), material.cpp:64-75, given its two draws:
 * `rand` for the bisection and u2 for the polarization. */
static void freq_index_draws(const table_t* t, double rand, double u2, int64_t* index, int* polar) {
    int64_t low = 0, high = NBINS - 1;
    int64_t mid = low + (high - low) / 2;
    while (high - low > 1) {
        if (rand < t->cdf[mid]) {
            high = mid;
        } else {
            low = mid;
        }
        mid = (low + high) / 2;
    }
    *index = high;
    *polar = (u2 <= t->frac[high]) ? POL_LA : POL_TA;
}

static void freq_index(const table_t* t, psim_stream* st, int64_t* index, int* polar) {
    double rand, u2; /* the bisection's draw and the branch draw, one block */
    psim_urand2(st, &rand, &u2);
    freq_index_draws(t, rand, u2, index, polar);
}

/* SensorController::initialUpdate / scatterUpdate with Material::getFreq / getVel and
 * Phonon::scatterUpdate, sensorController.cpp:28-36,52-55; material.cpp:77-84 */
static void resample(const material_t* mt, const table_t* t, phonon_t* p, psim_stream* st) {
    int64_t index;
    int polar;
    freq_index(t, st, &index, &polar);
    const double freq = mt->full ? mt->freq[index] : mt->freq[index] + (2. * psim_urand(st) - 1.) * mt->fw / 2.;
    p->freq_index = index;
    p->freq = freq;
    p->velocity = (polar == POL_LA) ? mt->vel_la[index] : mt->vel_ta[index];
    p->polar = polar;
}

/* Phonon::setRandDirection, phonon.cpp:28-31 */
static void set_rand_direction(phonon_t* p, psim_stream* st) {
    double u1, u2;
    psim_urand2(st, &u1, &u2);
    p->dx = 2. * u1 - 1.;
    p->dy = sqrt(1. - p->dx * p->dx) * psim_cos2pi(u2);
}

/* Surface::redirectPhonon (Lambertian about the normal), surface.cpp:23-30 */
static void redirect(pt n, phonon_t* p, psim_stream* st) {
    double rand, u2;
    psim_urand2(st, &rand, &u2);
    const double new_dx = sqrt(rand);
    const double new_dy = sqrt(1. - rand) * psim_cos2pi(u2);
    p->dx = n.x * new_dx - n.y * new_dy;
    p->dy = n.y * new_dx + n.x * new_dy;
}

/* Surface::boundaryHandlePhonon, surface.cpp:32-44 */
static void boundary_handle(pt n, double spec, phonon_t* p, psim_stream* st) {
    if (spec == 1. || psim_urand(st) < spec) {
        const double dx = p->dx, dy = p->dy;
        const double new_dx = -dx * n.x - dy * n.y;
        const double new_dy = -dx * n.y + dy * n.x;
        p->dx = n.x * new_dx - n.y * new_dy;
        p->dy = n.y * new_dx + n.x * new_dy;
    } else {
        redirect(n, p, st);
    }
}

/* Cell::handleSurfaceCollision -> CompositeSurface::handlePhonon -> Transition/Emit/boundary
 * handling, cell.cpp:103-108, compositeSurface.cpp:47-66, surface.cpp:61-65,71-109 */
static void handle_surface_collision(const or_model* m, phonon_t* p, pt poi, psim_stream* st) {
    const cell_t* c = &m->cells[p->cell];
    for (int b = 0; b < 3; ++b) {
        const csurf_t* cs = &c->b[b];
        if (!point_on_line(&cs->line, poi)) { continue; }
        for (int i = 0; i < cs->n_trans; ++i) {
            const trans_t* ts = &cs->trans[i];
            if (!point_on_line(&ts->line, poi)) { continue; }
            const int mat_p = m->sensors[c->sensor].mat;
            const int mat_t = m->sensors[m->cells[ts->cell].sensor].mat;
            if (mat_p == mat_t) {
                p->cell = ts->cell;
            } else {
                const material_t* mt = &m->mats[mat_t];
                const double max_freq = (p->polar == POL_LA) ? mt->w_max_la : mt->w_max_ta;
                if (p->freq > max_freq) {
                    redirect(ts->normal, p, st);
                } else {
                    p->cell = ts->cell;
                }
            }
            return;
        }
        for (int i = 0; i < cs->n_emit; ++i) {
            const emit_t* es = &cs->emit[i];
            if (!point_on_line(&es->line, poi)) { continue; }
            const double phonon_time = (double)p->lifestep * m->step_time;
            if (phonon_time < es->start || phonon_time + m->step_time > es->start + es->duration) {
                boundary_handle(es->normal, es->spec, p, st);
            } else {
                p->cell = -1;
            }
            return;
        }
        boundary_handle(cs->normal, cs->spec, p, st);
        return;
    }
}

/* getTime lambda of nextImpact, modelSimulator.cpp:96-98 */
static double impact_time_axis(double start, double end, double velocity, double max_time) {
    return (velocity > VELOCITY_EPS || velocity < -VELOCITY_EPS) ? (end - start) / velocity : max_time;
}

/* ModelSimulator::nextImpact, modelSimulator.cpp:87-122; returns 1 and the impact time on a hit */
static int next_impact(const or_model* m, phonon_t* p, double time, double* out_time, psim_stream* st,
    tally_t* ta) {
    (void)st;  /* nextImpact draws nothing */
    const double vx = p->dx * p->velocity;
    const double vy = p->dy * p->velocity;
    const pt start = { p->px, p->py };
    const pt end = { p->px + time * vx, p->py + time * vy };
    if (pt_eq(start, end)) { return 0; }
    line_t path;
    line_make(start, end, &path);
    const cell_t* c = &m->cells[p->cell];
    int hit = 0;
    pt impact = { 0, 0 };
    for (int b = 0; b < 3; ++b) {
        pt poi;
        if (line_get_intersection(&c->b[b].line, &path, &poi) && !pt_eq(poi, start)) {
            const double tx = impact_time_axis(start.x, poi.x, vx, time);
            const double ty = impact_time_axis(start.y, poi.y, vy, time);
            const double it = (tx <= ty) ? tx : ty;
            if (it <= time) {
                time = it;
                impact = poi;
                hit = 1;
            }
        }
    }
    if (hit) {
        p->px = impact.x;
        p->py = impact.y;
        ta->counters[OR_CNT_IMPACTS]++;
        *out_time = time;
        return 1;
    }
    return 0;
}

/* A flight: the phonon's step counter, the boundary its flight ends at (`cap`), the loop's
 * event count and the flight's start time. */
typedef struct {
    int64_t* step;
    int* events;
    int64_t cap;
    double age;
    int step_by_step;
} flight_t;

static void tally(const or_model* m, const phonon_t* p, int64_t idx, tally_t* ta);
static void tally_at(const or_model* m, const phonon_t* p, int s, int64_t idx, tally_t* ta);

/* The measurement events of the boundaries a flight passed up to time t
 * (modelSimulator.cpp:182-188): *step advances to the last boundary k < cap whose time
 * k * sim_time / N (the reference's step_times_) is at or before t, never backwards, and the
 * stream moves to the new step.  Every boundary passed is a phonon-step, and from
 * step_adjustment on it is tallied with the phonon's sensor and velocity of that moment
 * (Sensor::updateHeatParams, sensor.cpp:43-52).  Returns 1 when the step moved. */
static int advance_step(const or_model* m, phonon_t* p, int64_t* step, int64_t cap, double t, psim_stream* st,
    tally_t* ta) {
    const int64_t N = m->num_measurements;
    int64_t k = *step;
    /* a start at or below the answer: floor(t / dt) - 1 lies a whole step below t */
    const int64_t est = (int64_t)(t * ((double)N / m->sim_time)) - 1;
    if (est > k) { k = (est < cap - 1) ? est : cap - 1; }
    while (k + 1 < cap && (double)(k + 1) * m->sim_time / (double)N <= t) { ++k; }
    if (k == *step) { return 0; }
    for (int64_t j = (*step + 1 > m->step_adjustment ? *step + 1 : m->step_adjustment); j <= k; ++j) {
        if (m->planted_fault == 1) { /* test hook: one index late */
            tally(m, p, j - m->step_adjustment + 1, ta);
        } else if (m->planted_fault == 2) { /* test hook: the next sensor */
            tally_at(m, p, (m->cells[p->cell].sensor + 1) % m->n_sensors, j - m->step_adjustment, ta);
        } else {
            tally(m, p, j - m->step_adjustment, ta);
        }
    }
    ta->counters[OR_CNT_PHONON_STEPS] += k - *step;
    *step = k;
    p->lifestep = k;
    psim_stream_set_step(st, (uint32_t)k);
    return 1;
}

/* ModelSimulator::handleImpacts, modelSimulator.cpp:205-225.  Returns 0 when the phonon left
 * the system (std::nullopt), 1 otherwise with *drifted set. */
static int handle_impacts(const or_model* m, phonon_t* p, double drift_time, int sensor, double* drifted_out,
    int* relocated, psim_stream* st, tally_t* ta, const flight_t* fl) {
    double t;
    int hit = next_impact(m, p, drift_time, &t, st, ta);
    double drifted = 0.;
    int collisions = 0;
    while (hit) {
        /* the boundaries passed before the impact see the old direction; the surface sees the
         * step of the moment of impact */
        if (advance_step(m, p, fl->step, fl->cap, fl->age + (drifted + t), st, ta)) {
            *fl->events = 0;
            collisions = 0; /* the reference counts collisions per drift, and a drift ends at a boundary */
        }
        const pt poi = { p->px, p->py };
        handle_surface_collision(m, p, poi, st);
        if (p->cell < 0) { return 0; }
        drifted += t;
        if (++collisions > MAX_COLLISIONS) {
            double r1, r2;
            psim_urand2(st, &r1, &r2);
            const pt q = tri_rand_point(&m->cells[p->cell].tri, r1, r2); /* Phonon::setRandPoint */
            p->px = q.x;
            p->py = q.y;
            ta->counters[OR_CNT_RELOCATIONS]++;
            /* the phonon stays at the new point for the rest of the reference's drift: up to the
             * next boundary, or the flight's end if that comes first */
            const double frozen = (double)(*fl->step + 1) * m->sim_time / (double)m->num_measurements - fl->age;
            *drifted_out = (!fl->step_by_step && frozen < drift_time) ? frozen : drift_time;
            *relocated = 1;
            return 1;
        }
        if (m->cells[p->cell].sensor != sensor) {
            *drifted_out = drifted;
            return 1;
        }
        hit = next_impact(m, p, drift_time - drifted, &t, st, ta);
    }
    if (p->cell < 0) { return 0; }
    *drifted_out = drifted;
    return 1;
}

/* Material::relaxRates at the sensor temperature with the shared transcendentals,
 * material.cpp:54-56,207-239 */
static void transport_rates(const or_model* m, const phonon_t* p, double* r) {
    const sensor_t* s = &m->sensors[m->cells[p->cell].sensor];
    const material_t* mt = &m->mats[s->mat];
    const double freq = p->freq;
    if (p->polar == POL_LA) {
        r[0] = mt->b_l * freq * freq * s->t3;
        r[1] = mt->b_l * freq * freq * s->t3;
    } else {
        r[0] = (freq < mt->w) ? mt->b_tn * freq * s->t4 : 0.;
        r[1] = (freq >= mt->w) ? mt->b_tu * freq * freq / psim_sinh(HBAR * freq / (s->t_init * BOLTZ)) : 0.;
    }
    r[2] = mt->b_i * psim_pow4(freq);
}

/* ModelSimulator::scatter, modelSimulator.cpp:124-137, with its draw `rand` taken by the caller
 * (together with the next scattering-time draw, one block). */
static void scatter(const or_model* m, phonon_t* p, const double* r, double rand, psim_stream* st) {
    const double tau_inv = 0. + r[0] + r[1] + r[2];
    if (rand <= (r[0] + r[1]) / tau_inv) {
        const sensor_t* s = &m->sensors[m->cells[p->cell].sensor];
        resample(&m->mats[s->mat], s->scatter, p, st);
        if (rand > r[0] / tau_inv) { set_rand_direction(p, st); }
    } else if (r[2] > 0.) {
        set_rand_direction(p, st);
    }
}

/* Sensor::updateHeatParams, sensor.cpp:43-52 */
static void tally_at(const or_model* m, const phonon_t* p, int s, int64_t idx, tally_t* ta) {
    if (idx >= m->n_record) { return; }
    const double vx = p->dx * p->velocity;
    const double vy = p->dy * p->velocity;
    const int64_t k = (int64_t)s * m->n_record + idx;
    ta->counts[k] += p->sign;
    ta->flux[2 * k] += vx * (double)p->sign;
    ta->flux[2 * k + 1] += vy * (double)p->sign;
    ta->counters[OR_CNT_TALLIES]++;
}

static void tally(const or_model* m, const phonon_t* p, int64_t idx, tally_t* ta) {
    tally_at(m, p, m->cells[p->cell].sensor, idx, ta);
}

/* ModelSimulator::simulatePhonon, modelSimulator.cpp:139-198, with flights in place of one
 * drift per measurement step.
 *
 * A measurement event changes nothing about the phonon's motion (modelSimulator.cpp:182-188:
 * the step advances and, from step_adjustment_ on, the sensor records the phonon's sign and
 * velocity), so the phonon does not stop at each boundary.  Its flight ends at its scattering
 * time, at an impact that changes sensor, or at its cap: boundary step_adjustment - 1 in steady
 * state while the step lies below it (the recorded steps start there), else boundary N, where
 * it dies; from `stop` on it also ends at every flight_cap-th boundary after `stop`, where
 * flight_cap = psim_flight_cap(number of sensors) is the engine's rule (psim_gpu.h).  The
 * boundaries a flight passes are handled as events in time order: at each impact
 * before the surface acts (so they see the direction before it and the surface sees the step
 * of the impact) and at the end of the flight (advance_step).  Every draw therefore comes from
 * the stream of the step of its moment. */
static void simulate_phonon(const or_model* m, phonon_t* p, uint64_t seed, uint32_t id, tally_t* ta,
    int64_t* fate_step, int32_t* fate_cause, double* fate_xy) {
    const int64_t N = m->num_measurements;
    double phonon_age = p->lifetime;
    int64_t step = (int64_t)(phonon_age / m->step_time);
    int cause = 0;
    if (step >= N) {
        ta->counters[OR_CNT_NOT_STARTED]++;
        cause = 3;
    } else {
        p->lifestep = step;
        psim_stream st;
        psim_stream_init(&st, seed, psim_stream_tag(PSIM_TAG_TRANSPORT, m->stream), id, (uint32_t)step);
        double relax[3] = { 0, 0, 0 };
        double time_to_scatter = 0.;
        double time_to_measurement = 0.;
        int events = 0;
        int scatter_pending = 0; /* the scattering-time draw after a scatter is next_draw */
        double next_draw = 0.;
        const int64_t stop = m->step_adjustment >= 2 ? m->step_adjustment - 1 : 0;

        for (;;) {
            if (++events > MAX_EVENTS_PER_STEP) {
                ta->counters[OR_CNT_STALLED]++;
                cause = 4;
                break;
            }
            if (time_to_scatter <= 0.) {
                /* after a scatter, the second double of the scatter's block; else a block's first */
                const double u = scatter_pending ? next_draw : psim_urand(&st);
                transport_rates(m, p, relax);
                time_to_scatter = SCALING_FACTOR * -psim_log(u) / (0. + relax[0] + relax[1] + relax[2]);
            }
            scatter_pending = 0;
            int64_t cap;
            if (m->step_by_step) { /* modelSimulator.cpp:159: the next boundary, kept until it is reached */
                cap = step + 1;
                if (time_to_measurement <= 0.) {
                    time_to_measurement = (double)(step + 1) * m->sim_time / (double)N - phonon_age;
                }
            } else {
                if (step < stop) {
                    cap = stop;
                } else {
                    /* from `stop` on, every flight_cap-th boundary (psim_flight_cap, psim_gpu.h) */
                    const int64_t fc = psim_flight_cap(m->n_sensors);
                    cap = stop + fc * ((step - stop) / fc + 1);
                    if (cap > N) { cap = N; }
                }
                time_to_measurement = (double)cap * m->sim_time / (double)N - phonon_age;
            }
            double drift_time = (time_to_measurement < time_to_scatter) ? time_to_measurement : time_to_scatter;
            const int sensor = m->cells[p->cell].sensor;
            double drifted;
            int relocated = 0;
            const flight_t fl = { &step, &events, cap, phonon_age, m->step_by_step };
            if (handle_impacts(m, p, drift_time, sensor, &drifted, &relocated, &st, ta, &fl)) {
                if (m->cells[p->cell].sensor != sensor || relocated) { drift_time = drifted; }
                const double factor = p->velocity * (drift_time - drifted); /* Phonon::drift */
                p->px += p->dx * factor;
                p->py += p->dy * factor;
                phonon_age += drift_time;
                time_to_measurement -= drift_time;
                time_to_scatter -= drift_time;
                if (advance_step(m, p, &step, cap, phonon_age, &st, ta)) { events = 0; }
                if (time_to_measurement == 0. && cap == N) { /* boundary N: the run is over */
                    ta->counters[OR_CNT_PHONON_STEPS]++;
                    ta->counters[OR_CNT_EXPIRED]++;
                    cause = 2;
                    break;
                } else if (time_to_measurement == 0.) { /* boundary `cap` */
                    ta->counters[OR_CNT_PHONON_STEPS]++;
                    events = 0;
                    step = cap;
                    p->lifestep = step;
                    psim_stream_set_step(&st, (uint32_t)step);
                    if (step >= m->step_adjustment) { tally(m, p, step - m->step_adjustment, ta); }
                } else {
                    if (!m->phasor && time_to_scatter == 0.) {
                        ta->counters[OR_CNT_SCATTERS]++;
                        double u;
                        psim_urand2(&st, &u, &next_draw);
                        scatter_pending = 1;
                        scatter(m, p, relax, u, &st);
                    } else if (!relocated) { /* a new sensor: a new scattering time */
                        time_to_scatter = 0.;
                    }
                }
            } else {
                ta->counters[OR_CNT_ABSORBED]++;
                cause = 1;
                break;
            }
        }
    }
    if (fate_step) {
        fate_step[id] = (cause == 3) ? step : p->lifestep;
        fate_cause[id] = cause;
        fate_xy[2 * (int64_t)id] = p->px;
        fate_xy[2 * (int64_t)id + 1] = p->py;
    }
}

/* Phonon builders, phononBuilder.cpp:6-49, with EmitSurface::getPhononTime surface.cpp:67-69 */
static void build_phonon(const or_model* m, int64_t id, uint64_t seed, phonon_t* p) {
    int lo = 0, hi = m->n_sources - 1;
    while (lo < hi) { /* last source with offset <= id */
        const int mid = (lo + hi + 1) / 2;
        if (m->sources[mid].offset <= id) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    while (lo < m->n_sources - 1 && m->sources[lo].count == 0) { ++lo; }
    const source_t* src = &m->sources[lo];
    const cell_t* c = &m->cells[src->cell];
    const sensor_t* s = &m->sensors[c->sensor];
    material_t* mt = &m->mats[s->mat];
    psim_stream st;
    psim_stream_init(&st, seed, psim_stream_tag(PSIM_TAG_EMIT, m->stream), (uint32_t)id, 0);
    memset(p, 0, sizeof *p);
    p->cell = src->cell;
    if (src->kind == 0) { /* CellOriginBuilder::operator() */
        const double t_init = (m->sim_type == SIM_SS) ? s->t_steady : s->t_init;
        p->sign = (t_init > m->t_eq) ? 1 : -1;
        p->lifetime = 0.;
        resample(mt, s->base, p, &st);
        double r1, r2;
        psim_urand2(&st, &r1, &r2);
        const pt q = tri_rand_point(&c->tri, r1, r2);
        p->px = q.x;
        p->py = q.y;
        set_rand_direction(p, &st);
    } else { /* SurfaceOriginBuilder::operator() and PhasorBuilder::operator() */
        const emit_t* es = &c->b[src->b].emit[src->e];
        p->sign = (es->temp > m->t_eq) ? 1 : -1;
        p->lifetime = es->start + es->duration * psim_urand(&st);
        resample(mt, mat_table(mt, 1, mat_temp_index(mt, es->temp)), p, &st);
        const pt q = line_rand_point(&es->line, psim_urand(&st));
        p->px = q.x;
        p->py = q.y;
        redirect(es->normal, p, &st);
        if (m->phasor) {
            p->freq_index = 1;
            p->freq = 1;
            p->velocity = 1000;
            p->polar = POL_LA;
            p->dx = es->normal.x;
            p->dy = es->normal.y;
        }
    }
}

typedef struct {
    const or_model* m;
    uint64_t seed;
    int64_t begin, end;
    tally_t ta;
    int64_t* fate_step;
    int32_t* fate_cause;
    double* fate_xy;
} worker_t;

static void* worker_main(void* arg) {
    worker_t* w = (worker_t*)arg;
    for (int64_t id = w->begin; id < w->end; ++id) {
        phonon_t p;
        build_phonon(w->m, id, w->seed, &p);
        simulate_phonon(w->m, &p, w->seed, (uint32_t)id, &w->ta, w->fate_step, w->fate_cause, w->fate_xy);
    }
    return NULL;
}

int or_run(or_model* m, uint64_t seed, int nthreads, int64_t* counts, double* flux, int64_t* counters,
    int64_t* fate_step, int32_t* fate_cause, double* fate_xy) {
    if (!m->prepared) { return set_err(m, "or_prepare must be called first"); }
    /* Emit tables must exist before threads read them (mat_table caches lazily). */
    for (int ci = 0; ci < m->n_cells; ++ci) {
        for (int b = 0; b < 3; ++b) {
            for (int e = 0; e < m->cells[ci].b[b].n_emit; ++e) {
                material_t* mt = &m->mats[m->sensors[m->cells[ci].sensor].mat];
                (void)mat_table(mt, 1, mat_temp_index(mt, m->cells[ci].b[b].emit[e].temp));
            }
        }
    }
    if (nthreads < 1) { nthreads = 1; }
    const int64_t S = m->n_sensors, R = m->n_record;
    worker_t* ws = (worker_t*)calloc((size_t)nthreads, sizeof(worker_t));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    const int64_t n = m->total_phonons;
    for (int t = 0; t < nthreads; ++t) {
        ws[t].m = m;
        ws[t].seed = seed;
        ws[t].begin = n * t / nthreads;
        ws[t].end = n * (t + 1) / nthreads;
        ws[t].ta.counts = (int64_t*)calloc((size_t)(S * R), sizeof(int64_t));
        ws[t].ta.flux = (double*)calloc((size_t)(S * R * 2), sizeof(double));
        ws[t].fate_step = fate_step;
        ws[t].fate_cause = fate_cause;
        ws[t].fate_xy = fate_xy;
    }
    for (int t = 1; t < nthreads; ++t) { pthread_create(&th[t], NULL, worker_main, &ws[t]); }
    worker_main(&ws[0]);
    for (int t = 1; t < nthreads; ++t) { pthread_join(th[t], NULL); }
    memset(counts, 0, sizeof(int64_t) * (size_t)(S * R));
    memset(flux, 0, sizeof(double) * (size_t)(S * R * 2));
    memset(counters, 0, sizeof(int64_t) * OR_NUM_COUNTERS);
    for (int t = 0; t < nthreads; ++t) {
        for (int64_t k = 0; k < S * R; ++k) {
            counts[k] += ws[t].ta.counts[k];
            flux[2 * k] += ws[t].ta.flux[2 * k];
            flux[2 * k + 1] += ws[t].ta.flux[2 * k + 1];
        }
        for (int c = 0; c < OR_NUM_COUNTERS; ++c) { counters[c] += ws[t].ta.counters[c]; }
        free(ws[t].ta.counts);
        free(ws[t].ta.flux);
    }
    free(ws);
    free(th);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* SensorInterpreter, sensorInterpreter.cpp:19-112 */

/* inversion lambda: bisection on theoreticalEnergy(T) * area, sensorInterpreter.cpp:88-97 */
static double invert_energy(const or_model* m, const material_t* mt, double area, double energy) {
    double temp = 0., ub = m->ub, lb = m->lb;
    int iter = 0;
    while ((ub - lb >= INV_EPS) && (++iter != INV_MAX_ITERS)) {
        temp = (ub + lb) / 2.;
        const double de = (mt->base_sums[mat_temp_index(mt, temp)] * area) - energy;
        if (de < 0.) {
            lb = temp;
        } else {
            ub = temp;
        }
    }
    return temp;
}

/* SensorInterpreter::findTemperature, sensorInterpreter.cpp:80-112 */
static void find_temperature(const or_model* m, double eff_energy, int si, const int64_t* energies, int64_t n,
    double* temps) {
    const sensor_t* s = &m->sensors[si];
    const material_t* mt = &m->mats[s->mat];
    for (int64_t i = 0; i < n; ++i) {
        const double energy = eff_energy * (double)energies[i];
        if (m->t_eq != 0.) {
            temps[i] = energy / (s->area * s->heat_capacity) + m->t_eq;
        } else {
            temps[i] = invert_energy(m, mt, s->area, energy);
        }
    }
}

static void avg_and_stderr(const double* data, int64_t n, double* avg_out, double* err_out) {
    const double size = (double)n;
    const double avg = reduce4(data, n) / size;
    double* sq = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) { sq[i] = (avg - data[i]) * (avg - data[i]); }
    const double std_dev = sqrt(reduce4(sq, n) / size);
    free(sq);
    *avg_out = avg;
    *err_out = std_dev / sqrt(size);
}

/* Effective energy the reference uses for its results.  After the run, Model::runSimulation
 * calls resetRequired() (which sets SensorController::t_steady_ to each sensor's final
 * temperature, sensorController.cpp:57-62) and then refresh() (model.cpp:146-151,170), so in
 * steady state the interpreter's energy per phonon is recomputed with
 * getInitTemp() == t_steady_ (sensorController.h:63-67).  final_temp[s] must hold getFinalTemp. */
static double post_run_eff_energy(or_model* m, const double* final_temp) {
    if (m->sim_type != SIM_SS) { return m->eff_energy; }
    double* saved = (double*)malloc(sizeof(double) * (size_t)m->n_sensors);
    for (int si = 0; si < m->n_sensors; ++si) {
        saved[si] = m->sensors[si].t_steady;
        m->sensors[si].t_steady = final_temp[si];
    }
    double* per_cell = (double*)calloc((size_t)m->n_cells, sizeof(double));
    for (int ci = 0; ci < m->n_cells; ++ci) {
        per_cell[ci] = cell_init_energy(m, ci, m->t_eq) + cell_emit_energy(m, ci, m->t_eq);
    }
    const double total_energy = reduce4(per_cell, m->n_cells);
    free(per_cell);
    for (int si = 0; si < m->n_sensors; ++si) { m->sensors[si].t_steady = saved[si]; }
    free(saved);
    return total_energy / (double)m->num_phonons;
}

int or_interpret(or_model* m, const int64_t* counts, const double* flux, double* out, double* temps_out,
    double* fluxes_out) {
    const int64_t R = m->n_record;
    double* temps = (double*)malloc(sizeof(double) * (size_t)R);
    double* fx = (double*)malloc(sizeof(double) * (size_t)R);
    double* fy = (double*)malloc(sizeof(double) * (size_t)R);
    double* final_temp = (double*)malloc(sizeof(double) * (size_t)m->n_sensors);
    /* Model::resetRequired (model.cpp:250-272) with getFinalTemp (sensorInterpreter.cpp:68-74),
     * evaluated with the budget's energy per phonon. */
    for (int si = 0; si < m->n_sensors; ++si) {
        const sensor_t* s = &m->sensors[si];
        final_temp[si] = s->t_init;
        if (m->sim_type != SIM_TRANSIENT) {
            double ft = 0.;
            if (s->area != 0.) {
                find_temperature(m, m->eff_energy, si, counts + (int64_t)si * R + m->start_step, R - m->start_step,
                    temps);
                ft = accumulate(temps, (int)(R - m->start_step)) / (double)(R - m->start_step);
            }
            final_temp[si] = ft;
        }
    }
    const double eff = post_run_eff_energy(m, final_temp);
    for (int si = 0; si < m->n_sensors; ++si) {
        const sensor_t* s = &m->sensors[si];
        const int64_t* en = counts + (int64_t)si * R;
        /* SensorInterpreter::scaleHeatParams, sensorInterpreter.cpp:19-66: the first entry is
         * replaced by getInitTemp() (t_steady_ in steady state, t_init otherwise) */
        find_temperature(m, eff, si, en, R, temps);
        temps[0] = (m->sim_type == SIM_SS) ? final_temp[si] : s->t_init;
        const double f_factor = eff / s->area;
        for (int64_t k = 0; k < R; ++k) {
            fx[k] = flux[2 * ((int64_t)si * R + k)] * f_factor;
            fy[k] = flux[2 * ((int64_t)si * R + k) + 1] * f_factor;
        }
        double* o = out + 6 * si;
        avg_and_stderr(temps, R, &o[0], &o[1]);
        avg_and_stderr(fx, R, &o[2], &o[3]);
        avg_and_stderr(fy, R, &o[4], &o[5]);
        if (temps_out) { memcpy(temps_out + (int64_t)si * R, temps, sizeof(double) * (size_t)R); }
        if (fluxes_out) {
            for (int64_t k = 0; k < R; ++k) {
                fluxes_out[2 * ((int64_t)si * R + k)] = fx[k];
                fluxes_out[2 * ((int64_t)si * R + k) + 1] = fy[k];
            }
        }
    }
    free(temps);
    free(fx);
    free(fy);
    free(final_temp);
    return 0;
}

double or_result_eff_energy(or_model* m, const int64_t* counts) {
    const int64_t R = m->n_record;
    double* temps = (double*)malloc(sizeof(double) * (size_t)R);
    double* final_temp = (double*)malloc(sizeof(double) * (size_t)m->n_sensors);
    for (int si = 0; si < m->n_sensors; ++si) {
        const sensor_t* s = &m->sensors[si];
        final_temp[si] = s->t_init;
        if (m->sim_type != SIM_TRANSIENT) {
            double ft = 0.;
            if (s->area != 0.) {
                find_temperature(m, m->eff_energy, si, counts + (int64_t)si * R + m->start_step, R - m->start_step,
                    temps);
                ft = accumulate(temps, (int)(R - m->start_step)) / (double)(R - m->start_step);
            }
            final_temp[si] = ft;
        }
    }
    const double eff = post_run_eff_energy(m, final_temp);
    free(temps);
    free(final_temp);
    return eff;
}

/* ------------------------------------------------------------------------ */
/* Introspection for the tests */

int or_table(or_model* m, int material, int kind, double temp, double* cdf, double* frac, double* sum,
    int64_t* index) {
    if (material < 0 || material >= m->n_mat || !m->mats[material].temps) { return set_err(m, "no tables"); }
    material_t* mt = &m->mats[material];
    const int64_t idx = mat_temp_index(mt, temp);
    const table_t* t = mat_table(mt, kind, idx);
    memcpy(cdf, t->cdf, sizeof t->cdf);
    memcpy(frac, t->frac, sizeof t->frac);
    *sum = t->sum;
    *index = idx;
    return 0;
}

int or_material_arrays(or_model* m, int material, double* freq, double* vel_la, double* vel_ta, double* dens_la,
    double* dens_ta) {
    if (material < 0 || material >= m->n_mat) { return set_err(m, "no such material"); }
    const material_t* mt = &m->mats[material];
    memcpy(freq, mt->freq, sizeof mt->freq);
    memcpy(vel_la, mt->vel_la, sizeof mt->vel_la);
    memcpy(vel_ta, mt->vel_ta, sizeof mt->vel_ta);
    memcpy(dens_la, mt->dens_la, sizeof mt->dens_la);
    memcpy(dens_ta, mt->dens_ta, sizeof mt->dens_ta);
    return 0;
}

int64_t or_num_temps(const or_model* m, int material) { return m->mats[material].n_temps; }

int or_temps(const or_model* m, int material, double* temps, double* base_sums) {
    const material_t* mt = &m->mats[material];
    memcpy(temps, mt->temps, sizeof(double) * (size_t)mt->n_temps);
    memcpy(base_sums, mt->base_sums, sizeof(double) * (size_t)mt->n_temps);
    return 0;
}

int or_cell_geometry(const or_model* m, int ci, double* area, double* lines, double* normals, double* spec,
    int32_t* n_trans, int32_t* n_emit) {
    const cell_t* c = &m->cells[ci];
    *area = tri_area(&c->tri);
    for (int b = 0; b < 3; ++b) {
        const line_t* l = &c->b[b].line;
        double* o = lines + 7 * b;
        o[0] = l->p1.x;
        o[1] = l->p1.y;
        o[2] = l->p2.x;
        o[3] = l->p2.y;
        o[4] = l->slope;
        o[5] = l->intercept;
        o[6] = l->length;
        normals[2 * b] = c->b[b].normal.x;
        normals[2 * b + 1] = c->b[b].normal.y;
        spec[b] = c->b[b].spec;
        n_trans[b] = c->b[b].n_trans;
        n_emit[b] = c->b[b].n_emit;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Hooks for the oracle's own pinning tests */

int or_math(int fn, const double* x, double* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (fn) {
        case 0: y[i] = psim_log(x[i]); break;
        case 1: y[i] = psim_exp(x[i]); break;
        case 2: y[i] = psim_sinh(x[i]); break;
        case 3: y[i] = psim_cos2pi(x[i]); break;
        case 4: y[i] = psim_pow4(x[i]); break;
        default: return -1;
        }
    }
    return 0;
}

void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    psim_u32x4 c;
    for (int i = 0; i < 4; ++i) { c.v[i] = ctr[i]; }
    const psim_u32x4 r = psim_philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) { out[i] = r.v[i]; }
}

void or_stream(uint64_t seed, uint32_t tag, uint32_t id, uint32_t step, int64_t n, int pairs, double* out) {
    psim_stream st;
    psim_stream_init(&st, seed, tag, id, step);
    if (pairs) {
        for (int64_t i = 0; i + 1 < n; i += 2) { psim_urand2(&st, &out[i], &out[i + 1]); }
    } else {
        for (int64_t i = 0; i < n; ++i) { out[i] = psim_urand(&st); }
    }
}

/* Material::freqIndex with the draws u1 (bisection) and u2 (polarization). */
void or_freq_index(const double* cdf, const double* frac, double u1, double u2, int64_t* bin, int32_t* pol) {
    table_t t;
    memcpy(t.cdf, cdf, sizeof t.cdf);
    memcpy(t.frac, frac, sizeof t.frac);
    int64_t index = 0;
    int polar = 0;
    freq_index_draws(&t, u1, u2, &index, &polar);
    *bin = index;
    *pol = polar;
}

/* Directions: kind 0 Phonon::setRandDirection, kind 1 Surface::redirectPhonon about (nx, ny),
 * kind 2 the specular branch of Surface::boundaryHandlePhonon applied to (dx, dy) = out[2i..]. */
void or_directions(int kind, uint64_t seed, int64_t n, double nx, double ny, double* out) {
    const pt nrm = { nx, ny };
    for (int64_t i = 0; i < n; ++i) {
        phonon_t p;
        memset(&p, 0, sizeof p);
        psim_stream st;
        psim_stream_init(&st, seed, PSIM_TAG_TRANSPORT, (uint32_t)i, 7);
        if (kind == 0) {
            set_rand_direction(&p, &st);
        } else if (kind == 1) {
            redirect(nrm, &p, &st);
        } else {
            p.dx = out[2 * i];
            p.dy = out[2 * i + 1];
            boundary_handle(nrm, 1.0, &p, &st);
        }
        out[2 * i] = p.dx;
        out[2 * i + 1] = p.dy;
    }
}
