// Device-side phonon transport for gfx950.
//
// One lane owns one phonon.  The event loop is psim's ModelSimulator::simulatePhonon
// (psim/src/modelSimulator.cpp:139-198) with handleImpacts (:205-225), nextImpact (:87-122),
// scatter (:124-137), the surface handlers (psim/src/surface.cpp:23-109,
// compositeSurface.cpp:47-66, cell.cpp:103-108) and the geometry predicates of
// psim/src/geometry.cpp.  Every floating-point expression keeps the reference's operation
// order (the library is built with -ffp-contract=off), so a lane reproduces the CPU oracle's
// trajectory bit for bit given the same Philox stream.
//
// Phonon state lives in registers for a whole launch window of measurement steps and is
// written back to the structure-of-arrays pool only when the window ends.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "psim_gpu.h"
#include "psim_math.h"
#include "psim_rng.h"

namespace psim_dev {

constexpr double GEOEPS = __DBL_EPSILON__ * 1E9;  // utils.h:10
constexpr double SCALING_FACTOR = 1e9;  // modelSimulator.cpp:17
constexpr double VELOCITY_EPS = 0.01;  // modelSimulator.cpp:23
constexpr int MAX_COLLISIONS = 100;  // modelSimulator.cpp:25
constexpr double HBAR = 1.054517e-34;  // material.cpp:12
constexpr double BOLTZ = 1.38065e-23;  // material.cpp:13
constexpr int NBINS = PSIM_NUM_FREQ_BINS;
constexpr int GUIDE = 2048;  // guide-table cells per frequency table (a power of two)

enum : int { POL_LA = 0, POL_TA = 1 };
enum : int { FATE_ABSORBED = 1, FATE_EXPIRED = 2, FATE_NOT_STARTED = 3, FATE_STALLED = 4 };

// Phonon record word: bits 0-9 frequency bin, bit 10 TA, bit 11 negative sign, bit 12 started,
// bits 16-31 random-stream index (the model's index in a batch, 0 otherwise).
constexpr uint32_t META_BIN_MASK = 0x3ffu;
constexpr uint32_t META_TA = 1u << 10;
constexpr uint32_t META_NEG = 1u << 11;
constexpr uint32_t META_STARTED = 1u << 12;
constexpr int META_STREAM_SHIFT = 16;

struct Model {
    const psim_cell* __restrict__ cells;
    const psim_edge* __restrict__ edges;
    const psim_subsurface* __restrict__ subs;
    const psim_sensor* __restrict__ sensors;
    const psim_material* __restrict__ mats;
    const double* __restrict__ freq;
    const double* __restrict__ vel_la;
    const double* __restrict__ vel_ta;
    const double* __restrict__ cdf;
    const double* __restrict__ frac;
    // [n_tables][GUIDE + 1] start indices of the frequency search, or null when some table is not
    // monotone (then the bisection runs as written).
    const uint16_t* __restrict__ guide;
    const psim_source* __restrict__ sources;
    int32_t n_sources;
    int32_t n_sensors;
    int32_t full;
    int32_t phasor;
    int32_t N;  // measurement steps (validated to fit int32 at upload)
    int32_t adj;
    int32_t R;
    // Steady state: phonons whose step lies below `stop` (= adj - 1, else 0) fly across the
    // measurement boundaries below it without stopping; those boundaries only count steps.
    int32_t stop;
    double sim_time;
    double step_time;
    double t_eq;
};

// Birth time, cell (-1: never starts) and the cell's sensor of a pooled phonon.
struct __attribute__((aligned(16))) PoolAux {
    double age;
    int32_t cell;
    int32_t sensor;
};

// The emission pool, indexed by phonon id: five arrays of 8- and 16-byte records, so that a lane
// moves a phonon in five vector loads and the kernels keep five pointers instead of twelve.  A
// newborn phonon has no scattering time yet (0).
struct Pool {
    double2* __restrict__ pos;  // px, py
    double2* __restrict__ dir;  // dx, dy
    double2* __restrict__ fv;  // freq, vel
    PoolAux* __restrict__ aux;
    uint2* __restrict__ ms;  // meta, id within the phonon's random stream (id - source id_base)
};

struct Ph {
    double px, py, dx, dy, freq, vel;
    int32_t cell, bin, pol, sign;
    int32_t step;
};

struct Line {
    double p1x, p1y, p2x, p2y, slope, intercept, blx, bly, trx, try_;
};

// ---- geometry (psim/src/geometry.cpp) ------------------------------------------------

__device__ __forceinline__ bool pt_eq(double ax, double ay, double bx, double by) {  // :34-39
    const double dx = ax - bx;
    const double dy = ay - by;
    return dx * dx + dy * dy < GEOEPS * GEOEPS;
}

__device__ __forceinline__ double get_slope(double x1, double y1, double x2, double y2) {  // :325-327
    return (fabs(x1 - x2) < GEOEPS) ? 0. : (y1 - y2) / (x1 - x2);
}

__device__ __forceinline__ Line make_line(double x1, double y1, double x2, double y2) {  // :45-53, :267-280
    Line l;
    l.p1x = x1;
    l.p1y = y1;
    l.p2x = x2;
    l.p2y = y2;
    l.slope = get_slope(x1, y1, x2, y2);
    l.intercept = y1 - l.slope * x1;
    l.blx = (x2 <= x1) ? x2 : x1;
    l.trx = (x2 <= x1) ? x1 : x2;
    l.bly = (y2 <= y1) ? y2 : y1;
    l.try_ = (y2 <= y1) ? y1 : y2;
    return l;
}

// An edge's line from its endpoints, slope and intercept.  The bounding box is the min/max of the
// endpoints, as the host builds it (the same values as the record's box), so it is not loaded.
__device__ __forceinline__ Line edge_line(const psim_edge& e) {
    Line l;
    l.p1x = e.p1x;
    l.p1y = e.p1y;
    l.p2x = e.p2x;
    l.p2y = e.p2y;
    l.slope = e.slope;
    l.intercept = e.intercept;
    l.blx = (l.p2x <= l.p1x) ? l.p2x : l.p1x;
    l.trx = (l.p2x <= l.p1x) ? l.p1x : l.p2x;
    l.bly = (l.p2y <= l.p1y) ? l.p2y : l.p1y;
    l.try_ = (l.p2y <= l.p1y) ? l.p1y : l.p2y;
    return l;
}

// crossProduct of (p2 - p1) and (p - p1), :291-300, :311-313
__device__ __forceinline__ double line_cross(double x1, double y1, double x2, double y2, double px, double py) {
    const double ax = x2 - x1, ay = y2 - y1;
    const double bx = px - x1, by = py - y1;
    return ax * by - bx * ay;
}

__device__ __forceinline__ bool on_line(double x1, double y1, double x2, double y2, double px, double py) {
    return fabs(line_cross(x1, y1, x2, y2, px, py)) < GEOEPS;
}

// doesSegmentCrossLine(l1, l2), :302-306
__device__ __forceinline__ bool segment_crosses_line(const Line& l1, const Line& l2) {
    const double c1 = line_cross(l1.p1x, l1.p1y, l1.p2x, l1.p2y, l2.p1x, l2.p1y);
    const double c2 = line_cross(l1.p1x, l1.p1y, l1.p2x, l1.p2y, l2.p2x, l2.p2y);
    return fabs(c1) < GEOEPS || fabs(c2) < GEOEPS || ((c1 < 0.) != (c2 < 0.));
}

__device__ __forceinline__ bool rel_equals(double a, double b) {  // :108-110
    return fabs(a - b) <= ((fabs(a) < fabs(b) ? fabs(b) : fabs(a)) * GEOEPS);
}

// Line::getIntersection(self = edge, other = path), :104-138 with Line::intersects :92-95
// and boxesIntersect :282-289.
__device__ __forceinline__ bool get_intersection(const Line& s, const Line& o, double& px, double& py) {
    if (!(s.blx <= o.trx - GEOEPS && s.trx >= o.blx + GEOEPS && s.bly <= o.try_ - GEOEPS
            && s.try_ >= o.bly + GEOEPS)) {
        return false;
    }
    if (!segment_crosses_line(s, o) || !segment_crosses_line(o, s)) { return false; }
    if (rel_equals(s.p1x, s.p2x)) {  // self vertical: handle_vert(self.p1, other.p1, other.p2)
        if (rel_equals(o.p1x, o.p2x)) { return false; }
        const double m = get_slope(o.p1x, o.p1y, o.p2x, o.p2y);
        const double b = o.p1y - m * o.p1x;
        px = s.p1x;
        py = m * s.p1x + b;
        return true;
    }
    if (rel_equals(o.p1x, o.p2x)) {  // other vertical: handle_vert(other.p1, self.p1, self.p2)
        const double m = get_slope(s.p1x, s.p1y, s.p2x, s.p2y);
        const double b = s.p1y - m * s.p1x;
        px = o.p1x;
        py = m * o.p1x + b;
        return true;
    }
    if (rel_equals(s.slope, o.slope)) { return false; }
    const double x = (o.intercept - s.intercept) / (s.slope - o.slope);
    px = x;
    py = s.slope * x + s.intercept;
    return true;
}

// ---- phonon updates -------------------------------------------------------------------

// Material::freqIndex (material.cpp:64-75): a bisection over the cumulative table that keeps
// cdf[low] <= rand < cdf[high] and returns `high`.  For a non-decreasing table that is the first
// index whose value exceeds rand, clamped to [1, NBINS - 1] (bin 0 folds into bin 1, as in the
// reference).  With a guide table (guide[g] = first index whose value exceeds g / GUIDE, for
// g = 0..GUIDE) that index lies in [guide[g], guide[g + 1]] for g = floor(rand * GUIDE), which a
// short scan finds: one dependent load in place of ten.  Tables that are not non-decreasing
// have no guide and run the bisection.
__device__ __forceinline__ void freq_index(const double* __restrict__ cdf, const double* __restrict__ frac,
    const uint16_t* __restrict__ guide, psim_stream& st, int& bin, int& pol) {
    double rand, upol;  // the bisection's draw and the branch draw, one block
    psim_urand2(&st, &rand, &upol);
    int high;
    if (guide) {
        const int g = (int)(rand * (double)GUIDE);
        int i = guide[g];
        const int end = guide[g + 1];
        while (i < end && !(rand < cdf[i])) { ++i; }
        high = i < 1 ? 1 : (i > NBINS - 1 ? NBINS - 1 : i);
    } else {
        int low = 0;
        high = NBINS - 1;
        int mid = low + (high - low) / 2;
        while (high - low > 1) {
            if (rand < cdf[mid]) {
                high = mid;
            } else {
                low = mid;
            }
            mid = (low + high) / 2;
        }
    }
    bin = high;
    pol = (upol <= frac[high]) ? POL_LA : POL_TA;
}

__device__ __forceinline__ double bin_velocity(const Model& m, int mat, int bin, int pol) {  // material.cpp:82-84
    return (pol == POL_LA) ? m.vel_la[mat * NBINS + bin] : m.vel_ta[mat * NBINS + bin];
}

// SensorController::initialUpdate / scatterUpdate with Material::getFreq, sensorController.cpp:28-55,
// material.cpp:77-80
__device__ __forceinline__ void resample(const Model& m, int mat, int table, Ph& p, psim_stream& st) {
    int bin, pol;
    freq_index(m.cdf + (int64_t)table * NBINS, m.frac + (int64_t)table * NBINS,
        m.guide ? m.guide + (int64_t)table * (GUIDE + 1) : nullptr, st, bin, pol);
    const psim_material& mt = m.mats[mat];
    const double f0 = m.freq[mat * NBINS + bin];
    p.freq = m.full ? f0 : f0 + (2. * psim_urand(&st) - 1.) * mt.freq_width / 2.;
    p.bin = bin;
    p.pol = pol;
    p.vel = bin_velocity(m, mat, bin, pol);
}

__device__ __forceinline__ void set_rand_direction(Ph& p, psim_stream& st) {  // phonon.cpp:28-31
    double u1, u2;
    psim_urand2(&st, &u1, &u2);
    p.dx = 2. * u1 - 1.;
    p.dy = sqrt(1. - p.dx * p.dx) * psim_cos2pi(u2);
}

__device__ __forceinline__ void redirect(double nx, double ny, Ph& p, psim_stream& st) {  // surface.cpp:23-30
    double rand, u2;
    psim_urand2(&st, &rand, &u2);
    const double new_dx = sqrt(rand);
    const double new_dy = sqrt(1. - rand) * psim_cos2pi(u2);
    p.dx = nx * new_dx - ny * new_dy;
    p.dy = ny * new_dx + nx * new_dy;
}

__device__ __forceinline__ void boundary_handle(double nx, double ny, double spec, Ph& p,
    psim_stream& st) {  // surface.cpp:32-44
    if (spec == 1. || psim_urand(&st) < spec) {
        const double dx = p.dx, dy = p.dy;
        const double new_dx = -dx * nx - dy * ny;
        const double new_dy = -dx * ny + dy * nx;
        p.dx = nx * new_dx - ny * new_dy;
        p.dy = ny * new_dx + nx * new_dy;
    } else {
        redirect(nx, ny, p, st);
    }
}

// Cell::handleSurfaceCollision -> CompositeSurface::handlePhonon and the surface handlers.
__device__ __forceinline__ void handle_surface_collision(const Model& m, Ph& p, double poix, double poiy, psim_stream& st) {
    const psim_edge* e3 = m.edges + 3 * p.cell;
    for (int b = 0; b < 3; ++b) {
        const psim_edge& e = e3[b];
        if (!on_line(e.p1x, e.p1y, e.p2x, e.p2y, poix, poiy)) { continue; }
        const int n_sub = e.n_trans + e.n_emit;
        for (int i = 0; i < n_sub; ++i) {
            const psim_subsurface& s = m.subs[e.sub_begin + i];
            if (!on_line(s.p1x, s.p1y, s.p2x, s.p2y, poix, poiy)) { continue; }
            if (s.kind == 0) {  // TransitionSurface::handlePhonon, surface.cpp:71-109
                if (s.same_material) {
                    p.cell = s.target;
                } else {
                    const double max_freq = (p.pol == POL_LA) ? s.max_freq_la : s.max_freq_ta;
                    if (p.freq > max_freq) {
                        redirect(s.nx, s.ny, p, st);
                    } else {
                        p.cell = s.target;
                    }
                }
            } else {  // EmitSurface::handlePhonon, surface.cpp:61-65
                const double phonon_time = (double)p.step * m.step_time;
                if (phonon_time < s.start || phonon_time + m.step_time > s.start + s.duration) {
                    boundary_handle(s.nx, s.ny, s.spec, p, st);
                } else {
                    p.cell = -1;
                }
            }
            return;
        }
        boundary_handle(e.nx, e.ny, e.spec, p, st);
        return;
    }
}

struct LaneCounters {
    uint32_t steps, scatters, impacts, relocations, absorbed, expired, tallies, stalled;
};

// The geometric part of ModelSimulator::nextImpact (modelSimulator.cpp:87-117): the nearest edge
// crossing of the path p -> p + v * time.  On a hit the phonon is moved to the point of impact
// and the impact time is returned in out_time; the collision itself is handled separately
// (modelSimulator.cpp:118).
__device__ bool find_impact(const Model& m, Ph& p, double time, double& out_time, LaneCounters& c) {
    const double vx = p.dx * p.vel;
    const double vy = p.dy * p.vel;
    const double sx = p.px, sy = p.py;
    const double ex = p.px + time * vx, ey = p.py + time * vy;
    if (pt_eq(sx, sy, ex, ey)) { return false; }
    const Line path = make_line(sx, sy, ex, ey);
    const psim_edge* e3 = m.edges + 3 * p.cell;
    bool hit = false;
    double ix = 0., iy = 0.;
    for (int b = 0; b < 3; ++b) {
        const Line l = edge_line(e3[b]);
        double qx, qy;
        if (get_intersection(l, path, qx, qy) && !pt_eq(qx, qy, sx, sy)) {
            const double tx = (vx > VELOCITY_EPS || vx < -VELOCITY_EPS) ? (qx - sx) / vx : time;
            const double ty = (vy > VELOCITY_EPS || vy < -VELOCITY_EPS) ? (qy - sy) / vy : time;
            const double it = (tx <= ty) ? tx : ty;
            if (it <= time) {
                time = it;
                ix = qx;
                iy = qy;
                hit = true;
            }
        }
    }
    if (hit) {
        p.px = ix;
        p.py = iy;
        c.impacts++;
        out_time = time;
    }
    return hit;
}

// Material::relaxRates at the sensor temperature, material.cpp:54-56, 207-239
__device__ __forceinline__ void relax_rates(const Model& m, int sensor, const Ph& p, double* r) {
    const psim_sensor& s = m.sensors[sensor];
    const psim_material& mt = m.mats[s.material];
    const double freq = p.freq;
    if (p.pol == POL_LA) {
        r[0] = mt.b_l * freq * freq * s.t3;
        r[1] = mt.b_l * freq * freq * s.t3;
    } else {
        r[0] = (freq < mt.w) ? mt.b_tn * freq * s.t4 : 0.;
        r[1] = (freq >= mt.w) ? mt.b_tu * freq * freq / psim_sinh(HBAR * freq / (s.t_init * BOLTZ)) : 0.;
    }
    r[2] = mt.b_i * psim_pow4(freq);
}

// ModelSimulator::scatter, modelSimulator.cpp:124-137, with its draw `rand` taken by the caller
// (together with the next scattering-time draw, one block).
// The new direction of an Umklapp scatter (after the resample) and of an impurity scatter is one
// code path, so a wave holding both runs it once.
__device__ __forceinline__ void scatter(const Model& m, Ph& p, const double* r, double rand, psim_stream& st) {
    const double tau_inv = 0. + r[0] + r[1] + r[2];
    const bool resampled = rand <= (r[0] + r[1]) / tau_inv;
    if (resampled) {
        const psim_sensor& s = m.sensors[m.cells[p.cell].sensor];
        resample(m, s.material, s.scatter_table, p, st);
    }
    if (resampled ? rand > r[0] / tau_inv : r[2] > 0.) { set_rand_direction(p, st); }
}

}  // namespace psim_dev
