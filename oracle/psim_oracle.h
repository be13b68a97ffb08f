/*
 * CPU oracle for psim's phonon path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference (GwGibson/Psim, psim/src/ *.cpp):
 * model construction and validation, material tables, emission budget,
 * phonon emission, the per-phonon event loop (simulatePhonon /
 * handleImpacts / nextImpact / scatter), sensor tallies and the
 * SensorInterpreter.  Each function cites the reference lines it follows.
 *
 * Only the tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the CPU baseline -- never as
 * the product path.
 *
 * Randomness follows the documented Philox schedule in psim_rng.h instead of
 * the reference's unseeded mt19937 (psim/include/psim/utils.h:16-21), and the
 * transport loop uses the shared transcendentals in psim_math.h; with those
 * two substitutions every phonon trajectory is reproducible and the HIP
 * kernels must match it bit for bit.
 */
#ifndef PSIM_ORACLE_H
#define PSIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_model or_model;

/* settings[8] = num_runs, num_measurements, num_phonons, sim_time, t_eq,
 * sim_type (0 SS, 1 periodic, 2 transient), step_interval, phasor_sim. */
or_model* or_model_new(const double* settings, char* err, int errlen);
void or_model_free(or_model* m);
const char* or_error(const or_model* m);

/* params[13] = la[3], ta[3], max_freq_la, max_freq_ta, b_l, b_tn, b_tu, b_i, w */
int or_add_material(or_model* m, const char* name, const double* params);
int or_add_sensor(or_model* m, int64_t id, const char* material, double t_init);
/* tri[6] = x1 y1 x2 y2 x3 y3 */
int or_add_cell(or_model* m, const double* tri, int64_t sensor_id, double spec);
/* e[7] = x1 y1 x2 y2 temp duration start_time */
int or_add_emit(or_model* m, const double* e);

/* Bounds, material tables, effective energy and the seeded emission budget. */
int or_prepare(or_model* m, uint64_t seed);
/* on != 0: run the reference's transport loop as written, one drift per measurement step,
 * instead of flights across the boundaries (a check of the flights; never used otherwise). */
int or_set_step_by_step(or_model* m, int on);
/* Random-stream index (0..65535) of every draw: stream m reproduces model m of a batched run,
 * whose phonon ids count from 0 within the model.  Clears the prepared budget. */
int or_set_stream(or_model* m, int stream);
/* Test hook: plants a known mistake in the tallies of the boundaries a flight passes (1: one
 * index late, 2: into the next sensor; 0: none), so that a test can show its check detects it. */
int or_plant_flight_fault(or_model* m, int kind);

int or_num_sensors(const or_model* m);
int64_t or_num_record(const or_model* m);
int64_t or_total_phonons(const or_model* m);
double or_eff_energy(const or_model* m);
int or_num_sources(const or_model* m);
/* per source: cell, kind (0 cell-origin, 1 emit), energy, count */
int or_sources(const or_model* m, int64_t* cell, int32_t* kind, double* energy, int64_t* count);

/* Counter slots written by or_run. */
enum {
    OR_CNT_PHONON_STEPS = 0,
    OR_CNT_SCATTERS,
    OR_CNT_IMPACTS,
    OR_CNT_RELOCATIONS,
    OR_CNT_ABSORBED,
    OR_CNT_EXPIRED,
    OR_CNT_NOT_STARTED,
    OR_CNT_TALLIES,
    OR_CNT_STALLED,
    OR_NUM_COUNTERS
};

/* Simulates every phonon of the prepared budget (or_prepare must have been
 * called with the same seed).  counts[S*R] (int64) and flux[S*R*2] receive
 * the tallies, sensor-major in sensor insertion order; counters[OR_NUM_COUNTERS].
 * fate (optional, may be NULL): per phonon id, int64 step at death, int32
 * cause (1 absorbed, 2 expired, 3 not started, 4 stalled) and final x, y. */
int or_run(or_model* m, uint64_t seed, int nthreads, int64_t* counts, double* flux, int64_t* counters,
    int64_t* fate_step, int32_t* fate_cause, double* fate_xy);

/* SensorInterpreter::scaleHeatParams for every sensor.  out[S*6] = t_steady,
 * std_t, x_flux, std_x, y_flux, std_y; temps[S*R] and fluxes[S*R*2] receive
 * the per-step series (either may be NULL). */
int or_interpret(or_model* m, const int64_t* counts, const double* flux, double* out, double* temps, double* fluxes);
/* The energy per phonon the interpreter uses after the run (differs from the budget's in
 * deviational steady state, see psim_oracle.c). */
double or_result_eff_energy(or_model* m, const int64_t* counts);

/* Material table of kind 0 base, 1 emit, 2 scatter at temperature T (grid
 * lookup as Material::getTempIndex); cdf[1000], frac[1000]. */
int or_table(or_model* m, int material, int kind, double temp, double* cdf, double* frac, double* sum, int64_t* index);
int or_material_arrays(or_model* m, int material, double* freq, double* vel_la, double* vel_ta, double* dens_la,
    double* dens_ta);
int64_t or_num_temps(const or_model* m, int material);
int or_temps(const or_model* m, int material, double* temps, double* base_sums);

/* Geometry: per cell area, then per edge (3 per cell) line[7] (p1 p2 slope
 * intercept length), normal[2], spec, n_trans, n_emit. */
int or_cell_geometry(const or_model* m, int cell, double* area, double* lines, double* normals, double* spec,
    int32_t* n_trans, int32_t* n_emit);

/* Test hooks: the shared transcendentals (fn 0 log, 1 exp, 2 sinh, 3 cos2pi, 4 pow4), one
 * Philox4x32-10 block, a uniform stream, the frequency-bin bisection with given draws, and
 * direction sampling (0 isotropic, 1 Lambertian about n, 2 specular reflection of out[]). */
int or_math(int fn, const double* x, double* y, int64_t n);
void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
/* n draws of a stream: single draws (psim_urand), or n / 2 pair draws (psim_urand2) */
void or_stream(uint64_t seed, uint32_t tag, uint32_t id, uint32_t step, int64_t n, int pairs, double* out);
void or_freq_index(const double* cdf, const double* frac, double u1, double u2, int64_t* bin, int32_t* pol);
void or_directions(int kind, uint64_t seed, int64_t n, double nx, double ny, double* out);

#ifdef __cplusplus
}
#endif

#endif
