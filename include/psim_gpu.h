/*
 * psim_gpu.h -- C ABI of the MI355X phonon engine.
 *
 * The reference runs its phonon loop inside the process as a C++ member call:
 *   Model::runSimulation (psim/src/model.cpp:160-161)
 *     simulator_.initPhononBuilders(cells_, t_eq_, energy_per_phonon);
 *     simulator_.runSimulation(t_eq_);
 * with ModelSimulator declared in psim/include/psim/modelSimulator.h:12-41 and
 * the per-phonon loop in psim/src/modelSimulator.cpp:139-253.  There is no
 * plugin registry or FFI in the reference, so this header is the boundary a
 * maintainer binds instead of those two calls: the host flattens its Model
 * into the plain descriptors below, uploads them once, and gets the sensor
 * tallies the reference keeps in Sensor::inc_energy_ / inc_flux_
 * (psim/include/psim/sensor.h:73-74) back in caller-owned arrays.
 *
 * Two layers:
 *   psim_gpu_*    the engine: flattened geometry, tables and emission sources
 *                 in, tallies and counters out (replaces
 *                 ModelSimulator::initPhononBuilders + runSimulation);
 *   psim_model_*  the C++17 host behind a C ABI: reads the reference's JSON
 *                 input, builds and validates the model with the reference's
 *                 semantics, flattens it, runs the engine, interprets the
 *                 tallies (SensorInterpreter) and writes ss_/per_ tables
 *                 (OutputManager).
 *
 * Conventions: all functions return 0 on success and a negative PSIM_E_* code
 * on failure; no exception crosses the ABI; the message of the last failure
 * is returned by psim_gpu_last_error / psim_model_last_error.  Units are nm,
 * ns, K, m/s.  Arrays are caller-owned; the engine copies what it keeps.
 * A context is used by one host thread at a time.
 */
#ifndef PSIM_GPU_H
#define PSIM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSIM_NUM_FREQ_BINS 1000
/* Event-loop iterations a phonon may spend inside one measurement step before it is removed
 * (counted as stalled).  The reference has no such bound; typical phonons use about two. */
#define PSIM_MAX_EVENTS_PER_STEP 1000000

enum {
    PSIM_OK = 0,
    PSIM_E_INVALID = -1, /* bad argument or descriptor */
    PSIM_E_DEVICE = -2, /* HIP runtime error */
    PSIM_E_MODEL = -3, /* model construction / validation error (reference exceptions) */
    PSIM_E_IO = -4, /* file could not be read or written */
    PSIM_E_STATE = -5 /* call out of order */
};

/* ---- engine descriptors --------------------------------------------------- */

/* One of the three edges of a cell (order p1p2, p2p3, p3p1 as Cell::buildCompositeSurfaces,
 * psim/src/cell.cpp:114-124): the main surface line with the fields of Geometry::Line
 * (geometry.h:33-70), its inward normal and specularity, and the range of its sub-surfaces. */
typedef struct {
    double p1x, p1y, p2x, p2y;
    double slope, intercept;
    double blx, bly, trx, try_; /* bounding box (bottom-left, top-right) */
    double nx, ny;
    double spec;
    int32_t sub_begin; /* first sub-surface in psim_geom.subs */
    int32_t n_trans; /* transition sub-surfaces, searched first */
    int32_t n_emit; /* emitting sub-surfaces, searched next */
    int32_t pad_;
} psim_edge;

/* A sub-surface of an edge (CompositeSurface::transition_sub_surfaces_ / emit_sub_surfaces_,
 * psim/include/psim/compositeSurface.h:53-56). */
typedef struct {
    double p1x, p1y, p2x, p2y; /* its own line (used for the containment test) */
    double nx, ny; /* transition: its own normal; emit: the main surface's */
    double spec; /* emit: main surface specularity */
    double start, duration; /* emit: emitting window */
    double max_freq_la, max_freq_ta; /* transition: target material's maximum frequencies */
    int32_t kind; /* 0 transition, 1 emit */
    int32_t target; /* transition: neighbour cell */
    int32_t same_material; /* transition: neighbour has the owning cell's material */
    int32_t pad_;
} psim_subsurface;

typedef struct {
    double x1, y1, x2, y2, x3, y3; /* triangle (Cell::cell_) */
    int32_t sensor; /* index into psim_geom.sensors */
    int32_t pad_;
} psim_cell;

typedef struct {
    double t_init; /* relaxation-rate temperature (SensorController::t_steady_ during the run) */
    double t3, t4; /* pow(t_init, 3), pow(t_init, 4) as the reference's libm gives them */
    double area;
    int32_t material;
    int32_t base_table; /* index into psim_tables */
    int32_t scatter_table;
    /* Flight cap of the sensor's model (psim_flight_cap of that model's sensor count); 0 lets
     * psim_gpu_upload use psim_flight_cap(n_sensors) of the whole upload (one model). */
    int32_t flight_cap;
} psim_sensor;

/* Phonons fly across the measurement boundaries (they do not stop at each one); from the
 * first recorded boundary on, a flight also ends at every flight_cap-th boundary, so that the
 * engine's LDS tally windows of that length hold whole flights.  The cap of a model depends only on
 * its number of sensors: the number of steps of tallies (20 B per sensor, one spare row) that fit
 * 48 KB, at most 128; models with too many sensors for 16 such steps are tallied in HBM and take
 * 128.  It changes the rounding of positions (one flight is summed instead of several), not the
 * statistics; a CPU build must use the same rule to reproduce the engine's trajectories. */
static inline int32_t psim_flight_cap(int32_t n_sensors) {
    const int32_t s = n_sensors > 0 ? n_sensors : 1;
    const int32_t w = 49152 / (20 * s) - 1;
    return (w < 16 || w > 128) ? 128 : w;
}

typedef struct {
    int32_t n_cells, n_sensors, n_subs, pad_;
    const psim_cell* cells; /* [n_cells] */
    const psim_edge* edges; /* [3 * n_cells] */
    const psim_subsurface* subs; /* [n_subs] */
    const psim_sensor* sensors; /* [n_sensors] */
} psim_geom;

/* Material constants (psim/include/psim/material.h:94-116). */
typedef struct {
    double b_l, b_tn, b_tu, b_i, w;
    double w_max_la, w_max_ta;
    double freq_width;
} psim_material;

typedef struct {
    int32_t n_materials, n_tables;
    const psim_material* materials; /* [n_materials] */
    const double* freq; /* [n_materials][1000] bin centres */
    const double* vel_la; /* [n_materials][1000] */
    const double* vel_ta; /* [n_materials][1000] */
    const double* cdf; /* [n_tables][1000] normalised cumulative distribution */
    const double* la_frac; /* [n_tables][1000] LA share of each bin */
} psim_tables;

/* An emission source: a cell-origin or surface-origin builder
 * (psim/include/psim/phononBuilder.h:37-72); phonon ids [offset, offset + count).
 * A phonon's random streams are keyed by (stream, id - id_base): a single model uses stream 0
 * and id_base 0; a batch of models gives model m stream m and id_base = its first phonon id, so
 * every model draws exactly the numbers of a single run with the same stream index. */
typedef struct {
    int32_t kind; /* 0 cell origin, 1 surface origin */
    int32_t cell;
    int32_t table; /* base (cell) or emit (surface) table */
    int32_t sign; /* +1 or -1 */
    int64_t offset, count;
    double start, duration; /* surface: birth window */
    double p1x, p1y, p2x, p2y; /* surface: emitting line */
    double nx, ny; /* surface: inward normal */
    int32_t stream; /* 0 .. 65535 */
    int32_t pad_;
    int64_t id_base; /* <= offset */
} psim_source;

typedef struct {
    int32_t n_sources, pad_;
    const psim_source* sources;
} psim_sources;

typedef struct {
    int64_t num_measurements; /* N */
    int64_t step_adjustment; /* first recorded step (steady state 0.9 N, else 0) */
    int64_t num_record; /* tally length per sensor */
    int64_t total_phonons;
    double sim_time; /* ns */
    double step_time; /* sim_time / N */
    double t_eq;
    int32_t full; /* 1: full simulation (no frequency jitter) */
    int32_t phasor; /* 1: phasor simulation (no scattering) */
} psim_run_params;

/* Counter slots of psim_gpu_run. */
enum {
    PSIM_CNT_PHONON_STEPS = 0, /* measurement-boundary crossings (modelSimulator.cpp:182-183) */
    PSIM_CNT_SCATTERS,
    PSIM_CNT_IMPACTS,
    PSIM_CNT_RELOCATIONS, /* stuck phonons moved after MAX_COLLISIONS */
    PSIM_CNT_ABSORBED,
    PSIM_CNT_EXPIRED,
    PSIM_CNT_NOT_STARTED, /* born at or after the last measurement */
    PSIM_CNT_TALLIES,
    PSIM_CNT_STALLED, /* removed after PSIM_MAX_EVENTS_PER_STEP loop iterations in one step */
    PSIM_NUM_COUNTERS
};

typedef struct psim_gpu_ctx psim_gpu_ctx;

/* Replaces constructing ModelSimulator (modelSimulator.cpp:29-37): a context on one device. */
int psim_gpu_create(psim_gpu_ctx** ctx, int device);
void psim_gpu_destroy(psim_gpu_ctx* ctx);
const char* psim_gpu_last_error(const psim_gpu_ctx* ctx);

/* Replaces ModelSimulator::initPhononBuilders (modelSimulator.cpp:43-85) together with the
 * Model state it reads: copies geometry, tables, emission sources and run parameters to the
 * device.  May be called again to replace the model. */
int psim_gpu_upload(psim_gpu_ctx* ctx, const psim_geom* geom, const psim_tables* tables,
    const psim_sources* sources, const psim_run_params* params);

/* Replaces ModelSimulator::runSimulation (modelSimulator.cpp:39-41, 227-253): emits and
 * simulates every phonon of the uploaded budget with Philox seed `seed`, synchronously.
 * counts[S*R] (int64) and flux[S*R*2] receive the tallies (sensor-major); counters
 * [PSIM_NUM_COUNTERS] may be NULL. */
int psim_gpu_run(psim_gpu_ctx* ctx, uint64_t seed, int64_t* counts, double* flux, int64_t* counters);

/* Same as psim_gpu_run but leaves the tallies on the device (for timing with inputs resident);
 * psim_gpu_fetch copies the last run's tallies and counters to the host. */
int psim_gpu_run_resident(psim_gpu_ctx* ctx, uint64_t seed);
int psim_gpu_fetch(psim_gpu_ctx* ctx, int64_t* counts, double* flux, int64_t* counters);

/* Per-phonon fate of the last run (for trajectory comparisons): int64 step at death, int32
 * cause (1 absorbed, 2 expired, 3 never started) and final x, y.  Enable before the run. */
int psim_gpu_set_fate(psim_gpu_ctx* ctx, int enable);
int psim_gpu_fetch_fate(psim_gpu_ctx* ctx, int64_t* step, int32_t* cause, double* xy);

/* Timing of the last run measured with HIP events on the engine stream: total milliseconds
 * of the transport (step) kernel launches, their count, and the whole run. */
int psim_gpu_last_timing(const psim_gpu_ctx* ctx, double* step_kernel_ms, int64_t* step_launches,
    double* run_ms);

/* Engine tuning: measurement steps per transport launch (0 = automatic).  No length changes a
 * result.  By default a recording launch (LDS tallies) is as long as the model's flight cap
 * (psim_flight_cap), so it holds whole flights; a window that does not line up with the caps lets
 * flights pass its end, and their boundaries past it are tallied in HBM (same sums, slower). */
int psim_gpu_set_window(psim_gpu_ctx* ctx, int32_t steps_per_launch);

/* Engine tuning: transport blocks per compute unit (default 8). */
int psim_gpu_set_grid(psim_gpu_ctx* ctx, int32_t blocks_per_cu);

/* Engine tuning: threads per transport block, 128, 256 (default) or 512.  No size changes a
 * result.  No counterpart in the reference. */
int psim_gpu_set_block(psim_gpu_ctx* ctx, int32_t threads);

/* The launch shape in use: steps per untallied launch (0 = automatic), transport blocks per CU
 * and threads per transport block (any pointer may be NULL). */
int psim_gpu_get_launch_shape(const psim_gpu_ctx* ctx, int32_t* window, int32_t* blocks_per_cu, int32_t* block);

/* Engine tuning: where recorded steps are tallied.  AUTO chooses by model size; LDS keeps
 * per-block window partials in LDS (used whenever one step of tallies fits); HBM adds into the
 * tallies with FP64 atomics in one launch.  Integer counts are the same either way; flux sums
 * differ only in summation order.  No counterpart in the reference. */
#define PSIM_TALLY_AUTO 0
#define PSIM_TALLY_LDS 1
#define PSIM_TALLY_HBM 2
int psim_gpu_set_tally_mode(psim_gpu_ctx* ctx, int32_t mode);

/* Engine tuning: precision of the flux tallies.  FP64 (default) adds doubles, as the reference's
 * Sensor::updateHeatParams does (psim/src/sensor.cpp:43-52).  FP32_KAHAN keeps the LDS window
 * difference arrays and the HBM ones in FP32 (FP32 atomics), sums the window partials along the
 * steps and across blocks with Kahan compensation in FP32, and returns each tally as
 * sum - compensation in the same double arrays; integer counts are unaffected.  The launch plan
 * and trajectories are those of FP64 (its transport blocks are 256 threads), so a run of each on
 * one seed shows the error of the FP32 tallies against FP64.  No counterpart in the reference. */
#define PSIM_TALLY_FP64 0
#define PSIM_TALLY_FP32_KAHAN 1
int psim_gpu_set_tally_precision(psim_gpu_ctx* ctx, int32_t precision);

/* psim_gpu_run replayed from a HIP graph of the whole run (emission, sorts, every transport
 * window and reduction), captured on the first call and again after anything that changes the
 * launch sequence (upload, window, grid, fate recording).  The seed is data, so one graph serves
 * every seed.  Same results as psim_gpu_run; no counterpart in the reference (its loop is a
 * std::for_each, modelSimulator.cpp:227-253). */
int psim_gpu_run_graph(psim_gpu_ctx* ctx, uint64_t seed, int64_t* counts, double* flux, int64_t* counters);

/* The uploaded run replayed from its graph on the caller's HIP stream (a hipStream_t passed as
 * void*; NULL is the null stream), returning at once.  The results are in the engine's device
 * buffers once that stream reaches this point: synchronise it, then read them with
 * psim_gpu_fetch.  One run in flight per context: the runs share the engine's buffers.  For
 * callers that overlap the engine with their own work on a stream; the reference's loop is
 * synchronous (ModelSimulator::runSimulation, psim/src/modelSimulator.cpp:39-41). */
int psim_gpu_run_async(psim_gpu_ctx* ctx, uint64_t seed, void* stream);

/* Autotuner: runs the uploaded model once for each candidate launch shape (untallied window
 * length 0 / 1024 / 256 and 8 or 3 transport blocks per CU at the current block size, then 128,
 * 256 and 512 threads per block at the best of those; none changes any result), keeps the fastest
 * in the context and reports it (the block size through psim_gpu_get_launch_shape).  No
 * counterpart in the reference. */
int psim_gpu_autotune(psim_gpu_ctx* ctx, uint64_t seed, int32_t* window, int32_t* blocks_per_cu, double* best_ms);

/* ---- host model (C++17 behind the C ABI) ------------------------------------ */

typedef struct psim_model psim_model;

/* InputManager::deserialize (psim/src/inputManager.cpp:14-115) from JSON text in the
 * reference's schema; construction errors carry the reference's messages. */
int psim_model_from_json(psim_model** model, const char* json_text, char* err, int errlen);
void psim_model_destroy(psim_model* model);
const char* psim_model_last_error(const psim_model* model);

/* Sizes: number of sensors, recorded steps per sensor, runs. */
int psim_model_sizes(const psim_model* model, int32_t* n_sensors, int64_t* num_record, int64_t* num_runs);

/* Model::runSimulation (model.cpp:141-182) for every run with the engine of `ctx`; run r uses
 * Philox seed seed + r * 0x9E3779B97F4A7C15. */
int psim_model_run(psim_model* model, psim_gpu_ctx* ctx, uint64_t seed);

/* Results of run `run` (SensorInterpreter::scaleHeatParams, sensorInterpreter.cpp:19-66):
 * out[S*6] = t_steady, std, x_flux, std, y_flux, std (sensor id order); temps[S*R] and
 * fluxes[S*R*2] may be NULL; counts/flux[...] (raw tallies) may be NULL. */
int psim_model_results(const psim_model* model, int64_t run, double* out, double* temps, double* fluxes,
    int64_t* counts, double* raw_flux, int64_t* counters);

/* The pre-run part of Model::runSimulation for one run with Philox seed `seed`, without an
 * engine (model.cpp:144-160 and ModelSimulator::initPhononBuilders): temperature bounds,
 * material tables, the energy per phonon and the emission budget.  counts[counts_cap] receives
 * the phonon count of each source in builder order (per cell: cell origin, then its emitting
 * sub-surfaces); *n_sources their number. */
int psim_model_prepare(psim_model* model, uint64_t seed, double* eff_energy, int64_t* total_phonons, int64_t* counts,
    int32_t counts_cap, int32_t* n_sources);

/* psim_model_prepare followed by psim_gpu_upload of the flattened model to `ctx`, so that
 * psim_gpu_run_resident can then be timed with every input already in HBM. */
int psim_model_upload(psim_model* model, psim_gpu_ctx* ctx, uint64_t seed, double* eff_energy, int64_t* total_phonons);

/* Budget of the last prepared run: energy per phonon and phonon count. */
int psim_model_budget(const psim_model* model, double* eff_energy, int64_t* total_phonons);

/* OutputManager::exportResults (outputManager.cpp:13-38): ss_<stem>.txt or per_<stem>.txt next to
 * `json_path`, averaged over runs.  The periodic writer averages the per-step series over runs
 * (the reference's calculateAverages drops them and then faults, outputManager.cpp:40-93). */
int psim_model_export(const psim_model* model, const char* json_path, double time_s);

/* Random-stream index of the model's phonons (default 0).  Stream m reproduces model m of a
 * batch (psim_batch_run) exactly. */
int psim_model_set_stream(psim_model* model, int32_t stream);

/* Batched geometries (no counterpart in the reference, which runs model files one after the
 * other, psim/src/main.cpp:13-25): runs n independent models as one engine upload and one run.
 * Every model keeps its own tables, budget (its own energy per phonon) and results, readable with
 * psim_model_results(model, 0, ...); model m draws the random numbers of a single run with
 * stream m.  The models must share num_measurements, sim_time, sim_type, step_interval and the
 * phasor flag, and have num_runs 1.  `counters` (may be null) receives the batch totals. */
int psim_batch_run(psim_model* const* models, int32_t n, psim_gpu_ctx* ctx, uint64_t seed, int64_t* counters,
    char* err, int errlen);

#ifdef __cplusplus
}
#endif

#endif /* PSIM_GPU_H */
