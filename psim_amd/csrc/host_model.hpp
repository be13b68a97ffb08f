// C++17 host side of the MI355X phonon engine: the reference's Model / Material / Sensor /
// Cell setup, the drop-in GpuModelSimulator, the SensorInterpreter and the OutputManager.
//
// Class and member names follow the reference (psim/include/psim/*.h) so a caller of the
// reference finds the same construction order and semantics:
//   Model(ModelParams) -> setSimulationType -> addMaterial -> addSensor -> addCell ->
//   setEmitSurface -> runSimulation -> exportResults      (model.h:37-110)
// with the reference's validation errors (thrown as std::runtime_error carrying the same
// messages).  The particle loop is not run here: GpuModelSimulator replaces ModelSimulator
// (modelSimulator.h:12-41) and hands flattened descriptors to the HIP engine through the C ABI
// of include/psim_gpu.h.
#pragma once

#include <array>
#include <cstdint>
#include <filesystem>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "psim_gpu.h"

namespace psim {

enum class SimulationType { SteadyState, Periodic, Transient };

struct Point {
    double x, y;
};

struct Line {
    Point p1, p2;
    double slope, intercept;
    Point bl, tr;
    double length;
    Line() = default;
    Line(Point a, Point b);  // throws on identical points (Geometry::LineError)
};

struct Triangle {
    Point p1, p2, p3;
    Triangle(Point a, Point b, Point c);  // throws on a degenerate triangle (Geometry::TriangleError)
};

struct DispersionData {
    std::array<double, 3> LA_data{ 0., 0., 0. };
    std::array<double, 3> TA_data{ 0., 0., 0. };
    double w_max_la{ 0. };
    double w_max_ta{ 0. };
};

struct RelaxationData {
    double b_l{ 0. }, b_tn{ 0. }, b_tu{ 0. }, b_i{ 0. }, w{ 0. };
};

struct Table {
    std::array<double, PSIM_NUM_FREQ_BINS> cdf;
    std::array<double, PSIM_NUM_FREQ_BINS> la_frac;
    double sum;
};

class Material {
public:
    static constexpr int NUM_FREQ_BINS = PSIM_NUM_FREQ_BINS;
    Material(std::size_t id, const DispersionData& d, const RelaxationData& r);
    void setFullSimulation() noexcept { full_ = true; }
    bool full() const noexcept { return full_; }
    std::size_t id() const noexcept { return id_; }
    double max_freq_la() const noexcept { return d_.w_max_la; }
    double max_freq_ta() const noexcept { return d_.w_max_ta; }
    // Material::initializeTables (material.cpp:101-123): temperature grid and per-grid-point
    // energy; individual tables are built on demand.
    void initializeTables(double low, double high);
    std::size_t tempIndex(double temp) const noexcept;  // Material::getTempIndex
    const Table& table(int kind, std::size_t index) const;  // 0 base, 1 emit, 2 scatter
    double baseEnergy(double temp) const { return base_sums_[tempIndex(temp)]; }
    double emitEnergy(double temp) const { return table(1, tempIndex(temp)).sum; }
    psim_material descriptor() const;
    const std::array<double, NUM_FREQ_BINS>& frequencies() const { return freq_; }
    const std::array<double, NUM_FREQ_BINS>& velocitiesLA() const { return vel_la_; }
    const std::array<double, NUM_FREQ_BINS>& velocitiesTA() const { return vel_ta_; }
    const std::vector<double>& temps() const { return temps_; }
    const std::vector<double>& baseSums() const { return base_sums_; }

private:
    std::size_t id_;
    DispersionData d_;
    RelaxationData r_;
    double freq_width_;
    bool full_ = false;
    std::array<double, NUM_FREQ_BINS> freq_{}, dens_la_{}, dens_ta_{}, vel_la_{}, vel_ta_{};
    std::vector<double> temps_;
    std::vector<double> base_sums_;
    mutable std::map<std::size_t, std::shared_ptr<Table>> cache_;
    // bounds and mode the tables were last built for: the tables are a function of these and of
    // the material's constants alone, so a run with the same ones keeps them (and the cache)
    bool built_ = false, built_full_ = false;
    double built_low_ = 0., built_high_ = 0.;
    void phononDist(double temp, bool ta, double* out) const;
};

struct TransitionSurface {
    Line line;
    Point normal;
    std::size_t cell;
};

struct EmitSurface {
    Line line;
    Point normal;
    double spec, temp, duration, start;
};

struct CompositeSurface {
    Line line;
    Point normal;
    double spec;
    std::vector<TransitionSurface> transitions;
    std::vector<EmitSurface> emits;
    // CompositeSurface::verifySurfaceLine (compositeSurface.cpp:68-87): true if `l` lies in this
    // surface; throws on an overlap with an existing sub-surface.
    bool verify(const Line& l) const;
};

class Sensor {
public:
    Sensor(std::size_t id, const Material& mat, SimulationType type, std::size_t num_record, double t_init);
    std::size_t getID() const noexcept { return id_; }
    const Material& getMaterial() const noexcept { return *material_; }
    double getInitTemp() const noexcept;  // SteadyStateController returns t_steady_
    double getArea() const noexcept { return area_; }
    double getHeatCapacity() const noexcept { return heat_capacity_; }
    double t_init() const noexcept { return t_init_; }
    void addToArea(double a) noexcept { area_ += a; }
    void updateTables();  // SensorController::updateTables
    std::size_t tableIndex() const noexcept { return table_index_; }
    void setSteadyTemp(double t) noexcept { t_steady_ = t; }
    double steadyTemp() const noexcept { return t_steady_; }
    std::vector<int64_t> inc_energy;
    std::vector<std::array<double, 2>> inc_flux;

private:
    std::size_t id_;
    const Material* material_;
    SimulationType type_;
    double t_init_;
    double t_steady_;
    double area_ = 0.;
    double heat_capacity_ = 0.;
    std::size_t table_index_ = 0;
};

class Cell {
public:
    Cell(const Triangle& t, Sensor& sensor, double spec);
    Sensor& sensor() const noexcept { return *sensor_; }
    const Triangle& triangle() const noexcept { return tri_; }
    double getArea() const noexcept;
    std::array<CompositeSurface, 3>& boundaries() noexcept { return b_; }
    const std::array<CompositeSurface, 3>& boundaries() const noexcept { return b_; }
    int normSign() const noexcept;
    double getInitEnergy(double t_eq) const;
    double getEmitEnergy(double t_eq) const;

private:
    Triangle tri_;
    Sensor* sensor_;
    std::array<CompositeSurface, 3> b_;
};

struct ModelParams {
    std::size_t num_runs{ 1 };
    std::size_t num_cells{ 0 };
    std::size_t num_sensors{ 0 };
    std::size_t measurement_steps{ 0 };
    std::size_t num_phonons{ 0 };
    double simulation_time{ 0. };
    double t_eq{ 0. };
    bool phasor_sim{ false };
};

struct SensorMeasurements {
    std::size_t id{ 0 };
    double t_steady{ 0. }, std_t_steady{ 0. }, x_flux{ 0. }, std_x_flux{ 0. }, y_flux{ 0. }, std_y_flux{ 0. };
    std::vector<double> final_temps;
    std::vector<std::array<double, 2>> final_fluxes;
};

struct RunRecord {
    std::vector<SensorMeasurements> measurements;  // sorted by sensor id
    std::vector<int64_t> counts;  // raw tallies [S][R], sensor insertion order
    std::vector<double> flux;  // [S][R][2]
    std::array<int64_t, PSIM_NUM_COUNTERS> counters{};
    double eff_energy{ 0. };  // budget energy per phonon
    int64_t total_phonons{ 0 };
};

// A model flattened into the engine's descriptors (include/psim_gpu.h).
struct FlatModel {
    std::vector<psim_cell> cells;
    std::vector<psim_edge> edges;
    std::vector<psim_subsurface> subs;
    std::vector<psim_sensor> sensors;
    std::vector<psim_material> mats;
    std::vector<double> freq, vel_la, vel_ta, cdf, frac;
    std::vector<psim_source> sources;
    psim_run_params params{};
    int32_t n_tables = 0;
};

// Drop-in for ModelSimulator (modelSimulator.h:12-41): the same two calls, backed by the engine.
class GpuModelSimulator {
public:
    GpuModelSimulator(std::size_t measurement_steps, double simulation_time, bool phasor_sim);
    void setEngine(psim_gpu_ctx* ctx) noexcept { ctx_ = ctx; }
    void setSeed(uint64_t seed) noexcept { seed_ = seed; }
    void setStepAdjustment(std::size_t adj) noexcept { step_adjustment_ = adj; }
    void setRecordLength(std::size_t r) noexcept { num_record_ = r; }
    // Random-stream index of this model's phonons (0 unless it is model m of a batch).
    void setStream(int32_t stream) noexcept { stream_ = stream; }
    int32_t stream() const noexcept { return stream_; }
    // ModelSimulator::initPhononBuilders (modelSimulator.cpp:43-85): the emission budget per
    // source (Bernoulli rounding drawn from the Philox budget stream), then the flattened model
    // is uploaded to the engine.
    void initPhononBuilders(std::vector<Cell>& cells, std::vector<Sensor>& sensors, double t_eq, double eff_energy);
    // ModelSimulator::runSimulation (modelSimulator.cpp:39-41): the particle loop on the GPU;
    // tallies land in Sensor::inc_energy / inc_flux.
    void runSimulation(double t_eq, std::vector<Sensor>& sensors);
    void reset() noexcept { total_phonons_ = 0; }
    int64_t totalPhonons() const noexcept { return total_phonons_; }
    const std::array<int64_t, PSIM_NUM_COUNTERS>& counters() const noexcept { return counters_; }
    const std::vector<int64_t>& sourceCounts() const noexcept { return source_counts_; }
    const FlatModel& flat() const noexcept { return flat_; }

private:
    psim_gpu_ctx* ctx_ = nullptr;
    FlatModel flat_;
    int32_t stream_ = 0;
    std::vector<int64_t> source_counts_;
    uint64_t seed_ = 0;
    std::size_t measurement_steps_;
    double simulation_time_;
    double step_time_;
    bool phasor_sim_;
    std::size_t step_adjustment_ = 0;
    std::size_t num_record_ = 0;
    int64_t total_phonons_ = 0;
    std::array<int64_t, PSIM_NUM_COUNTERS> counters_{};
};

class Model {
public:
    explicit Model(const ModelParams& params);
    void setSimulationType(SimulationType type, std::size_t step_interval = 0);
    void addMaterial(const std::string& name, const Material& material);
    void addSensor(std::size_t id, const std::string& material_name, double t_init, SimulationType type);
    void addCell(const Triangle& triangle, std::size_t sensor_id, double spec = 1.);
    bool setEmitSurface(const Point& p1, const Point& p2, double temp, double duration, double start_time);
    // Model::runSimulation (model.cpp:141-182) with the engine of `ctx`; run r uses Philox seed
    // seed + r * 0x9E3779B97F4A7C15.
    void runSimulation(psim_gpu_ctx* ctx, uint64_t seed);
    // Runs the models together: one engine upload and one run (psim_batch_run).  Model m draws the
    // random numbers of its single run with stream m; each keeps its own budget and results.
    // `counters` receives the batch totals.
    static void runBatch(const std::vector<Model*>& models, psim_gpu_ctx* ctx, uint64_t seed,
        std::array<int64_t, PSIM_NUM_COUNTERS>& counters);
    void setStream(int32_t stream) noexcept { simulator_.setStream(stream); }
    void exportResults(const std::filesystem::path& filepath, double time) const;
    // Pre-run state of one run (bounds, tables, energy per phonon, emission budget) without an
    // engine; returns the budget's energy per phonon.
    double prepareBudget(uint64_t seed);
    // The same pre-run state uploaded to the engine of `ctx` (for timing runs with the model
    // resident in HBM); returns the budget's energy per phonon.
    double prepareAndUpload(psim_gpu_ctx* ctx, uint64_t seed);
    const GpuModelSimulator& simulator() const noexcept { return simulator_; }

    std::size_t numRuns() const noexcept { return num_runs_; }
    std::size_t numRecord() const noexcept { return num_record_; }
    const std::vector<Sensor>& sensors() const noexcept { return sensors_; }
    const std::vector<RunRecord>& runs() const noexcept { return runs_; }

private:
    SimulationType sim_type_ = SimulationType::SteadyState;
    std::size_t num_runs_, num_cells_, measurement_steps_, num_phonons_;
    double simulation_time_;
    double t_eq_;
    bool phasor_sim_;
    std::size_t start_step_ = 0;
    std::size_t step_interval_ = 0;
    std::size_t num_record_ = 0;
    double lb_ = 0., ub_ = 0.;
    GpuModelSimulator simulator_;
    std::vector<Cell> cells_;
    std::vector<Sensor> sensors_;
    std::vector<std::pair<std::string, std::unique_ptr<Material>>> materials_;
    std::vector<RunRecord> runs_;

    Sensor& getSensor(std::size_t id);
    double prepare(uint64_t seed);
    // The rest of one run once the sensors hold its tallies (model.cpp:161-181).
    void endRun(double eff0, int64_t total_phonons, const std::array<int64_t, PSIM_NUM_COUNTERS>& counters);
    double getTotalInitialEnergy() const;
    std::pair<double, double> setTemperatureBounds();
    std::vector<double> findTemperature(const Sensor& s, double eff_energy, std::size_t start) const;
    SensorMeasurements scaleHeatParams(const Sensor& s, double eff_energy) const;
};

// InputManager::deserialize (inputManager.cpp:14-115) from JSON text.
std::unique_ptr<Model> deserialize(const std::string& json_text);

}  // namespace psim
