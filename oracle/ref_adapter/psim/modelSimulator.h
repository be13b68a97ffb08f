// Test infrastructure: the MI355X engine as a drop-in for the reference's ModelSimulator.
//
// The reference's Model holds `ModelSimulator simulator_` (psim/include/psim/model.h:124) and calls
// only its constructor, setStepAdjustment, initPhononBuilders, runSimulation and reset
// (psim/src/model.cpp:39, 62, 160-161, 275).  This header declares a class of that name with
// exactly that interface, implemented over the engine's C ABI (include/psim_gpu.h) in
// oracle/ref_adapter/gpu_model_simulator.cpp.  `make -C oracle ref-gpu` compiles the reference's own,
// unmodified sources without modelSimulator.cpp, each translation unit force-including this header
// first: it carries the include guard of the reference's modelSimulator.h, so model.h's include of
// that header is skipped and Model::simulator_ is this class.  With the engine library linked, the
// result (psim_ref_gpu) is the reference's Model, sensors and SensorInterpreter around the engine's
// phonon loop.
#ifndef PSIM_MODELSIMULATOR_H
#define PSIM_MODELSIMULATOR_H

#include <cstddef>
#include <cstdint>
#include <vector>

#include "psim_gpu.h"

class Cell;
class Sensor;

class ModelSimulator {
public:
    ModelSimulator(std::size_t measurement_steps, double simulation_time, bool phasor_sim);
    ~ModelSimulator();
    ModelSimulator(const ModelSimulator&) = delete;
    ModelSimulator& operator=(const ModelSimulator&) = delete;

    // ModelSimulator::runSimulation (modelSimulator.cpp:39-41): the engine's run of the uploaded
    // model, its tallies written into the sensors (Sensor::inc_energy_ / inc_flux_).
    void runSimulation(double t_eq);
    // ModelSimulator::initPhononBuilders (modelSimulator.cpp:43-85): the budget and the upload.
    void initPhononBuilders(std::vector<Cell>& cells, double t_eq, double eff_energy) noexcept;
    void reset() noexcept { total_phonons_ = 0; }
    void setStepAdjustment(std::size_t step_adjustment) { step_adjustment_ = step_adjustment; }

    std::size_t total_phonons_{ 0 };  // the budget of the last upload (read by the reference driver)

private:
    psim_gpu_ctx* ctx_{ nullptr };
    std::size_t measurement_steps_;
    double simulation_time_;
    double step_time_;
    bool phasor_sim_;
    std::size_t step_adjustment_{ 0 };
    uint64_t seed_base_{ 0 };  // PSIM_REF_GPU_SEED; run r draws from seed_base_ + r * 0x9E3779B97F4A7C15
    uint64_t run_{ 0 };
    uint64_t seed_{ 0 };
    int error_{ 0 };
    std::vector<Sensor*> sensors_;  // in the Model's order (its sensors_ vector is reserved)
    std::size_t num_record_{ 0 };
};

#endif  // PSIM_MODELSIMULATOR_H
