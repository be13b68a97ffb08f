// psim_model_* C ABI over the C++17 host (include/psim_gpu.h).  No exception crosses the ABI.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "host_model.hpp"
#include "psim_gpu.h"

struct psim_model {
    std::unique_ptr<psim::Model> model;
    std::string err;
};

namespace {
int model_fail(psim_model* m, int code, const std::string& msg) {
    if (m) { m->err = msg; }
    return code;
}
}  // namespace

extern "C" {

int psim_model_from_json(psim_model** out, const char* json_text, char* err, int errlen) {
    if (!out || !json_text) { return PSIM_E_INVALID; }
    *out = nullptr;
    try {
        auto m = std::make_unique<psim_model>();
        m->model = psim::deserialize(json_text);
        *out = m.release();
        return PSIM_OK;
    } catch (const std::exception& e) {
        if (err && errlen > 0) { std::snprintf(err, (size_t)errlen, "%s", e.what()); }
        return PSIM_E_MODEL;
    }
}

void psim_model_destroy(psim_model* m) { delete m; }

const char* psim_model_last_error(const psim_model* m) { return m ? m->err.c_str() : "null model"; }

int psim_model_sizes(const psim_model* m, int32_t* n_sensors, int64_t* num_record, int64_t* num_runs) {
    if (!m) { return PSIM_E_INVALID; }
    if (n_sensors) { *n_sensors = (int32_t)m->model->sensors().size(); }
    if (num_record) { *num_record = (int64_t)m->model->numRecord(); }
    if (num_runs) { *num_runs = (int64_t)m->model->numRuns(); }
    return PSIM_OK;
}

int psim_model_run(psim_model* m, psim_gpu_ctx* ctx, uint64_t seed) {
    if (!m || !ctx) { return PSIM_E_INVALID; }
    try {
        m->model->runSimulation(ctx, seed);
        return PSIM_OK;
    } catch (const std::exception& e) { return model_fail(m, PSIM_E_DEVICE, e.what()); }
}

int psim_model_set_stream(psim_model* m, int32_t stream) {
    if (!m || stream < 0 || stream > 0xffff) { return PSIM_E_INVALID; }
    m->model->setStream(stream);
    return PSIM_OK;
}

int psim_batch_run(psim_model* const* models, int32_t n, psim_gpu_ctx* ctx, uint64_t seed, int64_t* counters,
    char* err, int errlen) {
    auto fail_msg = [&](int code, const std::string& msg) {
        if (err && errlen > 0) {
            std::strncpy(err, msg.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = '\0';
        }
        return code;
    };
    if (!models || n <= 0 || !ctx) { return fail_msg(PSIM_E_INVALID, "batch needs models and an engine"); }
    std::vector<psim::Model*> list;
    for (int32_t i = 0; i < n; ++i) {
        if (!models[i]) { return fail_msg(PSIM_E_INVALID, "null model in batch"); }
        list.push_back(models[i]->model.get());
    }
    try {
        std::array<int64_t, PSIM_NUM_COUNTERS> c{};
        psim::Model::runBatch(list, ctx, seed, c);
        if (counters) { std::memcpy(counters, c.data(), sizeof c); }
        return PSIM_OK;
    } catch (const std::exception& e) { return fail_msg(PSIM_E_DEVICE, e.what()); }
}

int psim_model_results(const psim_model* m, int64_t run, double* out, double* temps, double* fluxes, int64_t* counts,
    double* raw_flux, int64_t* counters) {
    if (!m) { return PSIM_E_INVALID; }
    const auto& runs = m->model->runs();
    if (run < 0 || (size_t)run >= runs.size()) { return PSIM_E_STATE; }
    const auto& r = runs[(size_t)run];
    const size_t R = m->model->numRecord();
    for (size_t i = 0; i < r.measurements.size(); ++i) {
        const auto& s = r.measurements[i];
        if (out) {
            const double v[6] = { s.t_steady, s.std_t_steady, s.x_flux, s.std_x_flux, s.y_flux, s.std_y_flux };
            std::memcpy(out + 6 * i, v, sizeof v);
        }
        if (temps) { std::memcpy(temps + i * R, s.final_temps.data(), sizeof(double) * R); }
        if (fluxes) {
            for (size_t k = 0; k < R; ++k) {
                fluxes[2 * (i * R + k)] = s.final_fluxes[k][0];
                fluxes[2 * (i * R + k) + 1] = s.final_fluxes[k][1];
            }
        }
    }
    if (counts) { std::memcpy(counts, r.counts.data(), sizeof(int64_t) * r.counts.size()); }
    if (raw_flux) { std::memcpy(raw_flux, r.flux.data(), sizeof(double) * r.flux.size()); }
    if (counters) { std::memcpy(counters, r.counters.data(), sizeof(int64_t) * PSIM_NUM_COUNTERS); }
    return PSIM_OK;
}

int psim_model_budget(const psim_model* m, double* eff_energy, int64_t* total_phonons) {
    if (!m) { return PSIM_E_INVALID; }
    const auto& runs = m->model->runs();
    if (runs.empty()) { return PSIM_E_STATE; }
    if (eff_energy) { *eff_energy = runs.back().eff_energy; }
    if (total_phonons) { *total_phonons = runs.back().total_phonons; }
    return PSIM_OK;
}

int psim_model_prepare(psim_model* m, uint64_t seed, double* eff_energy, int64_t* total_phonons, int64_t* counts,
    int32_t counts_cap, int32_t* n_sources) {
    if (!m) { return PSIM_E_INVALID; }
    try {
        const double eff = m->model->prepareBudget(seed);
        const auto& sim = m->model->simulator();
        if (eff_energy) { *eff_energy = eff; }
        if (total_phonons) { *total_phonons = sim.totalPhonons(); }
        const auto& sc = sim.sourceCounts();
        if (n_sources) { *n_sources = (int32_t)sc.size(); }
        if (counts) {
            for (size_t i = 0; i < sc.size() && (int32_t)i < counts_cap; ++i) { counts[i] = sc[i]; }
        }
        return PSIM_OK;
    } catch (const std::exception& e) { return model_fail(m, PSIM_E_MODEL, e.what()); }
}

int psim_model_upload(psim_model* m, psim_gpu_ctx* ctx, uint64_t seed, double* eff_energy, int64_t* total_phonons) {
    if (!m || !ctx) { return PSIM_E_INVALID; }
    try {
        const double eff = m->model->prepareAndUpload(ctx, seed);
        if (eff_energy) { *eff_energy = eff; }
        if (total_phonons) { *total_phonons = m->model->simulator().totalPhonons(); }
        return PSIM_OK;
    } catch (const std::exception& e) { return model_fail(m, PSIM_E_DEVICE, e.what()); }
}

int psim_model_export(const psim_model* m, const char* json_path, double time_s) {
    if (!m || !json_path) { return PSIM_E_INVALID; }
    try {
        m->model->exportResults(json_path, time_s);
        return PSIM_OK;
    } catch (const std::exception& e) {
        const_cast<psim_model*>(m)->err = e.what();
        return PSIM_E_IO;
    }
}

}  // extern "C"
