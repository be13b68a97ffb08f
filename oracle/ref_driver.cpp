// Test-infrastructure driver for the reference psim build in oracle/_ref/.
//
// It links the reference's own sources (psim/src/*.cpp except main.cpp and
// inputManager.cpp, which needs nlohmann_json), builds a Model through the
// reference's public API in the same order as InputManager::deserialize
// (psim/src/inputManager.cpp:55-105), runs Model::runSimulation once, and
// dumps the reference's internal state as JSON for the oracle's checks:
// raw tallies, the SensorInterpreter results, material tables and the cell
// surfaces.  Private members are read by redefining the access keywords after
// all standard headers are included; no reference file is modified.
//
// Input is a whitespace-separated model description written by
// oracle/refmodel.py from a psim JSON file (doubles printed with %.17g).

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <execution>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <optional>
#include <random>
#include <ranges>
#include <sstream>
#include <stack>
#include <string>
#include <unordered_map>
#include <variant>
#include <vector>

#define private public
#define protected public
#include "psim/cell.h"
#include "psim/geometry.h"
#include "psim/material.h"
#include "psim/model.h"
#include "psim/sensor.h"
#undef private
#undef protected

namespace {

struct MaterialIn {
    std::string name;
    DispersionData d;
    RelaxationData r;
};

void put_double(std::ostream& os, double v) {
    if (std::isnan(v)) {
        os << "\"nan\"";
    } else if (std::isinf(v)) {
        os << (v > 0 ? "\"inf\"" : "\"-inf\"");
    } else {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.17g", v);
        os << buf;
    }
}

void put_line(std::ostream& os, const Geometry::Line& l) {
    os << "[";
    put_double(os, l.p1.x); os << ",";
    put_double(os, l.p1.y); os << ",";
    put_double(os, l.p2.x); os << ",";
    put_double(os, l.p2.y); os << ",";
    put_double(os, l.slope); os << ",";
    put_double(os, l.intercept); os << ",";
    put_double(os, l.length); os << "]";
}

void put_vec(std::ostream& os, const Geometry::Vector2D& v) {
    os << "[";
    put_double(os, v.x); os << ",";
    put_double(os, v.y); os << "]";
}

void dump_table(std::ostream& os, const TableData& t) {
    os << "{\"sum\":";
    put_double(os, t.cumul_sum);
    os << ",\"cdf\":[";
    for (std::size_t i = 0; i < t.table.size(); ++i) {
        if (i) os << ",";
        put_double(os, t.table[i].first);
    }
    os << "],\"la_frac\":[";
    for (std::size_t i = 0; i < t.table.size(); ++i) {
        if (i) os << ",";
        put_double(os, t.table[i].second);
    }
    os << "]}";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: psim_ref <model.txt> <out.json> [--tables T1,T2,...] [--geometry] "
                     "[--export <model.json> <seconds>]\n";
        return 2;
    }
    std::vector<double> table_temps;
    bool dump_geometry = false;
    std::string export_path;  // Model::exportResults' file argument (model.cpp:184-186)
    double export_time = 0.;
    for (int i = 3; i < argc; ++i) {
        if (std::strcmp(argv[i], "--tables") == 0 && i + 1 < argc) {
            std::stringstream ss(argv[++i]);
            std::string tok;
            while (std::getline(ss, tok, ',')) table_temps.push_back(std::stod(tok));
        } else if (std::strcmp(argv[i], "--geometry") == 0) {
            dump_geometry = true;
        } else if (std::strcmp(argv[i], "--export") == 0 && i + 2 < argc) {
            export_path = argv[++i];
            export_time = std::stod(argv[++i]);
        }
    }
    std::ifstream in(argv[1]);
    std::string tag;
    std::size_t num_runs, num_meas, num_phonons, step_interval;
    double sim_time, t_eq;
    int sim_type, phasor;
    in >> tag >> num_runs >> num_meas >> num_phonons >> sim_time >> t_eq >> sim_type >> step_interval >> phasor;
    std::size_t n;
    in >> tag >> n;
    std::vector<MaterialIn> mats(n);
    for (auto& m : mats) {
        in >> m.name >> m.d.LA_data[0] >> m.d.LA_data[1] >> m.d.LA_data[2] >> m.d.TA_data[0] >> m.d.TA_data[1]
            >> m.d.TA_data[2] >> m.d.w_max_la >> m.d.w_max_ta >> m.r.b_l >> m.r.b_tn >> m.r.b_tu >> m.r.b_i >> m.r.w;
    }
    std::size_t n_sensors, n_cells, n_emit;
    in >> tag >> n_sensors;
    struct SIn { std::size_t id; std::string mat; double t_init; };
    std::vector<SIn> sensors(n_sensors);
    for (auto& s : sensors) in >> s.id >> s.mat >> s.t_init;
    in >> tag >> n_cells;
    struct CIn { double x1, y1, x2, y2, x3, y3; std::size_t sid; double spec; };
    std::vector<CIn> cells(n_cells);
    for (auto& c : cells) in >> c.x1 >> c.y1 >> c.x2 >> c.y2 >> c.x3 >> c.y3 >> c.sid >> c.spec;
    in >> tag >> n_emit;
    struct EIn { double x1, y1, x2, y2, temp, duration, start; };
    std::vector<EIn> emits(n_emit);
    for (auto& e : emits) in >> e.x1 >> e.y1 >> e.x2 >> e.y2 >> e.temp >> e.duration >> e.start;
    if (!in) {
        std::cerr << "bad model file\n";
        return 2;
    }

    std::ofstream out(argv[2]);
    try {
        const SimulationType type = sim_type == 1 ? SimulationType::Periodic
                                    : sim_type == 2 ? SimulationType::Transient
                                                    : SimulationType::SteadyState;
        ModelParams params{ num_runs, n_cells, n_sensors, num_meas, num_phonons, sim_time, t_eq, phasor != 0 };
        Model model(params);
        (type == SimulationType::SteadyState) ? model.setSimulationType(type)
                                              : model.setSimulationType(type, step_interval);
        std::size_t mid = 0;
        for (const auto& m : mats) {
            Material mat(mid++, m.d, m.r);
            if (t_eq == 0.) mat.setFullSimulation();
            model.addMaterial(m.name, mat);
        }
        for (const auto& s : sensors) model.addSensor(s.id, s.mat, s.t_init, type);
        for (const auto& c : cells) {
            Geometry::Triangle t{ { c.x1, c.y1 }, { c.x2, c.y2 }, { c.x3, c.y3 } };
            model.addCell(std::move(t), c.sid, c.spec);
        }
        for (const auto& e : emits) {
            if (!model.setEmitSurface({ e.x1, e.y1 }, { e.x2, e.y2 }, e.temp, e.duration, e.start)) {
                throw std::runtime_error("Unable to add emitting surface.");
            }
        }

        const auto t0 = std::chrono::steady_clock::now();
        model.runSimulation();
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // The reference's own writer (outputManager.cpp:13-38), as main.cpp:21 calls it.
        if (!export_path.empty()) model.exportResults(export_path, export_time);

        // The budget's energy per phonon: recompute the total energy with every controller's
        // t_steady_ back at t_init (the state before resetRequired ran).
        std::vector<double> t_steady;
        for (auto& s : model.sensors_) {
            t_steady.push_back(s.controller_->t_steady_);
            s.controller_->t_steady_ = s.controller_->t_init_;
        }
        const double eff_energy0 = model.getTotalInitialEnergy() / static_cast<double>(model.num_phonons_);
        for (std::size_t i = 0; i < model.sensors_.size(); ++i) model.sensors_[i].controller_->t_steady_ = t_steady[i];

        out << "{\"ok\":true,\"wall_s\":";
        put_double(out, wall);
        out << ",\"eff_energy\":";
        put_double(out, model.interpreter_.eff_energy_);
        out << ",\"eff_energy0\":";
        put_double(out, eff_energy0);
        out << ",\"t_steady\":[";
        for (std::size_t i = 0; i < t_steady.size(); ++i) {
            if (i) out << ",";
            put_double(out, t_steady[i]);
        }
        out << "]";
        out << ",\"t_eq\":";
        put_double(out, model.interpreter_.t_eq_);
        out << ",\"lb\":";
        put_double(out, model.interpreter_.lb_);
        out << ",\"ub\":";
        put_double(out, model.interpreter_.ub_);
        out << ",\"total_phonons\":" << model.simulator_.total_phonons_;
        out << ",\"sensors\":[";
        bool first = true;
        for (const auto& s : model.sensors_) {
            if (!first) out << ",";
            first = false;
            out << "{\"id\":" << s.getID() << ",\"area\":";
            put_double(out, s.getArea());
            out << ",\"heat_capacity\":";
            put_double(out, s.getHeatCapacity());
            out << ",\"energies\":[";
            const auto& en = s.getEnergies();
            for (std::size_t i = 0; i < en.size(); ++i) out << (i ? "," : "") << en[i];
            out << "],\"fluxes\":[";
            const auto& fl = s.getFluxes();
            for (std::size_t i = 0; i < fl.size(); ++i) {
                if (i) out << ",";
                out << "[";
                put_double(out, fl[i][0]);
                out << ",";
                put_double(out, fl[i][1]);
                out << "]";
            }
            out << "]}";
        }
        out << "],\"measurements\":[";
        first = true;
        for (const auto& m : model.outputManager_.measurements_[0]) {
            if (!first) out << ",";
            first = false;
            out << "{\"id\":" << m.id << ",\"t_steady\":";
            put_double(out, m.t_steady);
            out << ",\"std_t_steady\":";
            put_double(out, m.std_t_steady);
            out << ",\"x_flux\":";
            put_double(out, m.x_flux);
            out << ",\"std_x_flux\":";
            put_double(out, m.std_x_flux);
            out << ",\"y_flux\":";
            put_double(out, m.y_flux);
            out << ",\"std_y_flux\":";
            put_double(out, m.std_y_flux);
            out << ",\"final_temps\":[";
            for (std::size_t i = 0; i < m.final_temps.size(); ++i) {
                if (i) out << ",";
                put_double(out, m.final_temps[i]);
            }
            out << "]}";
        }
        out << "]";
        if (!table_temps.empty()) {
            const Material& mat = model.materials_.begin()->second;
            out << ",\"temps\":[";
            for (std::size_t i = 0; i < mat.temps_.size(); ++i) {
                if (i) out << ",";
                put_double(out, mat.temps_[i]);
            }
            out << "],\"frequencies\":[";
            for (std::size_t i = 0; i < mat.frequencies_.size(); ++i) {
                if (i) out << ",";
                put_double(out, mat.frequencies_[i]);
            }
            out << "],\"vel_la\":[";
            for (std::size_t i = 0; i < mat.velocities_la_.size(); ++i) {
                if (i) out << ",";
                put_double(out, mat.velocities_la_[i]);
            }
            out << "],\"vel_ta\":[";
            for (std::size_t i = 0; i < mat.velocities_ta_.size(); ++i) {
                if (i) out << ",";
                put_double(out, mat.velocities_ta_[i]);
            }
            out << "],\"base_sums\":[";
            for (std::size_t i = 0; i < mat.base_tables_.size(); ++i) {
                if (i) out << ",";
                put_double(out, mat.base_tables_[i].cumul_sum);
            }
            out << "],\"tables\":[";
            for (std::size_t i = 0; i < table_temps.size(); ++i) {
                if (i) out << ",";
                const auto idx = mat.getTempIndex(table_temps[i]);
                out << "{\"temp\":";
                put_double(out, table_temps[i]);
                out << ",\"index\":" << idx << ",\"base\":";
                dump_table(out, mat.base_tables_[idx]);
                out << ",\"emit\":";
                dump_table(out, mat.emit_tables_[idx]);
                out << ",\"scatter\":";
                dump_table(out, mat.scatter_tables_[idx]);
                out << "}";
            }
            out << "]";
        }
        if (dump_geometry) {
            out << ",\"cells\":[";
            for (std::size_t c = 0; c < model.cells_.size(); ++c) {
                const Cell& cell = model.cells_[c];
                if (c) out << ",";
                out << "{\"area\":";
                put_double(out, cell.getArea());
                out << ",\"edges\":[";
                for (std::size_t b = 0; b < 3; ++b) {
                    const auto& cs = cell.boundaries_[b];
                    if (b) out << ",";
                    out << "{\"line\":";
                    put_line(out, cs.main_surface_.getSurfaceLine());
                    out << ",\"normal\":";
                    put_vec(out, cs.main_surface_.getNormal());
                    out << ",\"spec\":";
                    put_double(out, cs.main_surface_.getSpecularity());
                    out << ",\"transitions\":[";
                    for (std::size_t t = 0; t < cs.transition_sub_surfaces_.size(); ++t) {
                        const auto& ts = cs.transition_sub_surfaces_[t];
                        if (t) out << ",";
                        // The neighbour is identified by its index in the model's cell vector.
                        const auto nb = static_cast<std::size_t>(&ts.cell_ - model.cells_.data());
                        out << "{\"line\":";
                        put_line(out, ts.getSurfaceLine());
                        out << ",\"normal\":";
                        put_vec(out, ts.getNormal());
                        out << ",\"cell\":" << nb << "}";
                    }
                    out << "],\"emits\":[";
                    for (std::size_t e = 0; e < cs.emit_sub_surfaces_.size(); ++e) {
                        const auto& es = cs.emit_sub_surfaces_[e];
                        if (e) out << ",";
                        out << "{\"line\":";
                        put_line(out, es.getSurfaceLine());
                        out << ",\"normal\":";
                        put_vec(out, es.getNormal());
                        out << ",\"temp\":";
                        put_double(out, es.getTemp());
                        out << "}";
                    }
                    out << "]}";
                }
                out << "]}";
            }
            out << "]";
        }
        out << "}\n";
    } catch (const std::exception& e) {
        std::string msg = e.what();
        std::replace(msg.begin(), msg.end(), '"', '\'');
        std::replace(msg.begin(), msg.end(), '\n', ' ');
        out << "{\"ok\":false,\"error\":\"" << msg << "\"}\n";
        return 1;
    }
    return 0;
}
