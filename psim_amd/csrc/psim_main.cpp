// psim_gpu: command-line driver with the reference's interface (psim/src/main.cpp:10-31):
// `psim_gpu file1.json [file2.json ...]` reads each model, runs it on GPU 0 and writes
// ss_<stem>.txt / per_<stem>.txt next to the input.  PSIM_SEED selects the Philox seed.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "psim_gpu.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cout << "Need filenames\n";
        std::cout << "done\n";
        return 0;
    }
    psim_gpu_ctx* ctx = nullptr;
    if (psim_gpu_create(&ctx, 0) != PSIM_OK) {
        std::cerr << "no GPU available\n";
        return 1;
    }
    const char* seed_env = std::getenv("PSIM_SEED");
    const uint64_t seed = seed_env ? std::strtoull(seed_env, nullptr, 10) : 0;
    int rc = 0;
    for (int i = 1; i < argc; ++i) {
        std::ifstream f(argv[i]);
        if (!f) {
            std::cerr << "Error opening file at \"" << argv[i] << "\"\n";
            std::cerr << "There was an error reading the data from the file at \"" << argv[i] << "\"\n";
            continue;
        }
        std::stringstream ss;
        ss << f.rdbuf();
        psim_model* model = nullptr;
        char err[1024];
        if (psim_model_from_json(&model, ss.str().c_str(), err, sizeof err) != PSIM_OK) {
            std::cerr << err << "\n";
            std::cerr << "There was an error reading the data from the file at \"" << argv[i] << "\"\n";
            continue;
        }
        const auto t0 = std::chrono::steady_clock::now();
        if (psim_model_run(model, ctx, seed) != PSIM_OK) {
            std::cerr << psim_model_last_error(model) << "\n";
            psim_model_destroy(model);
            rc = 1;
            continue;
        }
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "Time Taken: " << t << "[s]\n";
        if (psim_model_export(model, argv[i], t) != PSIM_OK) {
            std::cerr << psim_model_last_error(model) << "\n";
            rc = 1;
        }
        psim_model_destroy(model);
    }
    psim_gpu_destroy(ctx);
    std::cout << "done\n";
    return rc;
}
