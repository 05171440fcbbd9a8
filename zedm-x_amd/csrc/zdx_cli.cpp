// zdx — command-line driver with the interface of the reference's
// `zdem_cpu --config path.txt` (src/main.cpp:1691-1708, 1872-2502): the same
// config file, the same per-output-step files
// (output_dir/prefix_NNNNNN.vtk and .txt) and the same step line
//   step=S contacts=C step_ms=.. tri_ms=.. rebuild_ms=.. output_ms=.. tri_checks=.. segments=.. comps=.. loops=..
// followed by "done".  tri_ms and rebuild_ms are the device times of the
// narrow phase (with the world transform) and of the topology + forces.
// Options beyond the reference's: --mode plain|batched (default batched),
// --device N (HIP ordinal, default 0).
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "../../include/zdx.h"
#include "zdx_internal.h"

namespace {

void usage() { std::cout << "Usage: zdx --config path.txt [--mode plain|batched] [--device N]\n"; }

bool parse_args(int argc, char** argv, std::string& config, int& mode, int& device) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--config" && i + 1 < argc) {
            config = argv[++i];
        } else if (a == "--mode" && i + 1 < argc) {
            std::string m = argv[++i];
            if (m == "plain") mode = ZDX_MODE_PLAIN;
            else if (m == "batched") mode = ZDX_MODE_BATCHED;
            else { std::cerr << "Unknown mode: " << m << "\n"; return false; }
        } else if (a == "--device" && i + 1 < argc) {
            device = std::atoi(argv[++i]);
        } else if (a == "--help") {
            return false;
        } else {
            std::cerr << "Unknown arg: " << a << "\n";
            return false;
        }
    }
    return !config.empty();
}

std::string out_name(const zdx::SimConfig& cfg, int step, const char* ext) {
    std::ostringstream o;
    o << cfg.vtk_prefix << "_" << std::setw(6) << std::setfill('0') << step << ext;
    return (std::filesystem::path(cfg.output_dir) / o.str()).string();
}

}  // namespace

int main(int argc, char** argv) {
    std::string config;
    int mode = ZDX_MODE_BATCHED, device = 0;
    if (!parse_args(argc, argv, config, mode, device)) {
        usage();
        return 1;
    }
    zdx::SimConfig cfg;
    if (!zdx::parse_config(config, cfg)) {
        std::cerr << "Failed to open config: " << config << "\n";
        return 1;
    }
    zdx_ctx* ctx = nullptr;
    if (zdx_create(&ctx, device) != ZDX_OK) {
        std::cerr << "zdx: cannot open HIP device " << device << "\n";
        return 1;
    }
    zdx_sim* sim = nullptr;
    if (zdx_sim_load_config(ctx, config.c_str(), nullptr, &sim) != ZDX_OK) {
        std::cerr << "zdx: " << zdx_last_error(ctx) << "\n";
        zdx_destroy(ctx);
        return 1;
    }
    std::filesystem::create_directories(cfg.output_dir);
    const int interval = cfg.output_interval > 0 ? cfg.output_interval : 1;
    int rc = 0;
    double st[7 + ZDX_N_COUNTERS];
    for (int step = 0; step < cfg.steps; ++step) {
        const auto t0 = std::chrono::steady_clock::now();
        int64_t ctr[ZDX_N_COUNTERS] = {0};
        if (zdx_sim_step(sim, 1, mode, ctr) != ZDX_OK) {
            std::cerr << "zdx: step " << step << ": " << zdx_last_error(ctx) << "\n";
            rc = 1;
            break;
        }
        zdx_sim_stage_times(sim, st, 7 + ZDX_N_COUNTERS);
        if (step % interval != 0) continue;
        const auto to0 = std::chrono::steady_clock::now();
        if (zdx_sim_write_vtk(sim, out_name(cfg, step, ".vtk").c_str()) != ZDX_OK)
            std::cerr << "Failed to write VTK: " << out_name(cfg, step, ".vtk") << "\n";
        if (zdx_sim_write_state_txt(sim, out_name(cfg, step, ".txt").c_str()) != ZDX_OK)
            std::cerr << "Failed to write particle state: " << out_name(cfg, step, ".txt") << "\n";
        const auto t1 = std::chrono::steady_clock::now();
        auto ms = [](auto d) { return (long long)std::chrono::duration_cast<std::chrono::milliseconds>(d).count(); };
        std::cout << "step=" << step << " contacts=" << ctr[ZDX_CTR_CONTACTS] << " step_ms=" << ms(t1 - t0)
                  << " tri_ms=" << (long long)(st[0] + st[2]) << " rebuild_ms=" << (long long)st[3]
                  << " output_ms=" << ms(t1 - to0) << " tri_checks=" << ctr[ZDX_CTR_TRI_CHECKS]
                  << " segments=" << ctr[ZDX_CTR_SEGMENTS] << " comps=" << ctr[ZDX_CTR_COMPS]
                  << " loops=" << ctr[ZDX_CTR_LOOPS] << "\n";
    }
    if (rc == 0) std::cout << "done\n";
    zdx_sim_destroy(sim);
    zdx_destroy(ctx);
    return rc;
}
