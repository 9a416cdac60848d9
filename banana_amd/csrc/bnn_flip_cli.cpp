// bnn_flip_cli -- command-line driver, the equivalent of Simulator::doSimulation
// (Applications/QtApps/ParticleSimulation/Source/Simulator.cpp:42-69) without the UI thread:
// load a scene JSON, run frames StartFrame..FinalFrame, print one JSON line of stats per frame.
//
//   bnn_flip_cli <scene.json> [--seed N] [--frames K] [--dump particles.bin]
//
// --frames overrides FinalFrame; --dump writes n (uint64) then positions and velocities
// (n*3 float32 each) after the last frame.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bnn_flip_solver.hpp"

int main(int argc, char** argv) {
    if(argc < 2) {
        std::fprintf(stderr, "usage: %s <scene.json> [--seed N] [--frames K] [--dump file]\n", argv[0]);
        return 2;
    }
    std::string scene = argv[1], dump;
    uint64_t    seed   = 0;
    long        frames = -1;
    for(int i = 2; i + 1 < argc; i += 2) {
        if(!std::strcmp(argv[i], "--seed")) seed = std::strtoull(argv[i + 1], nullptr, 10);
        else if(!std::strcmp(argv[i], "--frames")) frames = std::strtol(argv[i + 1], nullptr, 10);
        else if(!std::strcmp(argv[i], "--dump")) dump = argv[i + 1];
        else {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            return 2;
        }
    }
    try {
        bnn::FLIP_Solver3D solver;
        solver.loadScene(scene, seed);
        bnn_flip_params p = solver.params();
        const uint32_t  last = frames >= 0 ? (uint32_t)frames : p.final_frame;
        std::printf("{\"solver\":\"%s\",\"particles\":%llu}\n", solver.getSolverName().c_str(), (unsigned long long)solver.getNParticles());
        for(uint32_t frame = solver.finishedFrame() + 1; frame <= last; ++frame) {
            auto t0 = std::chrono::steady_clock::now();
            solver.doFrameSimulation(frame);
            const double         ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            const bnn_flip_stats s  = solver.stats();
            std::printf("{\"frame\":%u,\"ms\":%.3f,\"substeps\":%u,\"last_substep\":%.6g,\"pcg_iterations\":%u,\"pcg_success\":%d}\n", frame,
                        ms, s.substep_count, (double)s.last_substep, s.pcg_iterations, s.pcg_success);
            std::fflush(stdout);
        }
        if(!dump.empty()) {
            std::vector<float> pos, vel;
            solver.getParticles(pos, vel);
            FILE* f = std::fopen(dump.c_str(), "wb");
            if(!f) throw bnn::flip_error(BNN_ERR_INVALID_ARG, "cannot write " + dump);
            const uint64_t n = pos.size() / 3;
            std::fwrite(&n, sizeof(n), 1, f);
            std::fwrite(pos.data(), sizeof(float), pos.size(), f);
            std::fwrite(vel.data(), sizeof(float), vel.size(), f);
            std::fclose(f);
        }
    } catch(const bnn::flip_error& e) {
        std::fprintf(stderr, "bnn_flip_cli: %s (status %d)\n", e.what(), (int)e.status);
        return 1;
    }
    return 0;
}
