// bnn_flip_solver.hpp -- header-only C++ adapter over the C ABI (bnn_flip.h) with the method names of
// Banana's FLIP_Solver<3,float> (FLIP_Solver.h:30-60, PIC_Solver.h:40-100, ParticleSolver.h:39-105),
// so reference-style callers (loadScene / doSimulation / doFrameSimulation) drop in.
//
// Errors throw bnn::flip_error (status + message) instead of exiting the process.
#pragma once

#include <algorithm>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bnn_flip.h"

namespace bnn {

class flip_error : public std::runtime_error {
public:
    flip_error(bnn_status s, const std::string& m) : std::runtime_error(m), status(s) {}
    bnn_status status;
};

class FLIP_Solver3D {
public:
    // FLIP_Solver.h:37-41
    static std::string solverName() { return "FLIP_3DSolver"; }
    std::string        getSolverName() const { return solverName(); }
    std::string        getSolverDescription() const { return "Simulation using FLIP-3D Solver"; }

    FLIP_Solver3D() = default;
    explicit FLIP_Solver3D(const bnn_flip_params& p) { create(p); }
    ~FLIP_Solver3D() { reset(); }
    FLIP_Solver3D(const FLIP_Solver3D&) = delete;
    FLIP_Solver3D& operator=(const FLIP_Solver3D&) = delete;
    FLIP_Solver3D(FLIP_Solver3D&& o) noexcept : m_h(std::exchange(o.m_h, nullptr)) {}
    FLIP_Solver3D& operator=(FLIP_Solver3D&& o) noexcept {
        if(this != &o) {
            reset();
            m_h = std::exchange(o.m_h, nullptr);
        }
        return *this;
    }

    // ParticleSolver::loadScene (ParticleSolver.cpp:30-168): scene JSON, seeded particle generation
    void loadScene(const std::string& sceneFile, uint64_t seed = 0) {
        std::ifstream f(sceneFile);
        if(!f) throw flip_error(BNN_ERR_INVALID_ARG, "Cannot open scene file: " + sceneFile);
        std::stringstream ss;
        ss << f.rdbuf();
        loadSceneFromString(ss.str(), seed);
    }
    // PIC_Solver::generateParticles (PIC_Solver.cpp:40-66): the latest memory state is loaded when
    // LoadMemoryState is set and one exists; otherwise the particles are generated (and sorted) and
    // frame 0 is saved when SaveFrameData is set
    void loadSceneFromString(const std::string& json, uint64_t seed = 0) {
        reset();
        bnn_flip*  h  = nullptr;
        bnn_status st = bnn_flip_create_from_json(json.c_str(), seed, &h);
        if(st != BNN_OK) throw flip_error(st, bnn_flip_last_error(nullptr));
        m_h = h;
        check(bnn_io_from_json(json.c_str(), &m_io));
        const bnn_flip_params p = params();
        const int32_t idx = m_io.load_memory_state ? bnn_latest_state_index(m_io.data_path, m_io.memory_state_folder, p.start_frame, p.final_frame)
                                                   : -1;
        if(idx >= 0) {
            loadMemoryState(filePath(m_io.memory_state_folder, idx), (uint32_t)idx);
            m_loadedState = idx;
        } else {
            saveFrameData();
        }
    }
    const bnn_io_params& ioParams() const { return m_io; }
    void                 setIoParams(const bnn_io_params& io) { m_io = io; }
    int32_t              loadedStateIndex() const { return m_loadedState; }
    void create(const bnn_flip_params& p) {
        reset();
        bnn_flip*  h  = nullptr;
        bnn_status st = bnn_flip_create(&p, &h);
        if(st != BNN_OK) throw flip_error(st, bnn_flip_last_error(nullptr));
        m_h = h;
    }

    // ParticleSolver::doSimulation (ParticleSolver.cpp:182-198)
    void doSimulation() {
        const bnn_flip_params p     = params();
        const uint32_t        done  = finishedFrame();
        const uint32_t        start = p.start_frame <= 1 ? done + 1 : std::min<uint32_t>(p.start_frame, done + 1);
        for(uint32_t frame = start; frame <= p.final_frame; ++frame) doFrameSimulation(frame);
    }
    // ParticleSolver::doFrameSimulation (ParticleSolver.cpp:200-219): sortParticles + advanceFrame,
    // which ends with saveFrameData and saveMemoryState (PIC_Solver.cpp:307-310)
    void doFrameSimulation(uint32_t frame) {
        check(bnn_flip_do_frame_simulation(h(), frame));
        saveFrameData();
        saveMemoryState();
    }

    // PIC_Solver::saveFrameData / saveMemoryState / loadMemoryState (PIC_Solver.cpp:158-262), files
    // named by DataIO::getFilePath; each returns the frame index written, or -1 when switched off
    int32_t saveFrameData() {
        if(!m_io.save_frame_data) return -1;
        const uint32_t f = finishedFrame();
        check(bnn_flip_save_frame_data(h(), filePath(m_io.frame_data_folder, (int32_t)f).c_str(), m_io.save_velocity, m_io.save_object_index));
        return (int32_t)f;
    }
    int32_t saveMemoryState() {
        const uint32_t f = finishedFrame();
        if(!m_io.save_memory_state || (m_io.frame_per_state && f % m_io.frame_per_state != 0)) return -1;
        check(bnn_flip_save_memory_state(h(), filePath(m_io.memory_state_folder, (int32_t)f).c_str()));
        return (int32_t)f;
    }
    void loadMemoryState(const std::string& path, uint32_t frame) { check(bnn_flip_load_memory_state(h(), path.c_str(), frame)); }

    // PIC_Solver override points (PIC_Solver.h:64-70, FLIP_Solver.h:49-52)
    void  advanceFrame() { check(bnn_flip_advance_frame(h())); }
    void  sortParticles() { check(bnn_flip_sort_particles(h())); }
    void  advanceVelocity(float dt) { check(bnn_flip_advance_velocity(h(), dt)); }
    void  moveParticles(float dt) { check(bnn_flip_move_particles(h(), dt)); }
    void  mapParticles2Grid() { check(bnn_flip_map_particles_to_grid(h())); }
    void  mapGrid2Particles() { check(bnn_flip_map_grid_to_particles(h())); }
    void  extrapolateVelocity() { check(bnn_flip_extrapolate_velocity(h())); }
    void  constrainGridVelocity() { check(bnn_flip_constrain_grid_velocity(h())); }
    void  computeCellWeights() { check(bnn_flip_compute_cell_weights(h())); }
    void  computeFluidSDF() { check(bnn_flip_compute_fluid_sdf(h())); }
    void  pressureProjection(float dt) { check(bnn_flip_pressure_projection(h(), dt)); }
    float timestepCFL() {
        float dt = 0;
        check(bnn_flip_timestep_cfl(h(), &dt));
        return dt;
    }
    // one body of the advanceFrame substep loop (PIC_Solver.cpp:292-294)
    void step(float dt) { check(bnn_flip_step(h(), dt)); }

    // particle data (ParticleSimulationData::positions / velocities, AoS float triples)
    void setParticles(const std::vector<float>& pos, const std::vector<float>& vel) {
        if(pos.size() != vel.size() || pos.size() % 3) throw flip_error(BNN_ERR_INVALID_ARG, "positions/velocities size mismatch");
        check(bnn_flip_set_particles(h(), pos.data(), vel.data(), pos.size() / 3));
    }
    uint64_t getNParticles() const { return bnn_flip_num_particles(m_h); }
    void     getParticles(std::vector<float>& pos, std::vector<float>& vel) {
        pos.resize(getNParticles() * 3);
        vel.resize(getNParticles() * 3);
        check(bnn_flip_get_particles(h(), pos.data(), vel.data()));
    }

    bnn_flip_params params() const {
        bnn_flip_params p;
        check(bnn_flip_get_params(h(), &p));
        return p;
    }
    bnn_flip_stats stats() {
        bnn_flip_stats s;
        check(bnn_flip_get_stats(h(), &s));
        return s;
    }
    uint32_t  finishedFrame() {
        return stats().finished_frame;
    }
    bnn_flip* handle() const { return m_h; }

private:
    bnn_flip*     m_h = nullptr;
    bnn_io_params m_io = default_io();
    int32_t       m_loadedState = -1;
    static bnn_io_params default_io() {
        bnn_io_params io;
        bnn_io_default_params(&io);
        io.load_memory_state = 0;  // set from the scene by loadScene (the reference's default there is true)
        return io;
    }
    std::string filePath(const char* folder, int32_t frame) const {
        char buf[1024];
        if(bnn_data_file_path(m_io.data_path, folder, frame, buf, sizeof(buf)) != BNN_OK)
            throw flip_error(BNN_ERR_INVALID_ARG, "data path too long");
        return buf;
    }
    bnn_flip* h() const {
        if(!m_h) throw flip_error(BNN_ERR_STATE, "solver not created (call loadScene or create first)");
        return m_h;
    }
    void check(bnn_status st) const {
        if(st != BNN_OK) throw flip_error(st, bnn_flip_last_error(m_h));
    }
    void reset() {
        if(m_h) bnn_flip_destroy(m_h);
        m_h = nullptr;
    }
};

}  // namespace bnn
