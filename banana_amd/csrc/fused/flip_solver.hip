// flip_solver.hip -- the C ABI of the MI355X FLIP substep (include/bnn_flip.h).
//
// A bnn_flip handle is one Scene plus a Runner bound to it alone (its own HIP stream); a
// bnn_flip_batch is a Runner that binds the scenes of several handles for one call; a bnn_server is
// the request loop of flip_server.hpp. All three drive the same kernels (flip_kernels.hpp).
#include "flip_server.hpp"

#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../scene_json.hpp"

struct bnn_flip {
    std::unique_ptr<bnn::Runner> run;
    std::unique_ptr<bnn::Scene>  s;
    std::string                  err;
    void                         bindSelf() {
        run->timing = s->prm.timing;
        run->bind({s.get()});
    }
};

struct bnn_flip_batch {
    std::unique_ptr<bnn::Runner> run;
    std::string                  err;
};

struct bnn_server {
    std::unique_ptr<bnn::Server> srv;
    std::string                  err;
};

namespace {
thread_local std::string g_createError;

template<class H, class F> bnn_status guarded(H* h, std::string& err, F&& f) {
    try {
        err.clear();
        return f();
    } catch(const bnn::Error& e) {
        err = e.what();
        return e.code;
    } catch(const std::bad_alloc&) {
        err = "host out of memory";
        return BNN_ERR_OUT_OF_MEMORY;
    } catch(const std::exception& e) {
        err = e.what();
        return BNN_ERR_DEVICE;
    }
}

template<class F> bnn_status guard(bnn_flip* h, F&& f) {
    if(!h || !h->s) return BNN_ERR_INVALID_ARG;
    return guarded(h, h->err, [&] {
        h->bindSelf();
        return f(*h->s, *h->run);
    });
}

bnn_status solve_status(const bnn::Scene& s) {
    return (!s.lastOk && s.prm.exit_if_pressure_projection_failed) ? BNN_ERR_NOT_CONVERGED : BNN_OK;
}

bnn_status stepOne(bnn::Scene& s, bnn::Runner& R, float dt) {
    s.curDt = dt;
    R.step(bnn::Runner::kAll);
    return solve_status(s);
}
}  // namespace

extern "C" {

void bnn_flip_default_params(bnn_flip_params* p) {
    if(!p) return;
    std::memset(p, 0, sizeof(*p));
    for(int d = 0; d < 3; ++d) {
        p->box_min[d] = -1.f;  // SimulationParameters::domainBMin default (ParticleSolverData.h:103-104)
        p->box_max[d] = 1.f;
    }
    p->cell_size                      = 1.0f / 64.0f;
    p->ratio_cell_size_pradius        = 2.0f;
    p->n_expand_cells                 = 2;
    p->min_timestep                   = 1.0e-6f;
    p->max_timestep                   = 1.0f / 30.0f;
    p->cfl_factor                     = 1.0f;
    p->frame_duration                 = (float)(1.0 / 30);
    p->cg_relative_tolerance          = 1e-15f;
    p->max_cg_iteration               = 5000;
    p->pic_flip_ratio                 = 0.97f;
    p->advection_steps                = 1;
    p->gravity_type                   = 0;
    p->apply_gravity                  = 1;
    p->boundary_reflection_multiplier = 1.0f;
    p->sort_frequency                 = 10;
    p->pressure_mode                  = BNN_PRESSURE_FP32;
    p->refine_tolerance               = 1e-10;
    p->preconditioner                 = BNN_PRECOND_MG;
    p->timing                         = 0;
    p->mg_sweeps                      = 2;
    p->mg_omega                       = 1.8f;
    p->mg_coarsest_sweeps             = 24;
    p->start_frame                    = 1;
    p->final_frame                    = 1;
}

bnn_status bnn_flip_create(const bnn_flip_params* p, bnn_flip** out) {
    if(!p || !out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    try {
        bnn::Scene::validate(*p);  // argument errors before any device call
        auto h = std::make_unique<bnn_flip>();
        h->run = std::make_unique<bnn::Runner>();
        h->s   = std::make_unique<bnn::Scene>(*p, h->run->stream);
        h->bindSelf();
        if(p->boundary_sdf_at_init) {
            h->run->boundarySDF(bnn::Runner::kAll);
            h->run->sync();
        }
        *out = h.release();
        return BNN_OK;
    } catch(const bnn::Error& e) {
        g_createError = e.what();
        return e.code;
    } catch(const std::exception& e) {
        g_createError = e.what();
        return BNN_ERR_DEVICE;
    }
}

bnn_status bnn_scene_from_json(const char* json_text, bnn_flip_params* p, bnn_scene_spec* spec, char* warnings, size_t cap) {
    if(!json_text || !p || !spec) return BNN_ERR_INVALID_ARG;
    try {
        bnn_flip_default_params(p);
        const bnn::json::Value root = bnn::json::Parser(json_text).parse();
        const std::string      w    = bnn::scene_from_json(root, *p, *spec);
        if(warnings && cap) {
            std::strncpy(warnings, w.c_str(), cap - 1);
            warnings[cap - 1] = 0;
        }
        return BNN_OK;
    } catch(const bnn::SceneLoadError& e) {
        g_createError = e.what();
        return e.code;
    } catch(const std::exception& e) {
        g_createError = e.what();
        return BNN_ERR_INVALID_ARG;
    }
}

bnn_status bnn_flip_create_from_json(const char* json_text, uint64_t seed, bnn_flip** out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    bnn_flip_params p;
    bnn_scene_spec  sp;
    bnn_status      st = bnn_scene_from_json(json_text, &p, &sp, nullptr, 0);
    if(st != BNN_OK) return st;
    st = bnn_flip_create(&p, out);
    if(st != BNN_OK) return st;
    st = bnn_flip_generate_scene(*out, &sp, seed);
    if(st != BNN_OK) {
        g_createError = bnn_flip_last_error(*out);
        bnn_flip_destroy(*out);
        *out = nullptr;
    }
    return st;
}

bnn_status bnn_flip_destroy(bnn_flip* h) {
    if(!h) return BNN_ERR_INVALID_ARG;
    try {
        delete h;
    } catch(...) {
        return BNN_ERR_DEVICE;
    }
    return BNN_OK;
}

const char* bnn_flip_last_error(const bnn_flip* h) { return h ? h->err.c_str() : g_createError.c_str(); }

bnn_status bnn_flip_get_params(const bnn_flip* h, bnn_flip_params* p) {
    if(!h || !h->s || !p) return BNN_ERR_INVALID_ARG;
    *p = h->s->prm;
    return BNN_OK;
}

bnn_status bnn_flip_grid_dims(const bnn_flip* h, int32_t n[3]) {
    if(!h || !h->s || !n) return BNN_ERR_INVALID_ARG;
    n[0] = h->s->g.nx;
    n[1] = h->s->g.ny;
    n[2] = h->s->g.nz;
    return BNN_OK;
}

bnn_status bnn_flip_grid_origin(const bnn_flip* h, float bmin[3], float* cell_size) {
    if(!h || !h->s || !bmin || !cell_size) return BNN_ERR_INVALID_ARG;
    bmin[0]    = h->s->g.bminx;
    bmin[1]    = h->s->g.bminy;
    bmin[2]    = h->s->g.bminz;
    *cell_size = h->s->g.dx;
    return BNN_OK;
}

bnn_status bnn_flip_set_particles(bnn_flip* h, const float* pos, const float* vel, uint64_t n) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.setParticles(pos, vel, n, R.stream);
        return BNN_OK;
    });
}

bnn_status bnn_generate_scene_host(const bnn_flip_params* p, const bnn_scene_spec* spec, uint64_t seed, float* pos, float* vel,
                                   uint64_t cap, uint64_t* n) {
    if(!p || !spec || !n) return BNN_ERR_INVALID_ARG;
    if((pos == nullptr) != (vel == nullptr)) return BNN_ERR_INVALID_ARG;
    *n = bnn::generate_scene(*p, *spec, seed, pos, vel, pos ? cap : 0);
    return BNN_OK;
}

bnn_status bnn_flip_generate_scene(bnn_flip* h, const bnn_scene_spec* spec, uint64_t seed) {
    if(!spec) return BNN_ERR_INVALID_ARG;
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        const uint64_t     n = bnn::generate_scene(s.prm, *spec, seed, nullptr, nullptr, 0);
        std::vector<float> pos(n * 3), vel(n * 3);
        bnn::generate_scene(s.prm, *spec, seed, pos.data(), vel.data(), n);
        s.setParticles(pos.data(), vel.data(), n, R.stream);
        return BNN_OK;
    });
}

uint64_t bnn_flip_num_particles(const bnn_flip* h) { return (h && h->s) ? h->s->np : 0; }

bnn_status bnn_flip_get_particles(bnn_flip* h, float* pos, float* vel) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.getParticles(pos, vel, R.stream);
        return BNN_OK;
    });
}

bnn_status bnn_flip_step(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) { return stepOne(s, R, dt); });
}

bnn_status bnn_flip_advance_frame(bnn_flip* h) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        R.advanceFrame(bnn::Runner::kAll);
        return solve_status(s);
    });
}

bnn_status bnn_flip_do_frame_simulation(bnn_flip* h, uint32_t frame) {
    (void)frame;
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        if(s.prm.enable_sort_particle) throw bnn::Error(BNN_ERR_UNSUPPORTED, "EnableSortParticle (z-sort) is not built yet");
        R.advanceFrame(bnn::Runner::kAll);
        return solve_status(s);
    });
}

bnn_status bnn_flip_timestep_cfl(bnn_flip* h, float* dt) {
    if(!dt) return BNN_ERR_INVALID_ARG;
    return guard(h, [&](bnn::Scene&, bnn::Runner& R) {
        std::vector<float> ts;
        R.timestepCFL(bnn::Runner::kAll, ts);
        *dt = ts[0];
        return BNN_OK;
    });
}

bnn_status bnn_flip_move_particles(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.curDt = dt;
        R.markDesc();
        R.move(bnn::Runner::kAll);
        R.sync();
        return BNN_OK;
    });
}

bnn_status bnn_flip_sort_particles(bnn_flip* h) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner&) {
        if(!s.prm.enable_sort_particle) return BNN_OK;  // PIC_Solver.cpp:316-318
        throw bnn::Error(BNN_ERR_UNSUPPORTED, "EnableSortParticle (z-sort) is not built yet");
    });
}

bnn_status bnn_flip_advance_velocity(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.curDt = dt;
        R.markDesc();
        R.beginStats();
        R.advanceVelocity(bnn::Runner::kAll);
        R.endStats(bnn::Runner::kAll);
        return solve_status(s);
    });
}

#define BNN_STAGE_CALL(name, body)                                      \
    bnn_status name(bnn_flip* h) {                                      \
        return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {            \
            (void)s;                                                    \
            body;                                                       \
            R.sync();                                                   \
            return BNN_OK;                                              \
        });                                                             \
    }

BNN_STAGE_CALL(bnn_flip_map_particles_to_grid, R.p2g(bnn::Runner::kAll))
BNN_STAGE_CALL(bnn_flip_map_grid_to_particles, if(s.np) R.g2p(bnn::Runner::kAll))
BNN_STAGE_CALL(bnn_flip_extrapolate_velocity, R.extrapolate(bnn::Runner::kAll))
BNN_STAGE_CALL(bnn_flip_constrain_grid_velocity, R.constrain(bnn::Runner::kAll))
BNN_STAGE_CALL(bnn_flip_compute_fluid_sdf, R.fluidSDF(bnn::Runner::kAll))
BNN_STAGE_CALL(bnn_flip_compute_cell_weights, s.weightsDirty = true; R.weights(bnn::Runner::kAll))
BNN_STAGE_CALL(bnn_flip_bin_particles, R.bin(bnn::Runner::kAll))

bnn_status bnn_flip_pressure_projection(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.curDt = dt;
        R.markDesc();
        R.beginStats();
        R.pressureProjection(bnn::Runner::kAll);
        R.endStats(bnn::Runner::kAll);
        return solve_status(s);
    });
}

bnn_status bnn_flip_get_grid(bnn_flip* h, int32_t field, void* dst) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.getGrid(field, dst, R.stream);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_pressure(bnn_flip* h, double* dst, uint64_t* n_active) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        if(!s.systemValid) throw bnn::Error(BNN_ERR_STATE, "no pressure solve yet");
        if(s.lastRefined) s.gatherActive<double>(s.x64, dst, n_active, R.stream);
        else s.gatherActive<float>(s.x, dst, n_active, R.stream);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_rhs(bnn_flip* h, double* dst, uint64_t* n_active) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        if(!s.systemValid) throw bnn::Error(BNN_ERR_STATE, "no pressure system yet");
        s.gatherActive<double>(s.rhs64, dst, n_active, R.stream);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_bins(bnn_flip* h, uint32_t* perm, uint32_t* cell_start) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.getBins(perm, cell_start, R.stream);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_stats(bnn_flip* h, bnn_flip_stats* out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    return guard(h, [&](bnn::Scene& s, bnn::Runner&) {
        *out = s.st;
        return BNN_OK;
    });
}

bnn_status bnn_flip_set_grid(bnn_flip* h, int32_t field, const void* src) {
    return guard(h, [&](bnn::Scene& s, bnn::Runner& R) {
        s.setGrid(field, src, R.stream);
        return BNN_OK;
    });
}

bnn_status bnn_flip_synchronize(bnn_flip* h) {
    return guard(h, [&](bnn::Scene&, bnn::Runner& R) {
        R.sync();
        return BNN_OK;
    });
}

bnn_status bnn_flip_footprint(const bnn_flip_params* p, uint64_t n_particles, uint64_t* bytes) {
    if(!p || !bytes) return BNN_ERR_INVALID_ARG;
    try {
        *bytes = bnn::Scene::footprint(*p, n_particles);
        return BNN_OK;
    } catch(const bnn::Error& e) {
        g_createError = e.what();
        return e.code;
    }
}

//--------------------------------------------------------------------------------------------------
// batches
//--------------------------------------------------------------------------------------------------
bnn_status bnn_flip_batch_create(bnn_flip_batch** out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    try {
        auto b = std::make_unique<bnn_flip_batch>();
        b->run = std::make_unique<bnn::Runner>();
        *out   = b.release();
        return BNN_OK;
    } catch(const bnn::Error& e) {
        g_createError = e.what();
        return e.code;
    } catch(const std::exception& e) {
        g_createError = e.what();
        return BNN_ERR_DEVICE;
    }
}

bnn_status bnn_flip_batch_destroy(bnn_flip_batch* b) {
    if(!b) return BNN_ERR_INVALID_ARG;
    delete b;
    return BNN_OK;
}

const char* bnn_flip_batch_last_error(const bnn_flip_batch* b) { return b ? b->err.c_str() : g_createError.c_str(); }

bnn_status bnn_flip_batch_set_timing(bnn_flip_batch* b, int32_t timing) {
    if(!b) return BNN_ERR_INVALID_ARG;
    b->run->timing = timing;
    return BNN_OK;
}

}  // extern "C"

namespace {
template<class F>
bnn_status batch_call(bnn_flip_batch* b, bnn_flip* const* scenes, const float* dts, uint32_t n, bnn_status* status, bool needDt, F&& f) {
    if(!b || (n && !scenes) || (needDt && n && !dts)) return BNN_ERR_INVALID_ARG;
    return guarded(b, b->err, [&] {
        std::vector<bnn::Scene*> sc(n);
        for(uint32_t i = 0; i < n; ++i) {
            if(!scenes[i] || !scenes[i]->s) throw bnn::Error(BNN_ERR_INVALID_ARG, "null scene handle");
            sc[i] = scenes[i]->s.get();
            for(uint32_t j = 0; j < i; ++j)
                if(sc[j] == sc[i]) throw bnn::Error(BNN_ERR_INVALID_ARG, "a scene is listed twice");
            if(needDt) sc[i]->curDt = dts[i];
        }
        if(!n) return BNN_OK;
        b->run->bind(sc);
        f(*b->run);
        b->run->sync();
        bnn_status worst = BNN_OK;
        for(uint32_t i = 0; i < n; ++i) {
            const bnn_status s = solve_status(*sc[i]);
            if(status) status[i] = s;
            if(s != BNN_OK) worst = s;
        }
        return worst;
    });
}
}  // namespace

extern "C" {

bnn_status bnn_flip_batch_step(bnn_flip_batch* b, bnn_flip* const* scenes, const float* dts, uint32_t n, bnn_status* status) {
    return batch_call(b, scenes, dts, n, status, true, [](bnn::Runner& R) { R.step(bnn::Runner::kAll); });
}

bnn_status bnn_flip_batch_advance_frame(bnn_flip_batch* b, bnn_flip* const* scenes, uint32_t n, bnn_status* status) {
    return batch_call(b, scenes, nullptr, n, status, false, [](bnn::Runner& R) { R.advanceFrame(bnn::Runner::kAll); });
}

bnn_status bnn_flip_batch_pressure_projection(bnn_flip_batch* b, bnn_flip* const* scenes, const float* dts, uint32_t n,
                                              bnn_status* status) {
    return batch_call(b, scenes, dts, n, status, true, [](bnn::Runner& R) {
        R.markDesc();
        R.beginStats();
        R.pressureProjection(bnn::Runner::kAll);
        R.endStats(bnn::Runner::kAll);
    });
}

//--------------------------------------------------------------------------------------------------
// serving
//--------------------------------------------------------------------------------------------------
bnn_status bnn_server_create(const bnn_server_config* cfg, bnn_server** out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    try {
        bnn_server_config c{};
        if(cfg) c = *cfg;
        else c.batching = 1;
        auto s = std::make_unique<bnn_server>();
        s->srv = std::make_unique<bnn::Server>(c);
        *out   = s.release();
        return BNN_OK;
    } catch(const bnn::Error& e) {
        g_createError = e.what();
        return e.code;
    } catch(const std::exception& e) {
        g_createError = e.what();
        return BNN_ERR_DEVICE;
    }
}

bnn_status bnn_server_destroy(bnn_server* s) {
    if(!s) return BNN_ERR_INVALID_ARG;
    delete s;
    return BNN_OK;
}

const char* bnn_server_last_error(const bnn_server* s) {
    static thread_local std::string copy;
    if(!s) return g_createError.c_str();
    copy = s->err.empty() ? s->srv->lastError() : s->err;
    return copy.c_str();
}

double bnn_server_now(const bnn_server* s) { return s ? s->srv->now() : 0.0; }

#define BNN_SERVER_GUARD(s, body)                                        \
    if(!(s)) return BNN_ERR_INVALID_ARG;                                 \
    return guarded(s, (s)->err, [&]() -> bnn_status { body; return BNN_OK; })

bnn_status bnn_server_add_scene(bnn_server* s, const bnn_flip_params* p, const bnn_scene_spec* spec, uint64_t seed, uint64_t* id) {
    if(!p || !spec || !id) return BNN_ERR_INVALID_ARG;
    BNN_SERVER_GUARD(s, *id = s->srv->addScene(*p, *spec, seed));
}

bnn_status bnn_server_release_scene(bnn_server* s, uint64_t id) { BNN_SERVER_GUARD(s, s->srv->releaseScene(id)); }

bnn_status bnn_server_submit(bnn_server* s, uint64_t scene, int32_t kind, float dt, uint64_t* id) {
    if(!id) return BNN_ERR_INVALID_ARG;
    BNN_SERVER_GUARD(s, *id = s->srv->submit(scene, kind, dt));
}

bnn_status bnn_server_wait(bnn_server* s, uint64_t id, double timeout, bnn_request_result* out) {
    BNN_SERVER_GUARD(s, if(!s->srv->wait(id, timeout, out)) return BNN_ERR_STATE);
}

bnn_status bnn_server_drain(bnn_server* s, double timeout) { BNN_SERVER_GUARD(s, if(!s->srv->drain(timeout)) return BNN_ERR_STATE); }

bnn_status bnn_server_get_particles(bnn_server* s, uint64_t scene, float* pos, float* vel, uint64_t* n) {
    BNN_SERVER_GUARD(s, s->srv->getParticles(scene, pos, vel, n));
}

bnn_status bnn_server_get_stats(bnn_server* s, bnn_server_stats* out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    BNN_SERVER_GUARD(s, *out = s->srv->stats());
}

bnn_status bnn_server_reset_stats(bnn_server* s) { BNN_SERVER_GUARD(s, s->srv->resetStats()); }

}  // extern "C"
