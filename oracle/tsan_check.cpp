// ThreadSanitizer check of the oracle's threaded path (test infrastructure, like the oracle itself).
// Built with -fsanitize=thread by `make tsan` in this folder and run by tests/test_tsan.py: a small
// dam break with a falling sphere and walls, stepped by the FP32 oracle on 4 threads, so the pool's
// chunked loops, the CAS float adds of the particle scatter (AtomicOperations.h:31-49), the CAS min of
// the fluid SDF and the reductions all run concurrently. Exit status 0 with no report is the pass.
#include "flip_oracle.cpp"

#include <cstdio>
#include <vector>

int main() {
    orc_params p{};
    for(int d = 0; d < 3; ++d) {
        p.box_min[d] = 0.f;
        p.box_max[d] = 1.f;
    }
    p.cell_size               = 1.f / 20.f;
    p.ratio_cell_size_pradius = 2.f;
    p.n_expand_cells          = 2;
    p.min_timestep = p.max_timestep = p.frame_duration = 1.f / 600.f;
    p.cfl_factor                                        = 1.f;
    p.cg_relative_tolerance                             = 1e-6f;
    p.max_cg_iteration                                  = 5000;
    p.pic_flip_ratio                                    = 0.97f;
    p.advection_steps                                   = 1;
    p.apply_gravity                                     = 1;
    p.boundary_reflection_multiplier                    = 1.f;
    p.boundary_sdf_at_init                              = 1;
    orc_scene_spec sp{};
    sp.n_boxes = 1;
    sp.box_max[0][0] = 0.4f;
    sp.box_max[0][1] = 0.6f;
    sp.box_max[0][2] = 1.f;
    sp.n_spheres        = 1;
    sp.sphere_center[0][0] = 0.7f;
    sp.sphere_center[0][1] = 0.75f;
    sp.sphere_center[0][2] = 0.5f;
    sp.sphere_radius[0]    = 0.12f;
    sp.sphere_vel[0][1]    = -2.f;
    sp.ppa                 = 2;
    sp.jitter              = 1.f;
    const uint64_t     n = orc_generate_scene(&p, &sp, 7, nullptr, nullptr);
    std::vector<float> pos(3 * n), vel(3 * n);
    orc_generate_scene(&p, &sp, 7, pos.data(), vel.data());
    void* h = orc_create(&p, 32, 4);
    orc_set_particles(h, pos.data(), vel.data(), n);
    int failed = 0;
    for(int s = 0; s < 3; ++s) failed |= orc_step(h, 1.f / 600.f);
    uint32_t iters  = 0;
    double   res    = 0;
    int32_t  ok     = 0;
    orc_pcg_stats(h, &iters, &res, &ok);
    orc_destroy(h);
    std::printf("tsan_check: %llu particles, 3 substeps on 4 threads, last solve %u iterations, success %d\n", (unsigned long long)n,
                iters, ok);
    return failed || !ok ? 1 : 0;
}
