// flip_host.hpp -- host utilities of the FLIP library: errors, device/pinned buffers, grid sizing
// and the seeded scene generator.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../../include/bnn_flip.h"

namespace bnn {

struct Error : std::runtime_error {
    bnn_status code;
    Error(bnn_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_OK(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if(e_ != hipSuccess)                                                                        \
            throw ::bnn::Error(e_ == hipErrorOutOfMemory ? BNN_ERR_OUT_OF_MEMORY : BNN_ERR_DEVICE, \
                               std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while(0)

static inline unsigned blocks(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// device buffer; reserve() grows it (contents are not preserved across growth). The caller makes
// sure no queued kernel still uses the old allocation when it grows.
struct DevBuf {
    void*  p     = nullptr;
    size_t bytes = 0;
    DevBuf()     = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if(p) (void)hipFree(p);
        p     = nullptr;
        bytes = 0;
    }
    // returns the number of bytes newly allocated
    size_t reserve(size_t b) {
        if(b <= bytes && p) return 0;
        release();
        if(b == 0) b = 256;
        HIP_OK(hipMalloc(&p, b));
        bytes = b;
        return b;
    }
    template<class T> T* as() const { return static_cast<T*>(p); }
};

struct PinnedBuf {
    void*  p     = nullptr;
    size_t bytes = 0;
    PinnedBuf()  = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if(p) (void)hipHostFree(p);
    }
    void reserve(size_t b) {
        if(b <= bytes && p) return;
        if(p) (void)hipHostFree(p);
        p = nullptr;
        HIP_OK(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
    }
    template<class T> T* as() const { return static_cast<T*>(p); }
};

//--------------------------------------------------------------------------------------------------
// Grid sizing: ParticleSolverData.cpp:247-253 (domain grown by NExpandCells*dx) + Grid.cpp:38-55
// (n = ceil(extent / dx))
//--------------------------------------------------------------------------------------------------
struct GridDims {
    float bmin[3], dx;
    int   n[3];
};
static inline GridDims grid_dims(const bnn_flip_params& p) {
    GridDims g;
    g.dx = p.cell_size;
    for(int d = 0; d < 3; ++d) {
        float ex  = g.dx * float(p.n_expand_cells);
        g.bmin[d] = p.box_min[d] - ex;
        float hi  = p.box_max[d] + ex;
        g.n[d]    = (int)std::ceil((hi - g.bmin[d]) / g.dx);
    }
    return g;
}

//--------------------------------------------------------------------------------------------------
// Seeded scene generator (host): one particle per sub-cell of the grid's sub-cell lattice whose
// centre lies inside an object, jittered inside the sub-cell; see bnn_scene_spec.
//--------------------------------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline double unit01(uint64_t seed, uint64_t id, uint32_t stream) {
    uint64_t h = mix64(seed ^ mix64(id * 0x2545F4914F6CDD1Dull + stream));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

static inline uint64_t generate_scene(const bnn_flip_params& prm, const bnn_scene_spec& sp, uint64_t seed, float* pos, float* vel,
                                      uint64_t cap) {
    const GridDims G    = grid_dims(prm);
    const int      ppa  = sp.ppa > 0 ? sp.ppa : 2;
    const double   h    = (double)G.dx / ppa;
    const int64_t  N[3] = {(int64_t)G.n[0] * ppa, (int64_t)G.n[1] * ppa, (int64_t)G.n[2] * ppa};
    const int      nb = std::min(std::max(sp.n_boxes, 0), 8), ns = std::min(std::max(sp.n_spheres, 0), 8);
    auto           in_obj = [&](int o, const double c[3]) {
        if(o < nb) {
            for(int d = 0; d < 3; ++d)
                if(c[d] < (double)sp.box_min[o][d] || c[d] > (double)sp.box_max[o][d]) return false;
            return true;
        }
        const int s  = o - nb;
        double    r2 = 0;
        for(int d = 0; d < 3; ++d) {
            double e = c[d] - (double)sp.sphere_center[s][d];
            r2 += e * e;
        }
        return r2 <= (double)sp.sphere_radius[s] * (double)sp.sphere_radius[s];
    };
    uint64_t count = 0;
    for(int o = 0; o < nb + ns; ++o) {
        int64_t lo[3], hi[3];
        for(int d = 0; d < 3; ++d) {
            double a, b;
            if(o < nb) {
                a = sp.box_min[o][d];
                b = sp.box_max[o][d];
            } else {
                a = (double)sp.sphere_center[o - nb][d] - sp.sphere_radius[o - nb];
                b = (double)sp.sphere_center[o - nb][d] + sp.sphere_radius[o - nb];
            }
            lo[d] = std::max<int64_t>(0, (int64_t)std::floor((a - G.bmin[d]) / h - 0.5) - 1);
            hi[d] = std::min<int64_t>(N[d] - 1, (int64_t)std::ceil((b - G.bmin[d]) / h - 0.5) + 1);
        }
        for(int64_t mz = lo[2]; mz <= hi[2]; ++mz)
            for(int64_t my = lo[1]; my <= hi[1]; ++my)
                for(int64_t mx = lo[0]; mx <= hi[0]; ++mx) {
                    const double c[3] = {(double)G.bmin[0] + ((double)mx + 0.5) * h, (double)G.bmin[1] + ((double)my + 0.5) * h,
                                         (double)G.bmin[2] + ((double)mz + 0.5) * h};
                    if(!in_obj(o, c)) continue;
                    bool taken = false;
                    for(int q = 0; q < o && !taken; ++q) taken = in_obj(q, c);
                    if(taken) continue;
                    if(pos && count < cap) {
                        const uint64_t gid = ((uint64_t)mz * (uint64_t)N[1] + (uint64_t)my) * (uint64_t)N[0] + (uint64_t)mx;
                        for(int d = 0; d < 3; ++d) {
                            const double u     = unit01(seed, gid, (uint32_t)d);
                            pos[count * 3 + d] = (float)(c[d] + (u - 0.5) * (double)sp.jitter * h);
                            double v0          = o < nb ? sp.box_vel[o][d] : sp.sphere_vel[o - nb][d];
                            if(sp.vel_sigma > 0) {
                                const double u1 = unit01(seed, gid, 3u + 2u * d), u2 = unit01(seed, gid, 4u + 2u * d);
                                v0 += (double)sp.vel_sigma * (std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2));
                            }
                            vel[count * 3 + d] = (float)v0;
                        }
                    }
                    ++count;
                }
    }
    return count;
}

}  // namespace bnn
