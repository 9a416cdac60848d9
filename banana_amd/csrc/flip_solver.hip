// flip_solver.hip -- host side of the MI355X FLIP substep and its C ABI (include/bnn_flip.h).
//
// bnn::FlipSolver owns one scene's device state and queues the stage kernels of
// FLIP_Solver<3,float>::advanceVelocity (FLIP_Solver.cpp:42-53) on its own HIP stream.
#include "flip_kernels.hpp"
#include "flip_batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <sys/stat.h>
#include <string>
#include <vector>

#include "../../include/bnn_flip.h"
#include "flip_boundary.hpp"
#include "flip_io.hpp"
#include "scene_json.hpp"

namespace bnn {

struct Error : std::runtime_error {
    bnn_status code;
    Error(bnn_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_OK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if(e_ != hipSuccess)                                                                             \
            throw ::bnn::Error(e_ == hipErrorOutOfMemory ? BNN_ERR_OUT_OF_MEMORY : BNN_ERR_DEVICE,      \
                               std::string(#expr) + ": " + hipGetErrorString(e_));                       \
    } while(0)

static inline unsigned blocks(size_t n, unsigned bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

// scratch of the z-sort: 256-bin digit histograms and their scanned offsets per 4096-key tile, the
// AND / OR key accumulators, and the scan's per-tile partials
struct ZsortScratch {
    uint32_t*           hist;  // 2 * nh words: histograms, then offsets
    unsigned long long* acc;   // 2 words
    uint32_t*           part;  // partWords words
};
struct ZsortLayout {
    size_t nh, partWords;
    size_t histOff, accOff, partOff, bytes;  // offsets into one contiguous block (256-B aligned)
};
static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
static inline ZsortLayout zsort_layout(uint64_t np) {
    ZsortLayout L{};
    const size_t ntiles = (size_t)((np + kRsTile - 1) / kRsTile);
    L.nh                = 256 * ntiles;
    L.partWords         = blocks(L.nh, kScanTile) + 1;
    L.accOff            = 0;
    L.histOff           = 256;
    L.partOff           = L.histOff + align256(L.nh * 8);
    L.bytes             = L.partOff + align256(L.partWords * 4);
    return L;
}

// the recorder of the batch this thread is running, if any: a buffer freed meanwhile first has the
// launches recorded so far issued and finished (they may use it)
static thread_local Recorder* tl_batch = nullptr;

// device buffer, grows on demand (contents are not preserved across growth)
struct DevBuf {
    void*  p     = nullptr;
    size_t bytes = 0;
    DevBuf()     = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if(p && tl_batch) tl_batch->flushAndSync();
        if(p) (void)hipFree(p);
        p     = nullptr;
        bytes = 0;
    }
    void reserve(size_t b) {
        if(b <= bytes && p) return;
        release();
        if(b == 0) b = 256;
        HIP_OK(hipMalloc(&p, b));
        bytes = b;
    }
    template<class T> T* as() const { return static_cast<T*>(p); }
    void swapWith(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
    }
};

struct ParticleBufs {
    DevBuf px, py, pz, vx, vy, vz, gp, gv, id;
    DevBuf c[9];  // APIC / AFLIP affine matrices (allocated only for those schemes)
    void   reserve(size_t n, bool affine) {
        for(DevBuf* b : {&px, &py, &pz, &vx, &vy, &vz}) b->reserve(n * 4);
        gp.reserve(n * 16);
        gv.reserve(n * 8);
        id.reserve(n * 4);
        if(affine)
            for(DevBuf& b : c) b.reserve(n * 4);
    }
    Affine affine() const {
        Affine a;
        for(int k = 0; k < 9; ++k) a.c[k] = c[k].as<float>();
        return a;
    }
    Particles view() const {
        return Particles{px.as<float>(), py.as<float>(), pz.as<float>(), vx.as<float>(), vy.as<float>(),
                         vz.as<float>(), gp.as<float4>(), gv.as<float2>(), id.as<uint32_t>()};
    }
};

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

struct GridDims {
    float bmin[3], dx;
    int   n[3];
};
// ParticleSolverData.cpp:247-253 (domain grown by NExpandCells*dx) + Grid.cpp:38-55 (n = ceil)
static GridDims grid_dims(const bnn_flip_params& p) {
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

static uint64_t generate_scene(const bnn_flip_params& prm, const bnn_scene_spec& sp, uint64_t seed, float* pos, float* vel,
                               uint64_t cap, uint16_t* obj = nullptr) {
    const GridDims G   = grid_dims(prm);
    const int      ppa = sp.ppa > 0 ? sp.ppa : 2;
    const double   h   = (double)G.dx / ppa;
    const int64_t  N[3] = {(int64_t)G.n[0] * ppa, (int64_t)G.n[1] * ppa, (int64_t)G.n[2] * ppa};
    const int      nb = std::min(std::max(sp.n_boxes, 0), 8), ns = std::min(std::max(sp.n_spheres, 0), 8);
    auto in_obj = [&](int o, const double c[3]) {
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
                            const double u = unit01(seed, gid, (uint32_t)d);
                            pos[count * 3 + d] = (float)(c[d] + (u - 0.5) * (double)sp.jitter * h);
                            double v0 = o < nb ? sp.box_vel[o][d] : sp.sphere_vel[o - nb][d];
                            if(sp.vel_sigma > 0) {
                                const double u1 = unit01(seed, gid, 3u + 2u * d), u2 = unit01(seed, gid, 4u + 2u * d);
                                v0 += (double)sp.vel_sigma * (std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2));
                            }
                            vel[count * 3 + d] = (float)v0;
                        }
                    }
                    if(obj && pos && count < cap) obj[count] = (uint16_t)o;  // generator index (ParticleSolverData.h objectIndex)
                    ++count;
                }
    }
    return count;
}

//--------------------------------------------------------------------------------------------------
class FlipSolver {
public:
    bnn_flip_params prm;
    GridGeom        g{};
    BoxDev          dom{};
    // the scene's additional boundary objects and the device set for their current transforms
    std::vector<BoundaryObjectHost> bobjs;
    BndSet                          bnd{};
    float           grav[3] = {0, -9.81f, 0};
    std::string     lastError;
    hipStream_t     stream = nullptr;

    // In a batch (flip_batch.hpp) launches and readbacks go to `rec` as lane `lane`, and a sync point
    // suspends the host coroutine, leaving its handle in `pending`.
    Recorder*               rec  = nullptr;
    uint32_t                lane = 0;
    std::coroutine_handle<> pending;

    template<class K> void launch(const K& k, uint32_t nblk) {
        if(nblk == 0) return;
        if(rec) rec->add(lane, k, nblk);
        else hipLaunchKernelGGL(k_one<K>, dim3(nblk), dim3(K::kThreads), 0, stream, k);
    }
    void fill32(void* p, uint32_t v, size_t bytes) {
        if(bytes) launch(k_fill32{static_cast<uint32_t*>(p), bytes / 4, v}, fill_blocks(bytes / 4));
    }
    void copy32(void* dst, const void* src, size_t bytes) {
        if(bytes) launch(k_copy32{static_cast<uint32_t*>(dst), static_cast<const uint32_t*>(src), bytes / 4}, fill_blocks(bytes / 4));
    }
    void readback(void* host, const void* dev, size_t bytes) {
        if(rec) rec->readback(lane, host, dev, bytes);
        else HIP_OK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream));
    }
    SyncAwait syncPoint() {
        if(rec) return SyncAwait{&pending};
        HIP_OK(hipStreamSynchronize(stream));
        return SyncAwait{nullptr};
    }
    // a phase boundary: alone nothing; in a batch the scene waits here until every scene of the
    // batch waits at a barrier or has finished, so that the next phase's launches of all scenes
    // merge (scenes reach the end of a solve after different numbers of iterations)
    bool      atBarrier = false;
    SyncAwait barrier() {
        if(!rec) return SyncAwait{nullptr};
        atBarrier = true;
        return SyncAwait{&pending};
    }

    // particles
    uint64_t     np = 0, cap = 0;
    ParticleBufs bufA, bufB;
    ParticleBufs *cur = &bufA, *alt = &bufB;
    DevBuf       key, perm;
    DevBuf       zsHist, zsAcc;  // z-sort digit histograms / offsets, key AND-OR
    bool         binned = false;     // cell_start / gp valid for the current particle positions
    bool         packedVel = false;  // the velocities in gp.w / gv equal the velocity arrays

    // grid
    size_t nU = 0, nV = 0, nW = 0, nC = 0, nN = 0;
    DevBuf u, v, w, ut, vt, wt, uo, vo, wo, wu, wv, ww;
    DevBuf vu, vv, vw, eu, ew;
    DevBuf phi, bsdf, cellStart, cellCount, scanPart, counters;
    uint8_t extGen = 2;  // the valid-flag value of faces the last extrapolation filled
    DevBuf occ;  // per 8^3 cell brick: holds a particle (binning), so P2G can skip empty neighbourhoods
    bool     weightsDirty    = true;   // boundary SDF changed since the weights were computed
    bool     weightsComputed = false;  // computeCellWeights has run at least once
    bool     bsdfNonzero     = false;  // the boundary SDF has been computed or set
    uint32_t solidFaces      = 0;      // faces with weight < FLT_MIN in the current weights

    // bricks + pressure system
    uint32_t nb = 0, nSlots = 0, slotCap = 0, nFluid = 0;
    DevBuf   bflag, bscan, brickSlot, slotBrick, nbrTab;
    DevBuf   x, r, z, s0, s1, q, rhs32, diag32, cx32, cy32, cz32, invd;
    DevBuf   rhs64, diag64, cx64, cy64, cz64, x64;
    DevBuf   x64prev, prevBrickSlot;  // the last refined solve's pressure and brick -> slot table
    bool     prevPressure = false;   // they hold a converged pressure of this grid
    bool     warmNext     = false;   // this solve starts from them (x64 filled by k_warm_start)
    DevBuf   pa, pb, pz, pst, ost;  // PCG partials (per block: pa, pb; per brick: pz), device states
    DevBuf   icInv, icLx, icLy, icLz;  // IC(0) factor (BNN_PRECOND_IC0 only)
    PcgState*   hostSt  = nullptr;  // pinned
    OuterState* hostOst = nullptr;
    uint32_t*   hostU32 = nullptr;
    int         device  = 0;  // the device current at creation (owns the stream and every buffer)
    // ParticleSimulationData::objectIndex / nObjects: the generator of each particle, in user order
    // (empty: every particle 0)
    std::vector<uint16_t> objIndex;
    uint32_t              nObjects = 1;
    std::vector<uint16_t> objScratch;  // same size as objIndex: the z-sort's reordering (no allocation there)
    std::vector<uint32_t> rankHost;
    bool        systemValid = false;
    bool        lastRefined = false;
    // launch shapes that leave results unchanged (set by the batch autotuner): P2G column segment
    // length, and the PCG iterations queued before the first stop-flag poll and at most per poll
    int      p2gSeg        = 8;
    uint32_t pcgChunkFirst = 4, pcgChunkMax = 64;

    // stats
    bnn_flip_stats st{};
    double         stageAcc[BNN_NUM_STAGES] = {};
    // frame bookkeeping (ParticleSolverData.h:43-50)
    float    frameTime     = 0;
    uint32_t finishedFrame = 0;

    explicit FlipSolver(const bnn_flip_params& p) : prm(p) {
        validate();
        const GridDims G = grid_dims(prm);
        g.nx = G.n[0];
        g.ny = G.n[1];
        g.nz = G.n[2];
        g.bminx = G.bmin[0];
        g.bminy = G.bmin[1];
        g.bminz = G.bmin[2];
        g.dx        = G.dx;
        g.halfdx    = 0.5f * G.dx;
        g.sdfRadius = G.dx * (float)(1.01 * std::sqrt(3) / 2.0);  // PIC_Data.cpp:43
        g.nbx = (g.nx + 7) / 8;
        g.nby = (g.ny + 7) / 8;
        g.nbz = (g.nz + 7) / 8;
        for(int d = 0; d < 3; ++d) {
            dom.lo[d] = prm.box_min[d];
            dom.hi[d] = prm.box_max[d];
        }
        bnd = make_bnd_set(dom.lo, dom.hi, bobjs);
        // gravity (ParticleSolverData.cpp:256-271)
        if(prm.gravity_type == 1) {
            const float* gd = prm.gravity_direction;
            float        l2 = gd[0] * gd[0] + gd[1] * gd[1] + gd[2] * gd[2];
            if(l2 < 1.1920929e-07f) {
                prm.gravity_type = 0;
            } else {
                float is = 1.0f / std::sqrt(l2);
                for(int d = 0; d < 3; ++d) grav[d] = 9.81f * (gd[d] * is);
            }
        }
        if(prm.gravity_type == 0) {
            grav[0] = 0;
            grav[1] = -9.81f;
            grav[2] = 0;
        }
        HIP_OK(hipGetDevice(&device));
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        allocGrid();
        HIP_OK(hipHostMalloc((void**)&hostSt, sizeof(PcgState), hipHostMallocDefault));
        HIP_OK(hipHostMalloc((void**)&hostOst, sizeof(OuterState), hipHostMallocDefault));
        HIP_OK(hipHostMalloc((void**)&hostU32, 64, hipHostMallocDefault));
        if(prm.boundary_sdf_at_init) computeBoundarySDF();
        HIP_OK(hipStreamSynchronize(stream));
    }

    ~FlipSolver() {
        if(stream) (void)hipStreamSynchronize(stream);
        for(hipEvent_t e : evPool) (void)hipEventDestroy(e);
        if(hostSt) (void)hipHostFree(hostSt);
        if(hostOst) (void)hipHostFree(hostOst);
        if(hostU32) (void)hipHostFree(hostU32);
        if(stream) (void)hipStreamDestroy(stream);
    }

    void validate() {
        if(!(prm.cell_size > 0)) throw Error(BNN_ERR_INVALID_ARG, "CellSize must be > 0");
        for(int d = 0; d < 3; ++d)
            if(!(prm.box_max[d] > prm.box_min[d])) throw Error(BNN_ERR_INVALID_ARG, "SimulationDomainBox: BoxMax must exceed BoxMin");
        if(prm.advection_steps < 1) throw Error(BNN_ERR_INVALID_ARG, "AdvectionSteps must be >= 1");
        if(prm.enable_sort_particle && prm.sort_frequency < 1) throw Error(BNN_ERR_INVALID_ARG, "SortFrequency must be >= 1");
        if(prm.gravity_type < 0 || prm.gravity_type > 3) throw Error(BNN_ERR_INVALID_ARG, "GravityType must be 0..3");
        if(prm.transfer_scheme < BNN_TRANSFER_FLIP || prm.transfer_scheme > BNN_TRANSFER_AFLIP)
            throw Error(BNN_ERR_INVALID_ARG, "unknown transfer scheme");
        if(!(prm.refine_inner_tolerance > 0.f && prm.refine_inner_tolerance < 1.f))
            throw Error(BNN_ERR_INVALID_ARG, "refine_inner_tolerance must be in (0, 1)");
        if(prm.preconditioner != BNN_PRECOND_JACOBI && prm.preconditioner != BNN_PRECOND_MG && prm.preconditioner != BNN_PRECOND_IC0)
            throw Error(BNN_ERR_INVALID_ARG, "unknown preconditioner");
        if(prm.preconditioner == BNN_PRECOND_MG && (prm.mg_sweeps < 1 || prm.mg_sweeps > 8 || prm.mg_coarsest_sweeps < 1 ||
                                                    prm.mg_coarsest_sweeps > 512 || !(prm.mg_omega > 0.f) || prm.mg_omega >= 2.f))
            throw Error(BNN_ERR_INVALID_ARG, "multigrid settings out of range");
        const GridDims G = grid_dims(prm);
        const double   C = (double)G.n[0] * G.n[1] * G.n[2];
        if(C > 1.0e9 || G.n[0] < 3 || G.n[1] < 3 || G.n[2] < 3) throw Error(BNN_ERR_INVALID_ARG, "grid size out of range");
    }

    void allocGrid() {
        nU = (size_t)(g.nx + 1) * g.ny * g.nz;
        nV = (size_t)g.nx * (g.ny + 1) * g.nz;
        nW = (size_t)g.nx * g.ny * (g.nz + 1);
        nC = (size_t)g.nx * g.ny * g.nz;
        nN = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
        nb = (uint32_t)(g.nbx * g.nby * g.nbz);
        for(auto* b : {&u, &ut, &uo, &wu}) b->reserve(nU * 4);
        for(auto* b : {&v, &vt, &vo, &wv}) b->reserve(nV * 4);
        for(auto* b : {&w, &wt, &wo, &ww}) b->reserve(nW * 4);
        // face-brick masks of the extrapolation (valid / near a valid face), per component in turn
        vu.reserve(nU);
        vv.reserve(nV);
        vw.reserve(nW);
        eu.reserve(nU);
        ev_.reserve(std::max(nV, (size_t)1));
        ew.reserve(nW);
        phi.reserve(nC * 4);
        bsdf.reserve(nN * 4);
        cellStart.reserve((nC + 1) * 4);
        cellCount.reserve(nC * 4);
        scanPart.reserve((blocks(std::max(nC, (size_t)nb), kScanTile) + 1) * 4);
        counters.reserve(256);
        bflag.reserve(nb * 4);
        occ.reserve((size_t)(nb + 3) / 4 * 4);
        bscan.reserve(nb * 4);
        brickSlot.reserve(nb * 4);
        pst.reserve(sizeof(PcgState));
        ost.reserve(sizeof(OuterState));
        // zero everything the reference zero-fills at resize (PIC_Data.cpp:101-117)
        for(auto* b : {&u, &v, &w, &ut, &vt, &wt, &uo, &vo, &wo, &wu, &wv, &ww, &vu, &vv, &vw, &eu, &ev_, &ew, &phi, &bsdf,
                       &cellStart, &cellCount, &counters, &pst, &ost})
            HIP_OK(hipMemsetAsync(b->p, 0, b->bytes, stream));
        weightsDirty = true;
    }
    DevBuf ev_;  // extrapolation flags of v (named apart from the event table)

    void ensureParticles(uint64_t n) {
        if(n <= cap) return;
        const uint64_t c = n + n / 8 + 1024;
        bufA.reserve(c, affineScheme());
        bufB.reserve(c, affineScheme());
        key.reserve(c * 4);
        perm.reserve(c * 4);
        cap = c;
    }

    void ensureSlots(uint32_t n) {
        if(n <= slotCap && slotCap) return;
        const uint32_t c  = std::max<uint32_t>(n + n / 4 + 16, 64);
        const size_t   nv = (size_t)c * 512;
        for(auto* b : {&x, &r, &z, &s0, &s1, &q, &rhs32, &diag32, &cx32, &cy32, &cz32, &invd}) b->reserve(nv * 4);
        slotBrick.reserve(c * 4);
        nbrTab.reserve((size_t)c * 6 * 4);
        pa.reserve((size_t)kMaxRedBlocks * 8);
        pb.reserve((size_t)std::max<uint32_t>(kMaxRedBlocks, c) * 8);  // per block, or per brick (k_pcg_update_pre)
        pz.reserve((size_t)c * 8);
        if(prm.preconditioner == BNN_PRECOND_IC0)
            for(auto* b : {&icInv, &icLx, &icLy, &icLz}) b->reserve(nv * 4);
        if(prm.pressure_mode == BNN_PRESSURE_REFINED)
            for(auto* b : {&rhs64, &diag64, &cx64, &cy64, &cz64, &x64}) b->reserve(nv * 8);
        else
            rhs64.reserve(nv * 8);
        slotCap = c;
    }

    // ---------------------------------------------------------------------------------------------
    void setParticles(const float* pos, const float* vel, uint64_t n) {
        if(n > 0xFFFFFFF0ull) throw Error(BNN_ERR_INVALID_ARG, "too many particles for 32-bit particle ids");
        if(n && (!pos || !vel)) throw Error(BNN_ERR_INVALID_ARG, "null particle array");
        for(uint64_t i = 0; i < 3 * n; ++i)
            if(!std::isfinite(pos[i]) || !std::isfinite(vel[i]))
                throw Error(BNN_ERR_INVALID_ARG, "particle " + std::to_string(i / 3) + " has a non-finite position or velocity");
        ensureParticles(std::max<uint64_t>(n, 1));
        std::vector<float> tmp(n);
        std::vector<uint32_t> ids(n);
        ParticleBufs& P = *cur;
        DevBuf* dst[6] = {&P.px, &P.py, &P.pz, &P.vx, &P.vy, &P.vz};
        for(int a = 0; a < 6; ++a) {
            const float* src = a < 3 ? pos : vel;
            for(uint64_t i = 0; i < n; ++i) tmp[i] = src[3 * i + (a % 3)];
            if(n) HIP_OK(hipMemcpyAsync(dst[a]->p, tmp.data(), n * 4, hipMemcpyHostToDevice, stream));
            HIP_OK(hipStreamSynchronize(stream));
        }
        for(uint64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
        if(n) HIP_OK(hipMemcpyAsync(P.id.p, ids.data(), n * 4, hipMemcpyHostToDevice, stream));
        if(affineScheme() && n)  // new particles start with C = 0 (APIC_Data: the matrices are zero-filled)
            for(DevBuf& b : P.c) HIP_OK(hipMemsetAsync(b.p, 0, n * 4, stream));
        HIP_OK(hipMemsetAsync(cellCount.p, 0, cellCount.bytes, stream));
        HIP_OK(hipStreamSynchronize(stream));
        np     = n;
        binned = false;
        objIndex.clear();
        objScratch.clear();
        rankHost.clear();
        nObjects = 1;
    }
    void setObjectIndices(std::vector<uint16_t>&& obj, uint32_t nObj) {
        if(obj.size() != np) throw Error(BNN_ERR_INVALID_ARG, "object_index has the wrong particle count");
        objIndex.swap(obj);
        objScratch.assign(np, 0);
        rankHost.assign(np, 0);
        nObjects = nObj;
    }

    bool affineScheme() const { return prm.transfer_scheme != BNN_TRANSFER_FLIP; }

    // the affine matrices in the caller's order, 9 floats per particle (row d = C_d)
    void getAffine(float* c) {
        if(!affineScheme()) throw Error(BNN_ERR_STATE, "the transfer scheme has no affine matrices (FLIP)");
        if(!np) return;
        std::vector<uint32_t> ids(np);
        std::vector<float>    a(np);
        HIP_OK(hipMemcpyAsync(ids.data(), cur->id.p, np * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for(int k = 0; k < 9; ++k) {
            HIP_OK(hipMemcpyAsync(a.data(), cur->c[k].p, np * 4, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            for(uint64_t i = 0; i < np; ++i) c[9 * (uint64_t)ids[i] + k] = a[i];
        }
    }
    void setAffine(const float* c) {
        if(!affineScheme()) throw Error(BNN_ERR_STATE, "the transfer scheme has no affine matrices (FLIP)");
        if(!np) return;
        std::vector<uint32_t> ids(np);
        std::vector<float>    a(np);
        HIP_OK(hipMemcpyAsync(ids.data(), cur->id.p, np * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for(int k = 0; k < 9; ++k) {
            for(uint64_t i = 0; i < np; ++i) a[i] = c[9 * (uint64_t)ids[i] + k];
            HIP_OK(hipMemcpyAsync(cur->c[k].p, a.data(), np * 4, hipMemcpyHostToDevice, stream));
            HIP_OK(hipStreamSynchronize(stream));
        }
    }

    void getParticles(float* pos, float* vel) {
        if(!np) return;
        std::vector<float>    a(np);
        std::vector<uint32_t> ids(np);
        HIP_OK(hipMemcpyAsync(ids.data(), cur->id.p, np * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        const DevBuf* src[6] = {&cur->px, &cur->py, &cur->pz, &cur->vx, &cur->vy, &cur->vz};
        for(int k = 0; k < 6; ++k) {
            float* out = k < 3 ? pos : vel;
            if(!out) continue;
            HIP_OK(hipMemcpyAsync(a.data(), src[k]->p, np * 4, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            for(uint64_t i = 0; i < np; ++i) out[3 * (uint64_t)ids[i] + (k % 3)] = a[i];
        }
    }

    // ---------------------------------------------------------------------------------------------
    // timing: HIP event pairs recorded on the solver stream, read once at the end of a step.
    // tag < BNN_NUM_STAGES: a stage; tag >= 100: kernel timer (tag - 100)
    struct TimerRec {
        int    tag;
        size_t b, e;
    };
    std::vector<hipEvent_t> evPool;
    size_t                  evUsed = 0;
    std::vector<TimerRec>   timerRecs;
    double                  ktAcc[BNN_NUM_KERNEL_TIMERS] = {};
    uint64_t                ktCnt[BNN_NUM_KERNEL_TIMERS] = {};
    size_t recordEvent() {
        if(evUsed == evPool.size()) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            evPool.push_back(e);
        }
        HIP_OK(hipEventRecord(evPool[evUsed], stream));
        return evUsed++;
    }
    void openTimer(int tag) { timerRecs.push_back(TimerRec{tag, recordEvent(), (size_t)-1}); }
    void closeTimer(int tag) {
        for(size_t i = timerRecs.size(); i-- > 0;)
            if(timerRecs[i].tag == tag && timerRecs[i].e == (size_t)-1) {
                timerRecs[i].e = recordEvent();
                return;
            }
    }
    void stageBegin(int s) {
        if(prm.timing) openTimer(s);
    }
    void stageEnd(int s) {
        if(prm.timing) closeTimer(s);
    }
    void ktBegin(int k) {
        if(prm.timing >= 2) openTimer(100 + k);
    }
    void ktEnd(int k) {
        if(prm.timing >= 2) closeTimer(100 + k);
    }
    void collectTimers() {
        if(timerRecs.empty()) return;
        HIP_OK(hipStreamSynchronize(stream));
        for(const auto& t : timerRecs) {
            if(t.e == (size_t)-1) continue;
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, evPool[t.b], evPool[t.e]));
            if(t.tag >= 100) {
                ktAcc[t.tag - 100] += ms;
                ktCnt[t.tag - 100] += 1;
            } else {
                stageAcc[t.tag] += ms;
            }
        }
        timerRecs.clear();
        evUsed = 0;
    }

    // ---------------------------------------------------------------------------------------------
    // part: ceil(n / kScanTile) + 1 words of partials (the solver's own buffer when null)
    void scan(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, uint32_t* part = nullptr) {
        const unsigned nt = blocks(n, kScanTile);
        if(!part) part = scanPart.as<uint32_t>();
        launch(k_scan_reduce{in, n, part}, nt);
        launch(k_scan_partials{part, nt, total}, 1);
        launch(k_scan_down{in, n, (const uint32_t*)part, out}, nt);
    }

    void computeBoundarySDF() {
        launch(k_boundary_sdf{bsdf.as<float>(), g, bnd}, blocks(nN));
        weightsDirty = true;
        bsdfNonzero  = true;
    }

    // PIC_Solver.cpp:353-373
    void moveParticles(float dt) {
        if(!np) return;
        VelGrid G{u.as<float>(), v.as<float>(), w.as<float>()};
        ktBegin(BNN_KT_MOVE);
        launch(k_move{cur->view(), (uint32_t)np, G, g, bnd, dt,
                           prm.advection_steps, prm.reflect_velocity_at_boundary, prm.boundary_reflection_multiplier}, blocks(np));
        ktEnd(BNN_KT_MOVE);
        binned = false;
    }

    uint64_t correctSeed() const { return mix64(prm.random_seed ^ mix64(st.steps + 1)); }

    // PIC_Solver::correctParticlePositions (PIC_Solver.cpp:377-485): bins the current positions,
    // writes the pushed positions to the spare buffer and copies them back
    void correctParticlePositions(float dt) {
        if(!np) return;
        requireBinned();
        const float   pr     = prm.cell_size / prm.ratio_cell_size_pradius;  // ParticleSolverData.cpp:232
        const float   radius = 2.0f * pr / (float)std::sqrt(3);
        CorrectParams c;
        c.radius2    = radius * radius;
        c.threshold  = 0.01f * radius;
        c.threshold2 = c.threshold * c.threshold;
        c.push       = c.threshold / dt;
        c.K          = radius * (dt / (float)prm.advection_steps) * prm.repulsive_force_stiffness;
        c.advSteps   = prm.advection_steps;
        c.seed       = correctSeed();
        launch(k_correct_positions{cur->view(), (const uint32_t*)cellStart.as<uint32_t>(),
                           (uint32_t)np, g, bnd, c, alt->px.as<float>(), alt->py.as<float>(), alt->pz.as<float>()}, blocks(np));
        copy32(cur->px.p, alt->px.p, np * 4);
        copy32(cur->py.p, alt->py.p, np * 4);
        copy32(cur->pz.p, alt->pz.p, np * 4);
        binned = false;
    }

    // cell-sorted binning: key + count, scan, scatter, per-cell ordering, permute
    void binParticles() {
        if(!np) {
            fill32(cellStart.p, 0, (nC + 1) * 4);
            fill32(occ.p, 0, occ.bytes);
            binned    = true;
            packedVel = true;
            return;
        }
        const ParticleBufs& P = *cur;
        fill32(occ.p, 0, occ.bytes);
        launch(k_cell_key{P.px.as<float>(), P.py.as<float>(), P.pz.as<float>(), (uint32_t)np, g, key.as<uint32_t>(), cellCount.as<uint32_t>(),
                          occ.as<uint8_t>()},
               blocks(np));
        scan(cellCount.as<uint32_t>(), cellStart.as<uint32_t>(), nC, cellStart.as<uint32_t>() + nC);
        launch(k_scatter{key.as<uint32_t>(), (uint32_t)np, cellStart.as<uint32_t>(), cellCount.as<uint32_t>(), perm.as<uint32_t>()}, blocks(np));
        launch(k_segsort{cellStart.as<uint32_t>(), (uint32_t)nC, perm.as<uint32_t>()}, blocks(nC));
        ktBegin(BNN_KT_PERMUTE);
        launch(k_permute{cur->view(), alt->view(), perm.as<uint32_t>(),
                           (uint32_t)np, g}, blocks(np));
        if(affineScheme())
            launch(k_permute_affine{cur->affine(), alt->affine(),
                               (const uint32_t*)perm.as<uint32_t>(), (uint32_t)np}, blocks(np));
        ktEnd(BNN_KT_PERMUTE);
        std::swap(cur, alt);
        binned    = true;
        packedVel = true;
    }

    void requireBinned() {
        if(!binned) binParticles();
    }

    // APIC_Solver.cpp:52-91 / AFLIP_Solver.cpp:48-90 as a column gather; the column segment length
    // p2gSeg (4, 8 or 16) does not change the result
    template<bool kAffine, int kSeg> void p2gLaunch(const FaceOut& out) {
        const size_t n = (size_t)(g.nx + 1) * (g.ny + 1) * (((size_t)g.nz + 1 + kSeg - 1) / kSeg);  // columns x segments
        launch(k_p2g_col<kAffine, kSeg>{cur->view(), kAffine ? cur->affine() : Affine{}, (const uint32_t*)cellStart.as<uint32_t>(), g, out,
                                        (const uint8_t*)occ.as<uint8_t>()},
               blocks(n));
    }
    template<bool kAffine> void p2gSegLaunch(const FaceOut& out) {
        if(p2gSeg == 4) p2gLaunch<kAffine, 4>(out);
        else if(p2gSeg == 16) p2gLaunch<kAffine, 16>(out);
        else p2gLaunch<kAffine, 8>(out);
    }
    void mapParticles2Grid() {
        requireBinned();
        if(!packedVel && np) {  // velocities changed since the binning (a G2P with no move after it)
            launch(k_pack_velocity{cur->view(), (uint32_t)np}, blocks(np));
            packedVel = true;
        }
        FaceOut out{u.as<float>(), v.as<float>(), w.as<float>(), vu.as<uint8_t>(), vv.as<uint8_t>(), vw.as<uint8_t>()};
        ktBegin(BNN_KT_P2G);
        if(affineScheme()) {
            p2gSegLaunch<true>(out);  // APIC_Solver.cpp:52-91, AFLIP_Solver.cpp:48-90
        } else {  // FLIP_Solver.cpp:57-95: one wavefront per 8^3 tile of nodes
            const uint32_t tiles = (uint32_t)(((g.nx + 8) / 8) * ((g.ny + 8) / 8) * ((g.nz + 8) / 8));
            launch(k_p2g_tile{cur->view(), (const uint32_t*)cellStart.as<uint32_t>(), g, out, (const uint8_t*)occ.as<uint8_t>()}, tiles);
        }
        ktEnd(BNN_KT_P2G);
    }

    // PIC_Solver.cpp:985-1090: the 4 Jacobi layers of each component in one launch over its 8^3 face
    // bricks. Faces filled by this call are marked valid with the generation value extGen (2 or 3,
    // alternating per call), which the same launch reads as invalid on entry.
    void extrapolateVelocity() {
        struct C {
            DevBuf *a, *val, *ext;
            int     sx, sy, sz;
        } comps[3] = {{&u, &vu, &eu, g.nx + 1, g.ny, g.nz}, {&v, &vv, &ev_, g.nx, g.ny + 1, g.nz}, {&w, &vw, &ew, g.nx, g.ny, g.nz + 1}};
        extGen = extGen == 2 ? 3 : 2;
        for(auto& c : comps) {
            const uint32_t nbk = (uint32_t)(((c.sx + 7) / 8) * ((c.sy + 7) / 8) * ((c.sz + 7) / 8));
            ktBegin(BNN_KT_EXTRAP);
            launch(k_extrap_tile{c.a->as<float>(), c.val->as<uint8_t>(), c.ext->as<uint8_t>(), c.sx, c.sy, c.sz, extGen}, nbk);
            ktEnd(BNN_KT_EXTRAP);
        }
    }

    Co<void> ensureWeights() {
        if(!weightsDirty) co_return;
        fill32(counters.p, 0, 4);
        launch(k_weights{(const float*)bsdf.as<float>(), g, wu.as<float>(), wv.as<float>(), ww.as<float>(), counters.as<uint32_t>()}, blocks(nN));
        readback(hostU32, counters.p, 4);
        co_await syncPoint();
        solidFaces      = hostU32[0];
        weightsDirty    = false;
        weightsComputed = true;
    }

    // PIC_Solver.cpp:1094-1115. Uses the weights as they stand: before the first pressure projection
    // they are the zero fill of PIC_Data.cpp:109 (FLIP_Solver.cpp:45 constrains before
    // computeCellWeights runs), so every extrapolated face qualifies then. A no-op while the boundary
    // SDF is identically zero (zero gradient) or no face has weight < FLT_MIN.
    void constrainGridVelocity() {
        if(!bsdfNonzero) return;
        if(weightsComputed && solidFaces == 0) return;
        DevBuf* vel[3] = {&u, &v, &w};
        DevBuf* tmp[3] = {&ut, &vt, &wt};
        DevBuf* ext[3] = {&eu, &ev_, &ew};
        DevBuf* wgt[3] = {&wu, &wv, &ww};
        size_t  cnt[3] = {nU, nV, nW};
        for(int d = 0; d < 3; ++d) {
            VelGrid G{u.as<float>(), v.as<float>(), w.as<float>()};
            // the new values of the faces that change go to tmp first: the interpolation around one
            // face reads its neighbours, which must still hold the values before this component's pass
            launch(k_constrain{G, g, d, (const uint8_t*)ext[d]->as<uint8_t>(),
                               (const float*)wgt[d]->as<float>(), (const float*)bsdf.as<float>(), tmp[d]->as<float>()}, blocks((cnt[d] + 3) / 4));
            launch(k_constrain_apply{vel[d]->as<float>(), (const float*)tmp[d]->as<float>(), (const uint8_t*)ext[d]->as<uint8_t>(),
                                     (const float*)wgt[d]->as<float>(), cnt[d]},
                   blocks((cnt[d] + 3) / 4));
        }
    }

    void backupGridVelocity() {  // FLIP_Data.h:48-52
        copy32(uo.p, u.p, nU * 4);
        copy32(vo.p, v.p, nV * 4);
        copy32(wo.p, w.p, nW * 4);
    }

    void addGravity(float dt) {  // PIC_Solver.cpp:505-529
        if(!prm.apply_gravity) return;
        if(prm.gravity_type == 0) {
            const float gdt = 9.81f * dt;
            launch(k_add_const{v.as<float>(), nV, gdt, 1}, blocks(nV));
        } else if(prm.gravity_type == 1) {
            DevBuf* vel[3] = {&u, &v, &w};
            size_t  cnt[3] = {nU, nV, nW};
            for(int d = 0; d < 3; ++d)
                launch(k_add_const{vel[d]->as<float>(), cnt[d], grav[d] * dt, 0}, blocks(cnt[d]));
        } else {
            DevBuf* vel[3] = {&u, &v, &w};
            for(int d = 0; d < 3; ++d)
                launch(k_gravity_center{vel[d]->as<float>(), g, d,
                                   prm.gravity_center[0], prm.gravity_center[1], prm.gravity_center[2], prm.gravity_type == 2 ? 1 : 0, dt}, blocks(nU));
        }
    }

    void computeFluidSDF() {  // PIC_Solver.cpp:601-689
        requireBinned();
        ktBegin(BNN_KT_FLUID_SDF);
        const unsigned tiles = blocks(g.nx, kTileX) * blocks(g.ny, kTileY) * blocks(g.nz, kTileZ);
        launch(k_fluid_sdf{cur->view(), (const uint32_t*)cellStart.as<uint32_t>(), g, (const float*)bsdf.as<float>(), phi.as<float>(),
                           (const uint8_t*)occ.as<uint8_t>()},
               tiles);
        ktEnd(BNN_KT_FLUID_SDF);
    }

    // active bricks + matrix-free system (PIC_Solver.cpp:693-893)
    Co<void> computeSystem(float dt) {
        // warm start: keep the previous pressure and its brick table before the slots are rebuilt
        warmNext = prm.warm_start && prm.pressure_mode == BNN_PRESSURE_REFINED && prevPressure;
        if(warmNext) {
            prevBrickSlot.reserve((size_t)nb * 4);
            copy32(prevBrickSlot.p, brickSlot.p, (size_t)nb * 4);
            x64prev.swapWith(x64);
        }
        prevPressure = false;
        fill32(counters.p, 0, 16);
        uint32_t* cnt = counters.as<uint32_t>();
        launch(k_brick_flags{(const float*)phi.as<float>(), g, bflag.as<uint32_t>(), cnt + 1}, nb);
        scan(bflag.as<uint32_t>(), bscan.as<uint32_t>(), nb, cnt + 2);
        readback(hostU32, counters.p, 16);
        co_await syncPoint();
        nFluid = hostU32[1];
        nSlots = hostU32[2];
        if(nSlots > nb) throw Error(BNN_ERR_DEVICE, "active brick count exceeds the brick grid");
        ensureSlots(std::max<uint32_t>(nSlots, 1));
        // the warm start swapped x64 with the previous-pressure buffer, which may be smaller than the
        // slot capacity (or not yet allocated): ensureSlots only grows buffers when the capacity grows
        if(prm.pressure_mode == BNN_PRESSURE_REFINED) x64.reserve((size_t)slotCap * 512 * 8);
        launch(k_brick_slots{(const uint32_t*)bflag.as<uint32_t>(), (const uint32_t*)bscan.as<uint32_t>(), nb, brickSlot.as<int32_t>(), slotBrick.as<int32_t>()}, blocks(nb));
        if(nSlots == 0) {
            warmNext    = false;
            systemValid = true;
            co_return;
        }
        launch(k_brick_nbrs{(const int32_t*)slotBrick.as<int32_t>(), nSlots, (const int32_t*)brickSlot.as<int32_t>(), g, nbrTab.as<int32_t>()}, blocks(nSlots));
        const bool refined = prm.pressure_mode == BNN_PRESSURE_REFINED;
        SystemOut  S{rhs64.as<double>(), refined ? diag64.as<double>() : nullptr, refined ? cx64.as<double>() : nullptr,
                    refined ? cy64.as<double>() : nullptr, refined ? cz64.as<double>() : nullptr, rhs32.as<float>(), diag32.as<float>(),
                    cx32.as<float>(), cy32.as<float>(), cz32.as<float>(), invd.as<float>()};
        VelGrid G{u.as<float>(), v.as<float>(), w.as<float>()};
        launch(k_assemble{(const int32_t*)slotBrick.as<int32_t>(), g, (const float*)phi.as<float>(), G, (const float*)wu.as<float>(), (const float*)wv.as<float>(), (const float*)ww.as<float>(), dt, S}, nSlots);
        if(warmNext) {
            const size_t need = (size_t)nSlots * 512 * 8;
            if(!x64.p || x64.bytes < need || !x64prev.p || !prevBrickSlot.p || prevBrickSlot.bytes < (size_t)nb * 4)
                throw Error(BNN_ERR_DEVICE, "warm start buffers are not allocated");
            launch(k_warm_start{(const int32_t*)slotBrick.as<int32_t>(), (const int32_t*)prevBrickSlot.as<int32_t>(),
                                (const double*)x64prev.as<double>(), (const float*)phi.as<float>(), g, nSlots, x64.as<double>()},
                   redBlocks());
        }
        if(prm.preconditioner == BNN_PRECOND_MG) co_await buildHierarchy();
        systemValid = true;
    }

    // ---------------------------------------------------------------------------------------------
    // multigrid hierarchy (rebuilt after every assembly when the MG preconditioner is selected)
    struct MgLevel {
        int      nx = 0, ny = 0, nz = 0, nbx = 0, nby = 0, nbz = 0;
        uint32_t nb = 0, nSlots = 0, slotCap = 0;
        DevBuf   flag, scan, brickSlot, slotBrick, nbr, diag, cx, cy, cz, inv, x, b, r, ptab;
    };
    DevBuf                                ptab0;  // parent table of the fine level
    std::vector<std::unique_ptr<MgLevel>> mg;  // mg[0] unused (the fine level lives in the PCG buffers)
    int                                   mgLevels = 0;
    static constexpr uint32_t             kCoarsestSlots = 1;  // coarsen until one brick remains

    LevelDev fineLevel() {
        return LevelDev{g.nx, g.ny, g.nz, g.nbx, g.nby, g.nbz, brickSlot.as<int32_t>(), slotBrick.as<int32_t>(), nbrTab.as<int32_t>(),
                        diag32.as<float>(), cx32.as<float>(), cy32.as<float>(), cz32.as<float>(), invd.as<float>(),
                        z.as<float>(), r.as<float>(), q.as<float>(), nSlots, ptab0.as<int32_t>()};
    }
    LevelDev levelDev(int l) {
        if(l == 0) return fineLevel();
        MgLevel& L = *mg[l];
        return LevelDev{L.nx, L.ny, L.nz, L.nbx, L.nby, L.nbz, L.brickSlot.as<int32_t>(), L.slotBrick.as<int32_t>(), L.nbr.as<int32_t>(),
                        L.diag.as<float>(), L.cx.as<float>(), L.cy.as<float>(), L.cz.as<float>(), L.inv.as<float>(),
                        L.x.as<float>(), L.b.as<float>(), L.r.as<float>(), L.nSlots, L.ptab.as<int32_t>()};
    }

    // Coarse levels halve the grid until one brick remains (or every axis fits one brick). All levels'
    // active bricks are flagged in one pass from the finest level's and counted together, so the
    // host waits once for the whole hierarchy.
    Co<void> buildHierarchy() {
        mgLevels = 1;
        if(mg.empty()) mg.emplace_back(nullptr);
        if(nSlots <= kCoarsestSlots) co_return;
        k_mg_flag_levels fl{};
        fl.fineSlotBrick = slotBrick.as<int32_t>();
        fl.nFine         = nSlots;
        fl.fnbx          = g.nbx;
        fl.fnby          = g.nby;
        int nx = g.nx, ny = g.ny, nz = g.nz, nCand = 0;
        while((nx > 8 || ny > 8 || nz > 8) && nCand + 1 < kMaxMgLevels) {
            const int l = ++nCand;
            if((int)mg.size() <= l) mg.emplace_back(new MgLevel());
            MgLevel& L = *mg[l];
            nx = (nx + 1) / 2;
            ny = (ny + 1) / 2;
            nz = (nz + 1) / 2;
            L.nx = nx; L.ny = ny; L.nz = nz;
            L.nbx = (nx + 7) / 8; L.nby = (ny + 7) / 8; L.nbz = (nz + 7) / 8;
            L.nb  = (uint32_t)(L.nbx * L.nby * L.nbz);
            L.flag.reserve(L.nb * 4);
            L.scan.reserve(L.nb * 4);
            L.brickSlot.reserve(L.nb * 4);
            fill32(L.flag.p, 0, L.nb * 4);
            fl.cnbx[l] = L.nbx;
            fl.cnby[l] = L.nby;
            fl.flag[l] = L.flag.as<uint32_t>();
        }
        if(nCand == 0) co_return;
        fl.nLevels = nCand;
        launch(fl, blocks(nSlots));
        uint32_t* count = counters.as<uint32_t>() + 16;  // level l's active brick count at count[l]
        for(int l = 1; l <= nCand; ++l) scan(mg[l]->flag.as<uint32_t>(), mg[l]->scan.as<uint32_t>(), mg[l]->nb, count + l);
        readback(hostU32, count, (size_t)(nCand + 1) * 4);
        co_await syncPoint();
        uint32_t prevSlots = nSlots;
        for(int l = 1; l <= nCand && prevSlots > kCoarsestSlots; ++l) {
            MgLevel&       L    = *mg[l];
            const LevelDev prev = levelDev(l - 1);
            L.nSlots            = hostU32[l];
            if(L.nSlots == 0 || L.nSlots > L.nb) throw Error(BNN_ERR_DEVICE, "multigrid level has an invalid brick count");
            if(L.nSlots > L.slotCap) {
                const uint32_t c  = L.nSlots + L.nSlots / 4 + 4;
                const size_t   nv = (size_t)c * 512 * 4;
                for(DevBuf* b : {&L.diag, &L.cx, &L.cy, &L.cz, &L.inv, &L.x, &L.b, &L.r}) b->reserve(nv);
                L.slotBrick.reserve(c * 4);
                L.nbr.reserve((size_t)c * 24);
                L.slotCap = c;
            }
            GridGeom lg = g;
            lg.nx = L.nx; lg.ny = L.ny; lg.nz = L.nz;
            lg.nbx = L.nbx; lg.nby = L.nby; lg.nbz = L.nbz;
            launch(k_brick_slots{(const uint32_t*)L.flag.as<uint32_t>(), (const uint32_t*)L.scan.as<uint32_t>(), L.nb, L.brickSlot.as<int32_t>(), L.slotBrick.as<int32_t>()}, blocks(L.nb));
            launch(k_brick_nbrs{(const int32_t*)L.slotBrick.as<int32_t>(), L.nSlots, (const int32_t*)L.brickSlot.as<int32_t>(), lg, L.nbr.as<int32_t>()}, blocks(L.nSlots));
            const LevelDev cur = levelDev(l);
            launch(k_mg_coarsen{prev, cur, L.diag.as<float>(), L.cx.as<float>(), L.cy.as<float>(), L.cz.as<float>(), L.inv.as<float>()}, L.nSlots);
            // the restriction writes only coarse cells under an active fine brick; the others stay 0
            fill32(L.b.p, 0, (size_t)L.nSlots * 512 * 4);
            // the finer level's parent table (its bricks' parents in this level)
            DevBuf& pt = l == 1 ? ptab0 : mg[l - 1]->ptab;
            pt.reserve((size_t)std::max<uint32_t>(prev.nSlots, 1u) * 32);
            launch(k_mg_parent_table{prev, levelDev(l), pt.as<int32_t>()}, blocks(prev.nSlots));
            prevSlots = L.nSlots;
            ++mgLevels;
        }
    }

    // z = M^-1 r on the fine level (b = r, x = z); symmetric V-cycle with the brick-hybrid smoother.
    // One block per brick. On each level the pre-smoothed x lives in L.r and the final x in L.x.
    // Returns true when the fine level's post-smoother also wrote the per-brick partials of z.r to pz
    // (false when the fine level is the coarsest: the caller then forms z.r itself).
    // finePreDone: the finest level's pre-smoother already ran (fused with the PCG update).
    bool vcycle(const PcgState* stp, bool finePreDone = false) {
        const int   nu    = prm.mg_sweeps;
        const float omega = prm.mg_omega;
        for(int l = 0; l + 1 < mgLevels; ++l) {
            const LevelDev L = levelDev(l), C = levelDev(l + 1);
            const unsigned G = std::max<uint32_t>(L.nSlots, 1u);
            if(!(l == 0 && finePreDone)) {
                ktBegin(l == 0 ? BNN_KT_MG_PRE : BNN_KT_MG_COARSE);
                launch(k_mg_smooth_brick<0, 0>{L, C, omega, nu, stp, (double*)nullptr}, G);
                ktEnd(l == 0 ? BNN_KT_MG_PRE : BNN_KT_MG_COARSE);
            }
            ktBegin(l == 0 ? BNN_KT_MG_RESTRICT : BNN_KT_MG_COARSE);
            launch(k_mg_residual_restrict{L, C, stp}, G);
            ktEnd(l == 0 ? BNN_KT_MG_RESTRICT : BNN_KT_MG_COARSE);
        }
        const LevelDev Lc = levelDev(mgLevels - 1);
        ktBegin(BNN_KT_MG_COARSE);
        if(Lc.nSlots == 1) {
            launch(k_mg_coarsest1{Lc, prm.mg_coarsest_sweeps, stp}, 1);
        } else {
            launch(k_mg_coarsest{Lc, prm.mg_coarsest_sweeps, stp}, 1);
        }
        ktEnd(BNN_KT_MG_COARSE);
        for(int l = mgLevels - 2; l >= 0; --l) {
            const LevelDev L = levelDev(l), C = levelDev(l + 1);
            const uint32_t G = std::max<uint32_t>(L.nSlots, 1u);
            if(l == 0) {  // finest level: correction as its own streaming pass, then smooth and form z.r
                ktBegin(BNN_KT_MG_PROLONG);
                launch(k_mg_prolong{L, C, omega, stp}, G);
                ktEnd(BNN_KT_MG_PROLONG);
                ktBegin(BNN_KT_MG_POST);
                launch(k_mg_smooth_brick<2, 1>{L, C, omega, nu, stp, pz.as<double>()}, G);
                ktEnd(BNN_KT_MG_POST);
            } else {
                ktBegin(BNN_KT_MG_COARSE);
                launch(k_mg_smooth_brick<1, 0>{L, C, omega, nu, stp, (double*)nullptr}, G);
                ktEnd(BNN_KT_MG_COARSE);
            }
        }
        return mgLevels > 1;
    }

    // IC(0) of the reference, one launch per hyperplane i + j + k = h (h = 0 .. nx + ny + nz - 3)
    Ic0Dev ic0Dev() {
        return Ic0Dev{brickSlot.as<int32_t>(), diag32.as<float>(), cx32.as<float>(), cy32.as<float>(), cz32.as<float>(),
                      icInv.as<float>(), icLx.as<float>(), icLy.as<float>(), icLz.as<float>()};
    }
    int ic0Planes() const { return g.nx + g.ny + g.nz - 2; }
    void formIC0() {  // PCGSolver.cpp:398-466 (formPreconditioner at the start of every solve)
        const Ic0Dev   D  = ic0Dev();
        const unsigned nb = blocks((size_t)g.nx * g.ny);
        for(int h = 0; h < ic0Planes(); ++h) launch(k_ic0_factor_plane{D, g, h}, nb);
    }
    // z = (L L^T)^-1 r (PCGSolver.cpp:243-269): forward planes ascending into z, then backward in place
    void applyIC0(const float* rv, float* zv, const PcgState* stp) {
        const Ic0Dev   D  = ic0Dev();
        const unsigned nb = blocks((size_t)g.nx * g.ny);
        const int      H  = ic0Planes();
        for(int h = 0; h < H; ++h) launch(k_ic0_lower_plane{D, g, h, rv, zv, stp}, nb);
        for(int h = H - 1; h >= 0; --h) launch(k_ic0_upper_plane{D, g, h, zv, stp}, nb);
    }

    unsigned redBlocks() const { return std::min<unsigned>(std::max<uint32_t>(nSlots, 1u), (unsigned)kMaxRedBlocks); }

    // Iterations to queue after a poll (the stop flag ends the ones not needed, but each still costs a
    // few launches): as many as a geometric fit of max|r| from iteration 0 (res0) to now (res after
    // iter iterations) says remain until max|r| < tol, at least 1 and at most pcgChunkMax; doubling the
    // last chunk while the fit has no progress to go by.
    uint32_t pcgNextChunk(float res0, float res, float tol, uint32_t iter, uint32_t last) const {
        const uint32_t doubled = std::min<uint32_t>(last * 2, pcgChunkMax);
        if(iter == 0 || !(res > 0.f) || !(res < res0) || !(tol > 0.f) || !(tol < res)) return doubled;
        const double rate = std::log((double)res / (double)res0) / (double)iter;  // log of the mean reduction per iteration
        const double need = std::ceil(std::log((double)tol / (double)res) / rate);
        if(!(need >= 1.0)) return 1u;
        return (uint32_t)std::min<double>(need, (double)pcgChunkMax);
    }

    // FP32 PCG on b -> x (PCGSolver.cpp:143-199); returns success
    Co<bool> pcg32(const float* b, float relTol, uint32_t maxIter, uint32_t& iters, float& res, float& bmax) {
        PcgState*      dst = pst.as<PcgState>();
        double*        dpa = pa.as<double>();
        double*        dpb = pb.as<double>();
        float *        sA = s0.as<float>(), *sB = s1.as<float>();
        const unsigned G   = redBlocks();
        const int      jac = prm.preconditioner == BNN_PRECOND_JACOBI ? 1 : 0;
        launch(k_pcg_init{b, (const float*)invd.as<float>(), x.as<float>(), r.as<float>(), z.as<float>(), sA, nSlots, dpa, dpb, jac}, G);
        // partials of z.r: per block (G) from k_pcg_init / k_pcg_dot_zr, or per brick from the V-cycle
        const double* rz  = dpa;
        int           nrz = (int)G;
        const bool    ic0 = prm.preconditioner == BNN_PRECOND_IC0;
        if(ic0) formIC0();
        auto          zr  = [&](const PcgState* stp, bool finePreDone) {
            if(ic0) {
                applyIC0(r.as<float>(), z.as<float>(), stp);
                launch(k_pcg_dot_zr{(const float*)z.as<float>(), (const float*)r.as<float>(), nSlots, stp, dpa}, G);
            } else if(vcycle(stp, finePreDone)) {
                rz  = pz.as<double>();
                nrz = (int)nSlots;
            } else {
                launch(k_pcg_dot_zr{(const float*)z.as<float>(), (const float*)r.as<float>(), nSlots, stp, dpa}, G);
            }
        };
        if(!jac) zr(nullptr, false);
        // multigrid over several levels: the update is fused with the finest pre-smoother, and its
        // max|r| partials are per brick
        const bool fuseUpdate = !jac && !ic0 && mgLevels > 1;
        launch(k_pcg_init_fin{dst, rz, nrz, (const double*)dpb, (int)G, relTol, maxIter}, 1);
        // the first chunk is queued without a poll (the stop flag ends it early when the system is
        // already solved); each later poll sizes the next chunk from max|r| at iteration 0, which is
        // max|b| since x starts at 0
        uint32_t chunk = pcgChunkFirst, launched = 0;
        for(bool first = true;; first = false) {
            if(!first) {
                readback(hostSt, dst, sizeof(PcgState));
                co_await syncPoint();
                if(hostSt->done) break;
                chunk = pcgNextChunk(hostSt->bmax, hostSt->res, hostSt->tol, hostSt->iter, chunk);
            }
            // never queue more iterations than can remain
            const uint32_t left = first ? maxIter : (hostSt->maxIter > hostSt->iter ? hostSt->maxIter - hostSt->iter : 1u);
            const uint32_t n    = std::min(chunk, left);
            for(uint32_t c = 0; c < n; ++c) {
                ktBegin(BNN_KT_PCG_SPMV);
                launch(k_pcg_spmv{(const int32_t*)nbrTab.as<int32_t>(), (const float*)z.as<float>(), (const float*)sA, sB, q.as<float>(), (const float*)diag32.as<float>(), (const float*)cx32.as<float>(), (const float*)cy32.as<float>(), (const float*)cz32.as<float>(), nSlots, (const PcgState*)dst, dpa}, G);
                ktEnd(BNN_KT_PCG_SPMV);
                launch(k_pcg_alpha{dst, (const double*)dpa, (int)G}, 1);
                ktBegin(BNN_KT_PCG_UPDATE);
                if(fuseUpdate)
                    launch(k_pcg_update_pre{(const float*)sB, x.as<float>(), fineLevel(), prm.mg_omega, prm.mg_sweeps, (const PcgState*)dst, dpb},
                           std::max<uint32_t>(nSlots, 1u));
                else
                    launch(k_pcg_update{(const float*)sB, (const float*)q.as<float>(), (const float*)invd.as<float>(), x.as<float>(), r.as<float>(), z.as<float>(), nSlots, (const PcgState*)dst, dpa, dpb, jac}, G);
                ktEnd(BNN_KT_PCG_UPDATE);
                if(!jac) zr(dst, fuseUpdate);
                launch(k_pcg_beta{dst, rz, nrz, (const double*)dpb, fuseUpdate ? (int)nSlots : (int)G}, 1);
                std::swap(sA, sB);
            }
            launched += n;
            if(launched > maxIter + 256) throw Error(BNN_ERR_DEVICE, "PCG device state stopped advancing");
        }
        iters = hostSt->iter;
        res   = hostSt->res;
        bmax  = hostSt->bmax;
        co_return hostSt->success != 0;
    }

    Co<bool> solveSystem() {  // PIC_Solver.cpp:897-907
        lastRefined = prm.pressure_mode == BNN_PRESSURE_REFINED;
        if(nSlots == 0) {
            st.pcg_iterations   = 0;
            st.outer_iterations = 0;
            st.pcg_residual     = 0;
            st.rhs_max          = 0;
            st.pcg_success      = 1;
            co_return true;
        }
        if(!lastRefined) {
            uint32_t it;
            float    res, bmax;
            bool     ok         = co_await pcg32(rhs32.as<float>(), prm.cg_relative_tolerance, prm.max_cg_iteration, it, res, bmax);
            st.pcg_iterations   = it;
            st.outer_iterations = 1;
            st.pcg_residual     = res;
            st.rhs_max          = bmax;
            st.pcg_success      = ok;
            co_return ok;
        }
        // FP64 defect correction around FP32 PCG: r = b - A x in FP64, solve A d = r in FP32, x += d
        const double   tol      = prm.refine_tolerance > 0 ? prm.refine_tolerance : 1e-10;
        const float    innerTol = prm.refine_inner_tolerance;
        const uint32_t maxOuter = 30;
        const unsigned G        = redBlocks();
        uint32_t       total = 0, outer = 0;
        bool           ok = false;
        double         bmax = 0, res = 0;
        for(int first = 1;; first = 0) {
            co_await barrier();  // the scenes' outer passes start together
            const int zeroX = first && !warmNext;  // x = 0 (else x64 holds the warm start or the sum so far)
            launch(k_residual64{(const int32_t*)nbrTab.as<int32_t>(), (const double*)x64.as<double>(), (const double*)rhs64.as<double>(), (const double*)diag64.as<double>(), (const double*)cx64.as<double>(), (const double*)cy64.as<double>(), (const double*)cz64.as<double>(), rhs32.as<float>(), nSlots, pa.as<double>(), pb.as<double>(), zeroX}, G);
            launch(k_residual64_fin{ost.as<OuterState>(), (const double*)pa.as<double>(), (const double*)pb.as<double>(), (int)G}, 1);
            readback(hostOst, ost.p, sizeof(OuterState));
            co_await syncPoint();
            if(first) {
                bmax = hostOst->bmax;
                if(bmax < 1.17549435e-38) {
                    fill32(x64.p, 0, (size_t)nSlots * 512 * 8);
                    ok  = true;
                    res = 0;
                    break;
                }
            }
            res = hostOst->res;
            if(res <= tol * bmax) {
                if(zeroX) fill32(x64.p, 0, (size_t)nSlots * 512 * 8);
                ok = true;
                break;
            }
            if(outer >= maxOuter) break;
            // the inner solve need only reduce the residual by the factor still missing to the bar
            // (with a margin of 4), when that is less than refine_inner_tolerance asks for
            const double missing = tol * bmax / res;
            const float  passTol = missing > 4.0 * innerTol ? (float)std::min(0.25 * missing, 0.5) : innerTol;
            uint32_t     it;
            float        r32, b32;
            co_await pcg32(rhs32.as<float>(), passTol, prm.max_cg_iteration, it, r32, b32);
            total += it;
            launch(k_accumulate64{x64.as<double>(), (const float*)x.as<float>(), nSlots, zeroX}, G);
            ++outer;
            if(it == 0) break;  // inner solve made no progress
        }
        st.pcg_iterations   = total;
        st.outer_iterations = outer;
        st.pcg_residual     = res;
        st.rhs_max          = bmax;
        st.pcg_success      = ok;
        prevPressure        = ok;
        co_return ok;
    }

    void updateProjectedVelocity(float dt) {  // PIC_Solver.cpp:911-981
        DevBuf* vel[3] = {&u, &v, &w};
        DevBuf* val[3] = {&vu, &vv, &vw};
        DevBuf* wgt[3] = {&wu, &wv, &ww};
        size_t  cnt[3] = {nU, nV, nW};
        for(int d = 0; d < 3; ++d) {
            if(lastRefined && nSlots) {
                PressView<double> pv{x64.as<double>(), brickSlot.as<int32_t>()};
                launch(k_update_velocity<double>{vel[d]->as<float>(),
                                   val[d]->as<uint8_t>(), (const float*)wgt[d]->as<float>(), (const float*)phi.as<float>(), pv, g, d, dt}, blocks(cnt[d]));
            } else {
                PressView<float> pv{x.as<float>(), brickSlot.as<int32_t>()};
                launch(k_update_velocity<float>{vel[d]->as<float>(),
                                   val[d]->as<uint8_t>(), (const float*)wgt[d]->as<float>(), (const float*)phi.as<float>(), pv, g, d, dt}, blocks(cnt[d]));
            }
        }
    }

    Co<bool> pressureProjection(float dt) {  // PIC_Solver.cpp:533-539
        stageBegin(BNN_STAGE_WEIGHTS);
        co_await ensureWeights();
        stageEnd(BNN_STAGE_WEIGHTS);
        stageBegin(BNN_STAGE_PHI);
        computeFluidSDF();
        stageEnd(BNN_STAGE_PHI);
        stageBegin(BNN_STAGE_ASSEMBLY);
        co_await computeSystem(dt);
        stageEnd(BNN_STAGE_ASSEMBLY);
        stageBegin(BNN_STAGE_PCG);
        co_await barrier();
        bool ok = co_await solveSystem();
        co_await barrier();
        stageEnd(BNN_STAGE_PCG);
        stageBegin(BNN_STAGE_UPDATE);
        updateProjectedVelocity(dt);
        stageEnd(BNN_STAGE_UPDATE);
        co_return ok;
    }

    void mapGrid2Particles() {  // FLIP_Solver.cpp:99-133
        if(!np) return;
        requireBinned();
        VelGrid G{u.as<float>(), v.as<float>(), w.as<float>()};
        OldGrid O{uo.as<float>(), vo.as<float>(), wo.as<float>()};
        ktBegin(BNN_KT_G2P);
        if(prm.transfer_scheme == BNN_TRANSFER_APIC)
            launch(k_g2p_affine<1>{cur->view(), cur->affine(), (uint32_t)np, G, O, g,
                               prm.pic_flip_ratio}, blocks(np));
        else if(prm.transfer_scheme == BNN_TRANSFER_AFLIP)
            launch(k_g2p_affine<2>{cur->view(), cur->affine(), (uint32_t)np, G, O, g,
                               prm.pic_flip_ratio}, blocks(np));
        else
            launch(k_g2p{cur->view(), (uint32_t)np, G, O, g, prm.pic_flip_ratio}, blocks(np));
        ktEnd(BNN_KT_G2P);
        packedVel = false;  // new velocities; the copy in gp / gv is refreshed before the next P2G
    }

    Co<bool> advanceVelocity(float dt) {  // FLIP_Solver.cpp:42-53 (FLIP, AFLIP); APIC_Solver.cpp:41-48 (APIC)
        stageBegin(BNN_STAGE_BIN);
        requireBinned();
        stageEnd(BNN_STAGE_BIN);
        stageBegin(BNN_STAGE_P2G);
        mapParticles2Grid();
        stageEnd(BNN_STAGE_P2G);
        if(prm.transfer_scheme == BNN_TRANSFER_APIC) {  // no extrapolation, constrain or backup before the solve
            stageBegin(BNN_STAGE_CONSTRAIN);
            addGravity(dt);
            stageEnd(BNN_STAGE_CONSTRAIN);
            bool ok = co_await pressureProjection(dt);
            stageBegin(BNN_STAGE_EXTRAPOLATE);
            extrapolateVelocity();
            stageEnd(BNN_STAGE_EXTRAPOLATE);
            stageBegin(BNN_STAGE_CONSTRAIN);
            constrainGridVelocity();
            stageEnd(BNN_STAGE_CONSTRAIN);
            stageBegin(BNN_STAGE_G2P);
            mapGrid2Particles();
            stageEnd(BNN_STAGE_G2P);
            co_return ok;
        }
        stageBegin(BNN_STAGE_EXTRAPOLATE);
        extrapolateVelocity();
        stageEnd(BNN_STAGE_EXTRAPOLATE);
        stageBegin(BNN_STAGE_CONSTRAIN);
        constrainGridVelocity();
        backupGridVelocity();
        addGravity(dt);
        stageEnd(BNN_STAGE_CONSTRAIN);
        bool ok = co_await pressureProjection(dt);
        stageBegin(BNN_STAGE_EXTRAPOLATE);
        extrapolateVelocity();
        stageEnd(BNN_STAGE_EXTRAPOLATE);
        stageBegin(BNN_STAGE_CONSTRAIN);
        constrainGridVelocity();
        stageEnd(BNN_STAGE_CONSTRAIN);
        stageBegin(BNN_STAGE_G2P);
        mapGrid2Particles();
        stageEnd(BNN_STAGE_G2P);
        co_return ok;
    }

    void beginStats() {
        for(double& a : stageAcc) a = 0;
        for(double& a : ktAcc) a = 0;
        for(uint64_t& c : ktCnt) c = 0;
    }
    Co<void> endStats(float dt) {
        co_await syncPoint();
        collectTimers();
        for(int s = 0; s < BNN_NUM_STAGES; ++s) st.stage_ms[s] = prm.timing ? stageAcc[s] : 0.0;
        for(int k = 0; k < BNN_NUM_KERNEL_TIMERS; ++k) {
            st.kernel_ms[k]       = ktAcc[k];
            st.kernel_launches[k] = ktCnt[k];
        }
        // algorithmic (compulsory) bytes per stage
        const double Np = (double)np, C = (double)nC, F = (double)(nU + nV + nW), V = (double)nN, A = (double)nFluid;
        st.stage_bytes[BNN_STAGE_MOVE]        = 24 * Np + 4 * F;
        st.stage_bytes[BNN_STAGE_BIN]         = 72 * Np + 12 * C;
        st.stage_bytes[BNN_STAGE_P2G]         = 24 * Np + 5 * F;
        st.stage_bytes[BNN_STAGE_EXTRAPOLATE] = 2 * 11 * F;
        st.stage_bytes[BNN_STAGE_CONSTRAIN]   = 8 * F + 8 * F / 3;
        st.stage_bytes[BNN_STAGE_WEIGHTS]     = 0;
        st.stage_bytes[BNN_STAGE_PHI]         = 12 * Np + 13 * C + 4 * V;
        st.stage_bytes[BNN_STAGE_ASSEMBLY]    = 9 * C + 4 * C + 8 * F + 24 * A;
        // PCG per inner iteration: SpMV 32 B and update 24 B per fluid cell, plus the preconditioner:
        // Jacobi 8 B (fused into the update); the V-cycle's four finest passes 89 B (pre-smoother 24,
        // residual-restriction 24.5, correction 12.5, post-smoother 28) and the coarser levels about
        // 89/7 B (each level 1/8 of the one above), once per iteration and once per inner solve; the
        // two IC(0) sweeps 2 x 24 B. FP64 defect correction: 72 B per outer pass.
        const double it = st.pcg_iterations, outer = st.outer_iterations;
        double       pcgB = 56 * A * it;
        if(prm.preconditioner == BNN_PRECOND_JACOBI) pcgB += 8 * A * it;
        else if(prm.preconditioner == BNN_PRECOND_MG) pcgB += (89 + 89.0 / 7) * A * (it + outer);
        else pcgB += 48 * A * (it + outer);
        if(lastRefined) pcgB += 72 * A * (outer + 1);
        st.stage_bytes[BNN_STAGE_PCG]         = pcgB;
        st.stage_bytes[BNN_STAGE_UPDATE]      = 4 * C + 4 * A + 13 * F;
        st.stage_bytes[BNN_STAGE_G2P]         = 36 * Np + 8 * F + 12 * F;
        st.stage_bytes[BNN_STAGE_OTHER]       = 0;
        st.num_particles     = np;
        st.num_cells         = nC;
        st.num_active_cells  = nFluid;
        st.num_active_bricks = nSlots;
        st.last_substep      = dt;
    }

    // a served pressure-solve request: pressureProjection on the current grid velocities, with stats
    Co<bool> projectionRequest(float dt) {
        if(!(dt > 0)) throw Error(BNN_ERR_INVALID_ARG, "dt must be > 0");
        beginStats();
        bool ok = co_await pressureProjection(dt);
        co_await endStats(dt);
        co_return ok;
    }

    Co<bool> step(float dt) {  // PIC_Solver.cpp:292-294 (moveParticles, advanceVelocity)
        if(!(dt > 0)) throw Error(BNN_ERR_INVALID_ARG, "dt must be > 0");
        beginStats();
        stageBegin(BNN_STAGE_MOVE);
        moveParticles(dt);
        stageEnd(BNN_STAGE_MOVE);
        if(prm.correct_position) {  // PIC_Solver.cpp:295 (after moveParticles, before advanceVelocity)
            stageBegin(BNN_STAGE_OTHER);
            correctParticlePositions(dt);
            stageEnd(BNN_STAGE_OTHER);
        }
        stageBegin(BNN_STAGE_BIN);
        binParticles();
        stageEnd(BNN_STAGE_BIN);
        bool ok = co_await advanceVelocity(dt);
        co_await endStats(dt);
        ++st.steps;
        co_return ok;
    }

    Co<float> timestepCFL() {  // PIC_Solver.cpp:342-349
        fill32(counters.p, 0, 4);
        DevBuf* vel[3] = {&u, &v, &w};
        size_t  cnt[3] = {nU, nV, nW};
        for(int d = 0; d < 3; ++d)
            launch(k_maxabs{(const float*)vel[d]->as<float>(), cnt[d], counters.as<uint32_t>()}, std::min<unsigned>(blocks(cnt[d]), 2048));
        readback(hostU32, counters.p, 4);
        co_await syncPoint();
        float m;
        std::memcpy(&m, hostU32, 4);
        float maxVel = std::max(kFltMin, m);
        float ts     = maxVel > kFltMin ? prm.cfl_factor * (g.dx / maxVel) : 3.40282347e+38f;
        co_return std::min(std::max(ts, prm.min_timestep), prm.max_timestep);
    }

    Co<bool> advanceFrame() {  // PIC_Solver.cpp:267-311
        const float fd = prm.frame_duration;
        frameTime      = 0;
        uint32_t count = 0;
        bool     ok    = true;
        float    last  = 0;
        while(frameTime < fd) {
            if(finishedFrame > 0) advanceScene(frameTime / fd);  // PIC_Solver.cpp:279-281
            float sub = co_await timestepCFL();
            float rem = fd - frameTime;
            if(frameTime + sub >= fd) sub = rem;
            else if(frameTime + 1.5f * sub >= fd) sub = rem * 0.5f;
            ok &= co_await step(sub);
            last = sub;
            frameTime += sub;
            ++count;
        }
        ++finishedFrame;
        st.finished_frame = finishedFrame;
        st.substep_count = count;
        st.last_substep  = last;
        co_return ok;
    }

    // the scene's additional boundary objects (ParticleSolver::generateBoundaries)
    void setBoundaryObjects(const bnn_boundary_object* objs, uint32_t n) {
        if(n > BNN_MAX_BOUNDARY_OBJECTS) throw Error(BNN_ERR_INVALID_ARG, "too many boundary objects");
        if(n && !objs) throw Error(BNN_ERR_INVALID_ARG, "null boundary object array");
        std::vector<BoundaryObjectHost> v(n);
        try {
            for(uint32_t i = 0; i < n; ++i) v[i].load(objs[i]);
        } catch(const std::invalid_argument& e) {
            throw Error(BNN_ERR_INVALID_ARG, e.what());
        }
        bobjs.swap(v);
        bnd = make_bnd_set(dom.lo, dom.hi, bobjs);
        if(bsdfNonzero || prm.boundary_sdf_at_init) computeBoundarySDF();
        HIP_OK(hipStreamSynchronize(stream));
    }
    bool hasDynamicBoundaries() const {
        for(const auto& o : bobjs)
            if(o.isDynamic()) return true;
        return false;
    }
    // PIC_Solver::advanceScene (PIC_Solver.cpp:70-118): animated boundary objects move to
    // (finishedFrame + fraction) and the boundary SDF is recomputed from every object
    void advanceScene(float fraction) {
        if(!hasDynamicBoundaries()) return;
        for(auto& o : bobjs) o.update(finishedFrame, (double)fraction);
        bnd = make_bnd_set(dom.lo, dom.hi, bobjs);
        computeBoundarySDF();
    }

    // PIC_Solver.cpp:315-328: z-sort at frame 0 and then every SortFrequency frames, if enabled
    bool sortDue() const {
        return prm.enable_sort_particle && !(finishedFrame > 0 && (finishedFrame + 1) % prm.sort_frequency != 0);
    }
    void sortParticles() {
        if(sortDue()) zSortParticles();
    }

    // NeighborSearch::z_sort + PointSet::sort_field (NeighborSearch.cpp:72-86, PointSet.h:114-131):
    // stable by the caller's index (the reference's tbb::parallel_sort leaves ties unordered). The
    // spare particle buffer holds the keys; only the external ids change.
    void zSortParticles() {
        if(!np) return;
        const ZsortLayout L = zsort_layout(np);
        zsHist.reserve(L.nh * 4 * 2);
        zsAcc.reserve(16);
        scanPart.reserve(std::max<size_t>(scanPart.bytes, L.partWords * 4));
        zSortWith(ZsortScratch{zsHist.as<uint32_t>(), zsAcc.as<unsigned long long>(), scanPart.as<uint32_t>()});
    }

    // the z-sort with its scratch (histograms + offsets, key accumulators, scan partials) given,
    // sized by zsort_layout(np); allocates nothing
    void zSortWith(const ZsortScratch& sc) {
        if(!np) return;
        const uint32_t n      = (uint32_t)np;
        const uint32_t ntiles = (uint32_t)((np + kRsTile - 1) / kRsTile);
        const size_t   nh     = (size_t)256 * ntiles;
        uint32_t*      lo[2]  = {alt->px.as<uint32_t>(), alt->vx.as<uint32_t>()};
        uint32_t*      hi[2]  = {alt->py.as<uint32_t>(), alt->vy.as<uint32_t>()};
        uint32_t*      val[2] = {alt->pz.as<uint32_t>(), alt->vz.as<uint32_t>()};
        const float    inv    = (float)(1.0 / (double)prm.cell_size);  // NeighborSearch.cpp:33
        launch(k_zsort_keys{cur->view(), n, inv, lo[0], hi[0], val[0]}, blocks(np));
        launch(k_zsort_acc_init{sc.acc}, 1);
        launch(k_zsort_andor{(const uint32_t*)lo[0], (const uint32_t*)hi[0], n, sc.acc}, std::min<unsigned>(blocks(np), 1024));
        HIP_OK(hipMemcpyAsync(hostU32, sc.acc, 16, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        uint64_t acc[2];
        std::memcpy(acc, hostU32, 16);
        const uint64_t varying = acc[0] ^ acc[1];
        uint32_t*      hist    = sc.hist;
        uint32_t*      offs    = hist + nh;
        int            s       = 0;
        for(int d = 0; d < 8; ++d) {
            if(((varying >> (8 * d)) & 0xFFull) == 0) continue;  // every key has the same digit here
            const int word = d >> 2, shift = (d & 3) * 8;
            launch(k_rs_hist{(const uint32_t*)(word ? hi[s] : lo[s]), n, shift, ntiles, hist}, ntiles);
            scan(hist, offs, nh, nullptr, sc.part);
            launch(k_rs_scatter{(const uint32_t*)lo[s], (const uint32_t*)hi[s], (const uint32_t*)val[s], n, word, shift, ntiles, (const uint32_t*)offs, lo[1 - s], hi[1 - s], val[1 - s]}, ntiles);
            s = 1 - s;
        }
        uint32_t* rank = alt->gp.as<uint32_t>();
        launch(k_zsort_rank{(const uint32_t*)val[s], n, rank}, blocks(np));
        launch(k_zsort_rename{cur->id.as<uint32_t>(), n, (const uint32_t*)rank}, blocks(np));
        if(objIndex.size() == np && rankHost.size() == np && objScratch.size() == np) {
            // the user order changed: rank[old user index] = new user index
            HIP_OK(hipMemcpyAsync(rankHost.data(), rank, np * 4, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            for(uint64_t i = 0; i < np; ++i) objScratch[rankHost[i]] = objIndex[i];
            objIndex.swap(objScratch);
        }
        HIP_OK(hipStreamSynchronize(stream));
    }

    // positions in the brick layout of the active cells, in the reference's activeCellIdx order
    std::vector<int64_t> activeOffsets() {
        std::vector<float>   ph(nC);
        std::vector<int32_t> bs(nb);
        HIP_OK(hipMemcpyAsync(ph.data(), phi.p, nC * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(bs.data(), brickSlot.p, nb * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        std::vector<int64_t> off;
        for(int k = 0; k < g.nz; ++k)
            for(int j = 0; j < g.ny; ++j)
                for(int i = 0; i < g.nx; ++i) {
                    if(!(ph[((size_t)k * g.ny + j) * g.nx + i] < 0)) continue;
                    const int32_t s = bs[((k >> 3) * g.nby + (j >> 3)) * g.nbx + (i >> 3)];
                    off.push_back(s < 0 ? -1 : (int64_t)s * 512 + ((k & 7) * 8 + (j & 7)) * 8 + (i & 7));
                }
        return off;
    }

    // z = M^-1 r on the current system, r and z in activeCellIdx order; uses the PCG's r, z, q
    void applyPreconditioner(const double* rin, double* zout) {
        const std::vector<int64_t> off = activeOffsets();
        if(!nSlots) {
            for(size_t c = 0; c < off.size(); ++c) zout[c] = 0;
            return;
        }
        std::vector<float> rv((size_t)nSlots * 512, 0.f), zv((size_t)nSlots * 512);
        for(size_t c = 0; c < off.size(); ++c)
            if(off[c] >= 0) rv[off[c]] = (float)rin[c];
        HIP_OK(hipMemcpyAsync(r.p, rv.data(), rv.size() * 4, hipMemcpyHostToDevice, stream));
        if(prm.preconditioner == BNN_PRECOND_IC0) {
            formIC0();
            applyIC0(r.as<float>(), z.as<float>(), nullptr);
        } else if(prm.preconditioner == BNN_PRECOND_MG) {
            vcycle(nullptr);  // writes z (with a single level, the coarsest solve alone does)
        } else {
            launch(k_mul_arrays{(const float*)invd.as<float>(), (const float*)r.as<float>(), z.as<float>(), rv.size()}, blocks(rv.size()));
        }
        HIP_OK(hipMemcpyAsync(zv.data(), z.p, zv.size() * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for(size_t c = 0; c < off.size(); ++c) zout[c] = off[c] >= 0 ? (double)zv[off[c]] : 0.0;
    }

    // readback of a per-cell pressure-system vector in the reference's activeCellIdx order
    template<class T> void gatherActive(const DevBuf& vec, double* dst, uint64_t* nActive) {
        std::vector<float>   ph(nC);
        std::vector<int32_t> bs(nb);
        HIP_OK(hipMemcpyAsync(ph.data(), phi.p, nC * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(bs.data(), brickSlot.p, nb * 4, hipMemcpyDeviceToHost, stream));
        std::vector<T> xv((size_t)nSlots * 512);
        if(nSlots) HIP_OK(hipMemcpyAsync(xv.data(), vec.p, xv.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        uint64_t cnt = 0;
        for(int k = 0; k < g.nz; ++k)
            for(int j = 0; j < g.ny; ++j)
                for(int i = 0; i < g.nx; ++i) {
                    if(!(ph[((size_t)k * g.ny + j) * g.nx + i] < 0)) continue;
                    if(dst) {
                        const int32_t s = bs[((k >> 3) * g.nby + (j >> 3)) * g.nbx + (i >> 3)];
                        dst[cnt]        = s < 0 ? 0.0 : (double)xv[(size_t)s * 512 + ((k & 7) * 8 + (j & 7)) * 8 + (i & 7)];
                    }
                    ++cnt;
                }
        if(nActive) *nActive = cnt;
    }

    void getGrid(int f, void* dst) {
        if(!dst) throw Error(BNN_ERR_INVALID_ARG, "null destination");
        const DevBuf* src = nullptr;
        size_t        bytes = 0;
        switch(f) {
            case BNN_FIELD_U: src = &u; bytes = nU * 4; break;
            case BNN_FIELD_V: src = &v; bytes = nV * 4; break;
            case BNN_FIELD_W: src = &w; bytes = nW * 4; break;
            case BNN_FIELD_WEIGHT_U: src = &wu; bytes = nU * 4; break;
            case BNN_FIELD_WEIGHT_V: src = &wv; bytes = nV * 4; break;
            case BNN_FIELD_WEIGHT_W: src = &ww; bytes = nW * 4; break;
            case BNN_FIELD_VALID_U: src = &vu; bytes = nU; break;
            case BNN_FIELD_VALID_V: src = &vv; bytes = nV; break;
            case BNN_FIELD_VALID_W: src = &vw; bytes = nW; break;
            case BNN_FIELD_EXTRAP_U: src = &eu; bytes = nU; break;
            case BNN_FIELD_EXTRAP_V: src = &ev_; bytes = nV; break;
            case BNN_FIELD_EXTRAP_W: src = &ew; bytes = nW; break;
            case BNN_FIELD_OLD_U: src = &uo; bytes = nU * 4; break;
            case BNN_FIELD_OLD_V: src = &vo; bytes = nV * 4; break;
            case BNN_FIELD_OLD_W: src = &wo; bytes = nW * 4; break;
            case BNN_FIELD_PHI: src = &phi; bytes = nC * 4; break;
            case BNN_FIELD_BOUNDARY_SDF: src = &bsdf; bytes = nN * 4; break;
            default: throw Error(BNN_ERR_INVALID_ARG, "unknown grid field");
        }
        HIP_OK(hipMemcpyAsync(dst, src->p, bytes, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        if(isValidField(f)) {  // any nonzero flag is valid (the extrapolation marks its faces 2 or 3)
            uint8_t* d = static_cast<uint8_t*>(dst);
            for(size_t i = 0; i < bytes; ++i) d[i] = d[i] != 0;
        }
    }
    static bool isValidField(int f) { return f == BNN_FIELD_VALID_U || f == BNN_FIELD_VALID_V || f == BNN_FIELD_VALID_W; }

    void setGrid(int f, const void* src) {
        if(!src) throw Error(BNN_ERR_INVALID_ARG, "null source");
        DevBuf* dst   = nullptr;
        size_t  bytes = 0;
        switch(f) {
            case BNN_FIELD_U: dst = &u; bytes = nU * 4; break;
            case BNN_FIELD_V: dst = &v; bytes = nV * 4; break;
            case BNN_FIELD_W: dst = &w; bytes = nW * 4; break;
            case BNN_FIELD_VALID_U: dst = &vu; bytes = nU; break;
            case BNN_FIELD_VALID_V: dst = &vv; bytes = nV; break;
            case BNN_FIELD_VALID_W: dst = &vw; bytes = nW; break;
            case BNN_FIELD_EXTRAP_U: dst = &eu; bytes = nU; break;
            case BNN_FIELD_EXTRAP_V: dst = &ev_; bytes = nV; break;
            case BNN_FIELD_EXTRAP_W: dst = &ew; bytes = nW; break;
            case BNN_FIELD_OLD_U: dst = &uo; bytes = nU * 4; break;
            case BNN_FIELD_OLD_V: dst = &vo; bytes = nV * 4; break;
            case BNN_FIELD_OLD_W: dst = &wo; bytes = nW * 4; break;
            case BNN_FIELD_PHI: dst = &phi; bytes = nC * 4; break;
            case BNN_FIELD_BOUNDARY_SDF: dst = &bsdf; bytes = nN * 4; weightsDirty = true; bsdfNonzero = true; break;
            default: throw Error(BNN_ERR_INVALID_ARG, "field cannot be set");
        }
        std::vector<uint8_t> flags;
        if(isValidField(f)) {  // stored as 0 / 1
            flags.resize(bytes);
            for(size_t i = 0; i < bytes; ++i) flags[i] = static_cast<const uint8_t*>(src)[i] != 0;
            src = flags.data();
        }
        HIP_OK(hipMemcpyAsync(dst->p, src, bytes, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    // ---------------------------------------------------------------------------------------------
    // frame output and memory states in the reference's BananaParticleData files
    float particleRadius() const { return prm.cell_size / prm.ratio_cell_size_pradius; }  // ParticleSolverData.cpp:232

    // ParticleHelpers::compress (ParticleHelpers.cpp:47-79): per-component min / max over the
    // particles, then (uint16)(65535 * ((v - min) / (max - min))) in float; the attribute's bytes are
    // the 3 minima, the 3 maxima (float) and the 3 uint16 of every particle
    static std::vector<uint8_t> compress3(const std::vector<float>& v, uint64_t n) {
        float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
        for(int d = 0; d < 3 && n; ++d) {
            mn[d] = mx[d] = v[d];
            for(uint64_t i = 1; i < n; ++i) {
                mn[d] = std::min(mn[d], v[3 * i + d]);
                mx[d] = std::max(mx[d], v[3 * i + d]);
            }
        }
        std::vector<uint8_t> out(24 + 6 * n);
        std::memcpy(out.data(), mn, 12);
        std::memcpy(out.data() + 12, mx, 12);
        uint16_t* q = reinterpret_cast<uint16_t*>(out.data() + 24);
        for(int d = 0; d < 3; ++d) {
            const float diff = mx[d] - mn[d];
            for(uint64_t i = 0; i < n; ++i) {
                // a component with max == min has no spread to encode (the reference divides by zero)
                const float t = diff > 0.f ? 65535.0f * ((v[3 * i + d] - mn[d]) / diff) : 0.f;
                q[3 * i + d]  = (uint16_t)t;
            }
        }
        return out;
    }

    std::vector<uint16_t> objectIndices() const { return objIndex.size() == np ? objIndex : std::vector<uint16_t>(np, 0); }

    // PIC_Solver::saveFrameData (PIC_Solver.cpp:232-262): particle radius, compressed positions and,
    // as the OptionalSavingData keys ask, compressed velocities and the object indices
    void saveFrameData(const std::string& path, bool withVelocity, bool withObjectIndex) {
        BpdFile f;
        f.nParticles     = np;
        const float rad  = particleRadius();
        f.addFixed<float>("particle_radius", "real", 1, &rad);
        std::vector<float> pos(3 * np), vel(withVelocity ? 3 * np : 0);
        getParticles(pos.data(), withVelocity ? vel.data() : nullptr);
        f.addRaw(false, "particle_position", "compressed_real", 4, 3, compress3(pos, np));
        if(withVelocity) f.addRaw(false, "particle_velocity", "compressed_real", 4, 3, compress3(vel, np));
        if(withObjectIndex) {
            const std::vector<uint16_t> obj = objectIndices();
            f.addFixed<uint32_t>("NObjects", "uint", 1, &nObjects);
            f.addParticle<uint16_t>("object_index", "uint16", 1, obj.data());
        }
        std::string err;
        if(!bpd_write(path, std::move(f), err)) throw Error(BNN_ERR_INVALID_ARG, err);
    }

    // PIC_Solver::saveMemoryState (PIC_Solver.cpp:206-228): grid resolution, the three face-velocity
    // grids, particle radius, object count; per particle: object index, position, velocity
    void saveMemoryState(const std::string& path) {
        BpdFile f;
        f.nParticles          = np;
        const uint32_t res[3] = {(uint32_t)g.nx, (uint32_t)g.ny, (uint32_t)g.nz};
        f.addFixed<uint32_t>("grid_resolution", "uint", 3, res);
        const DevBuf* vg[3]   = {&u, &v, &w};
        const size_t  cnt[3]  = {nU, nV, nW};
        for(int d = 0; d < 3; ++d) {
            std::vector<float> a(cnt[d]);
            HIP_OK(hipMemcpyAsync(a.data(), vg[d]->p, cnt[d] * 4, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            f.addFixed<float>("grid_velocity_" + std::to_string(d), "real", (uint32_t)cnt[d], a.data());
        }
        const float rad = particleRadius();
        f.addFixed<uint32_t>("NObjects", "uint", 1, &nObjects);
        f.addFixed<float>("particle_radius", "real", 1, &rad);
        std::vector<float> pos(3 * np), vel(3 * np);
        getParticles(pos.data(), vel.data());
        const std::vector<uint16_t> obj = objectIndices();
        f.addParticle<uint16_t>("object_index", "uint16", 1, obj.data());
        f.addParticle<float>("particle_position", "real", 3, pos.data());
        f.addParticle<float>("particle_velocity", "real", 3, vel.data());
        std::string err;
        if(!bpd_write(path, std::move(f), err)) throw Error(BNN_ERR_INVALID_ARG, err);
    }

    // PIC_Solver::loadMemoryState (PIC_Solver.cpp:158-202) from one file; `frame` becomes the
    // finished frame, as the reference sets it to the loaded state's index
    void loadMemoryState(const std::string& path, uint32_t frame) {
        BpdFile     f;
        std::string err;
        if(!bpd_read(path, f, err)) throw Error(BNN_ERR_INVALID_ARG, err);
        auto need = [&](const char* name, size_t bytes) -> const BpdAttr& {
            const BpdAttr* a = f.find(name);
            if(!a) throw Error(BNN_ERR_INVALID_ARG, path + ": no attribute " + name);
            if(a->data.size() != bytes)
                throw Error(BNN_ERR_INVALID_ARG, path + ": attribute " + name + " has " + std::to_string(a->data.size()) + " bytes, expected " +
                                                     std::to_string(bytes));
            return *a;
        };
        uint32_t res[3];
        std::memcpy(res, need("grid_resolution", 12).data.data(), 12);
        if(res[0] != (uint32_t)g.nx || res[1] != (uint32_t)g.ny || res[2] != (uint32_t)g.nz)
            throw Error(BNN_ERR_INVALID_ARG, path + ": grid resolution differs from the scene's");
        float rad;
        std::memcpy(&rad, need("particle_radius", 4).data.data(), 4);
        if(!(std::fabs(rad - particleRadius()) < 1.1920929e-07f))  // __BNN_REQUIRE_APPROX_NUMBERS(.., MEpsilon)
            throw Error(BNN_ERR_INVALID_ARG, path + ": particle radius differs from the scene's");
        uint32_t nobj;
        std::memcpy(&nobj, need("NObjects", 4).data.data(), 4);
        const uint64_t n   = f.nParticles;
        const BpdAttr& pa_ = need("particle_position", 12 * n);
        const BpdAttr& va_ = need("particle_velocity", 12 * n);
        const BpdAttr& oa_ = need("object_index", 2 * n);
        const BpdAttr* gv[3] = {&need("grid_velocity_0", nU * 4), &need("grid_velocity_1", nV * 4), &need("grid_velocity_2", nW * 4)};
        setParticles(reinterpret_cast<const float*>(pa_.data.data()), reinterpret_cast<const float*>(va_.data.data()), n);
        std::vector<uint16_t> obj(n);
        if(n) std::memcpy(obj.data(), oa_.data.data(), 2 * n);
        setObjectIndices(std::move(obj), nobj);
        setGrid(BNN_FIELD_U, gv[0]->data.data());
        setGrid(BNN_FIELD_V, gv[1]->data.data());
        setGrid(BNN_FIELD_W, gv[2]->data.data());
        finishedFrame    = frame;
        st.finished_frame = frame;
    }

    void getBins(uint32_t* permOut, uint32_t* startOut) {
        if(!binned) throw Error(BNN_ERR_STATE, "no binning for the current particle positions");
        if(permOut && np) HIP_OK(hipMemcpyAsync(permOut, perm.p, np * 4, hipMemcpyDeviceToHost, stream));
        if(startOut) HIP_OK(hipMemcpyAsync(startOut, cellStart.p, (nC + 1) * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
};

}  // namespace bnn

#include "flip_serving.hpp"

//==================================================================================================
// C ABI
//==================================================================================================
struct bnn_flip {
    std::unique_ptr<bnn::FlipSolver> s;
    std::string                      err;
};

namespace {
template<class F> bnn_status guard(bnn_flip* h, F&& f) {
    if(!h || !h->s) return BNN_ERR_INVALID_ARG;
    try {
        h->err.clear();
        return f(*h->s);
    } catch(const bnn::Error& e) {
        h->err = e.what();
        return e.code;
    } catch(const std::bad_alloc&) {
        h->err = "host out of memory";
        return BNN_ERR_OUT_OF_MEMORY;
    } catch(const std::exception& e) {
        h->err = e.what();
        return BNN_ERR_DEVICE;
    }
}
thread_local std::string g_createError;
}  // namespace

//--------------------------------------------------------------------------------------------------
// Scratch of the allocation-free z-sort (bnn_flip_sort_particles_noalloc): one array per device,
// held until the process ends. Host code never uses the name `scratch` as an address (on the host it
// is a stand-in, not device memory): each device's address comes from hipGetSymbolAddress, taken
// once with that device current.
//--------------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[BNN_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex    g_scratchMutex;
void*         g_scratchPtr[kMaxDevices] = {};

void* scratch_for_device(int dev) {
    if(dev < 0 || dev >= kMaxDevices) throw bnn::Error(BNN_ERR_DEVICE, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    if(!g_scratchPtr[dev]) {
        int cur = 0;
        HIP_OK(hipGetDevice(&cur));
        if(cur != dev) HIP_OK(hipSetDevice(dev));
        void*            ptr = nullptr;
        const hipError_t e   = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if(cur != dev) (void)hipSetDevice(cur);
        if(e != hipSuccess || !ptr)
            throw bnn::Error(BNN_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchPtr[dev] = ptr;
    }
    return g_scratchPtr[dev];
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
    p->repulsive_force_stiffness      = 50.f;  // ParticleSolverData.h:143
    p->refine_inner_tolerance         = 1e-4f;
    p->timing                         = 0;
    p->mg_sweeps                      = 4;
    p->mg_omega                       = 1.8f;
    p->mg_coarsest_sweeps             = 8;
    p->start_frame                    = 1;
    p->final_frame                    = 1;
}

bnn_status bnn_flip_create(const bnn_flip_params* p, bnn_flip** out) {
    if(!p || !out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    try {
        auto* h = new bnn_flip();
        h->s    = std::make_unique<bnn::FlipSolver>(*p);
        *out    = h;
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

bnn_status bnn_boundaries_from_json(const char* json_text, bnn_boundary_object* objs, uint32_t cap, uint32_t* n) {
    if(!json_text || !n || (cap && !objs)) return BNN_ERR_INVALID_ARG;
    try {
        const bnn::json::Value           root = bnn::json::Parser(json_text).parse();
        std::vector<bnn_boundary_object> v;
        bnn::boundaries_from_json(root, v);
        *n = (uint32_t)v.size();
        if(v.size() > cap) {
            g_createError = "more boundary objects than the array holds";
            return BNN_ERR_INVALID_ARG;
        }
        for(size_t i = 0; i < v.size(); ++i) objs[i] = v[i];
        return BNN_OK;
    } catch(const bnn::SceneLoadError& e) {
        g_createError = e.what();
        return e.code;
    } catch(const std::exception& e) {
        g_createError = e.what();
        return BNN_ERR_INVALID_ARG;
    }
}

bnn_status bnn_flip_set_boundary_objects(bnn_flip* h, const bnn_boundary_object* objs, uint32_t n) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.setBoundaryObjects(objs, n);
        return BNN_OK;
    });
}

bnn_status bnn_flip_set_finished_frame(bnn_flip* h, uint32_t frame) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.finishedFrame      = frame;
        s.st.finished_frame  = frame;
        return BNN_OK;
    });
}

bnn_status bnn_flip_create_from_json(const char* json_text, uint64_t seed, bnn_flip** out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    bnn_flip_params p;
    bnn_scene_spec  sp;
    bnn_status      st = bnn_scene_from_json(json_text, &p, &sp, nullptr, 0);
    if(st != BNN_OK) return st;
    bnn_boundary_object objs[BNN_MAX_BOUNDARY_OBJECTS];
    uint32_t            nobj = 0;
    st                       = bnn_boundaries_from_json(json_text, objs, BNN_MAX_BOUNDARY_OBJECTS, &nobj);
    if(st != BNN_OK) return st;
    st = bnn_flip_create(&p, out);
    if(st != BNN_OK) return st;
    if(nobj) st = bnn_flip_set_boundary_objects(*out, objs, nobj);
    if(st == BNN_OK) st = bnn_flip_generate_scene(*out, &sp, seed);
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
    return guard(h, [&](bnn::FlipSolver& s) {
        s.setParticles(pos, vel, n);
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
    return guard(h, [&](bnn::FlipSolver& s) {
        const uint64_t        n = bnn::generate_scene(s.prm, *spec, seed, nullptr, nullptr, 0);
        std::vector<float>    pos(n * 3), vel(n * 3);
        std::vector<uint16_t> obj(n);
        bnn::generate_scene(s.prm, *spec, seed, pos.data(), vel.data(), n, obj.data());
        s.setParticles(pos.data(), vel.data(), n);
        s.setObjectIndices(std::move(obj), (uint32_t)std::max(0, std::min(spec->n_boxes, 8)) + (uint32_t)std::max(0, std::min(spec->n_spheres, 8)));
        s.sortParticles();  // PIC_Solver::generateParticles sorts the new particles (PIC_Solver.cpp:57)
        return BNN_OK;
    });
}

uint64_t bnn_flip_num_particles(const bnn_flip* h) { return (h && h->s) ? h->s->np : 0; }

bnn_status bnn_flip_get_particles(bnn_flip* h, float* pos, float* vel) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.getParticles(pos, vel);
        return BNN_OK;
    });
}

static bnn_status solve_status(const bnn::FlipSolver& s, bool ok) {
    return (!ok && s.prm.exit_if_pressure_projection_failed) ? BNN_ERR_NOT_CONVERGED : BNN_OK;
}
static bnn_status bnn_solve_status_of(const bnn::FlipSolver& s, bool ok) { return solve_status(s, ok); }

bnn_status bnn_flip_step(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::FlipSolver& s) { return solve_status(s, run_now(s.step(dt))); });
}

bnn_status bnn_flip_advance_frame(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) { return solve_status(s, run_now(s.advanceFrame())); });
}

bnn_status bnn_flip_do_frame_simulation(bnn_flip* h, uint32_t frame) {
    (void)frame;
    return guard(h, [&](bnn::FlipSolver& s) {
        s.sortParticles();
        return solve_status(s, run_now(s.advanceFrame()));
    });
}

bnn_status bnn_flip_timestep_cfl(bnn_flip* h, float* dt) {
    if(!dt) return BNN_ERR_INVALID_ARG;
    return guard(h, [&](bnn::FlipSolver& s) {
        *dt = run_now(s.timestepCFL());
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_affine(bnn_flip* h, float* c) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!c && s.np) throw bnn::Error(BNN_ERR_INVALID_ARG, "null destination");
        s.getAffine(c);
        return BNN_OK;
    });
}

bnn_status bnn_flip_set_affine(bnn_flip* h, const float* c) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!c && s.np) throw bnn::Error(BNN_ERR_INVALID_ARG, "null source");
        s.setAffine(c);
        return BNN_OK;
    });
}

bnn_status bnn_flip_correct_particle_positions(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!(dt > 0)) throw bnn::Error(BNN_ERR_INVALID_ARG, "dt must be > 0");
        if(s.prm.correct_position) s.correctParticlePositions(dt);
        return BNN_OK;
    });
}

bnn_status bnn_flip_move_particles(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.moveParticles(dt);
        return BNN_OK;
    });
}

bnn_status bnn_flip_sort_particles_noalloc(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!s.sortDue() || !s.np) return BNN_OK;
        const bnn::ZsortLayout L = bnn::zsort_layout(s.np);
        if(L.bytes > (size_t)BNN_NOALLOC_SCRATCH_BYTES)
            throw bnn::Error(BNN_ERR_INVALID_ARG, "bnn_flip_sort_particles_noalloc: " + std::to_string(s.np) +
                                                      " particles need " + std::to_string(L.bytes) +
                                                      " bytes of scratch, more than BNN_NOALLOC_SCRATCH_BYTES (" +
                                                      std::to_string((size_t)BNN_NOALLOC_SCRATCH_BYTES) + ")");
        unsigned char* base = static_cast<unsigned char*>(scratch_for_device(s.device));
        s.zSortWith(bnn::ZsortScratch{reinterpret_cast<uint32_t*>(base + L.histOff),
                                      reinterpret_cast<unsigned long long*>(base + L.accOff),
                                      reinterpret_cast<uint32_t*>(base + L.partOff)});
        return BNN_OK;
    });
}

bnn_status bnn_flip_sort_particles(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.sortParticles();
        return BNN_OK;
    });
}

bnn_status bnn_flip_advance_velocity(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.beginStats();
        bool ok = run_now(s.advanceVelocity(dt));
        run_now(s.endStats(dt));
        return solve_status(s, ok);
    });
}

bnn_status bnn_flip_map_particles_to_grid(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.mapParticles2Grid();
        return BNN_OK;
    });
}

bnn_status bnn_flip_map_grid_to_particles(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.mapGrid2Particles();
        return BNN_OK;
    });
}

bnn_status bnn_flip_extrapolate_velocity(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.extrapolateVelocity();
        return BNN_OK;
    });
}

bnn_status bnn_flip_constrain_grid_velocity(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.constrainGridVelocity();
        return BNN_OK;
    });
}

bnn_status bnn_flip_pressure_projection(bnn_flip* h, float dt) {
    return guard(h, [&](bnn::FlipSolver& s) { return solve_status(s, run_now(s.pressureProjection(dt))); });
}

bnn_status bnn_flip_compute_fluid_sdf(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.computeFluidSDF();
        return BNN_OK;
    });
}

bnn_status bnn_flip_compute_cell_weights(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.weightsDirty = true;
        run_now(s.ensureWeights());
        return BNN_OK;
    });
}

bnn_status bnn_flip_bin_particles(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.binParticles();
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_grid(bnn_flip* h, int32_t field, void* dst) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.getGrid(field, dst);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_pressure(bnn_flip* h, double* dst, uint64_t* n_active) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!s.systemValid) throw bnn::Error(BNN_ERR_STATE, "no pressure solve yet");
        if(s.lastRefined) s.gatherActive<double>(s.x64, dst, n_active);
        else s.gatherActive<float>(s.x, dst, n_active);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_rhs(bnn_flip* h, double* dst, uint64_t* n_active) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!s.systemValid) throw bnn::Error(BNN_ERR_STATE, "no pressure system yet");
        s.gatherActive<double>(s.rhs64, dst, n_active);
        return BNN_OK;
    });
}

bnn_status bnn_flip_apply_preconditioner(bnn_flip* h, const double* r, double* z) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!r || !z) throw bnn::Error(BNN_ERR_INVALID_ARG, "null vector");
        if(!s.systemValid) throw bnn::Error(BNN_ERR_STATE, "no pressure system yet");
        s.applyPreconditioner(r, z);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_ic0_factor(bnn_flip* h, double* dst, uint64_t* n_active) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(s.prm.preconditioner != BNN_PRECOND_IC0) throw bnn::Error(BNN_ERR_STATE, "the preconditioner is not IC(0)");
        if(!s.systemValid) throw bnn::Error(BNN_ERR_STATE, "no pressure system yet");
        if(s.nSlots) s.formIC0();
        s.gatherActive<float>(s.icInv, dst, n_active);
        return BNN_OK;
    });
}

bnn_status bnn_data_file_path(const char* data_path, const char* folder, int32_t frame, char* out, size_t cap) {
    if(!data_path || !folder || !out || !cap) return BNN_ERR_INVALID_ARG;
    const std::string p = bnn::bpd_file_path(data_path, folder, frame);
    if(p.size() + 1 > cap) return BNN_ERR_INVALID_ARG;
    std::memcpy(out, p.c_str(), p.size() + 1);
    return BNN_OK;
}

int32_t bnn_latest_state_index(const char* data_path, const char* folder, uint32_t start_frame, uint32_t final_frame) {
    if(!data_path || !folder) return -1;
    auto exists = [&](int32_t i) {
        struct stat sb;
        return ::stat(bnn::bpd_file_path(data_path, folder, i).c_str(), &sb) == 0;
    };
    const int32_t last = (int32_t)start_frame - 1;  // PIC_Solver.cpp:165-167
    if(last > 1 && exists(last)) return last;
    for(int32_t i = (int32_t)final_frame; i >= 1; --i)  // DataIO::getLatestFileIndex (DataIO.h:90-98)
        if(exists(i)) return i;
    return -1;
}

bnn_status bnn_flip_save_frame_data(bnn_flip* h, const char* path, int32_t save_velocity, int32_t save_object_index) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!path) throw bnn::Error(BNN_ERR_INVALID_ARG, "null path");
        s.saveFrameData(path, save_velocity != 0, save_object_index != 0);
        return BNN_OK;
    });
}

bnn_status bnn_flip_save_memory_state(bnn_flip* h, const char* path) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!path) throw bnn::Error(BNN_ERR_INVALID_ARG, "null path");
        s.saveMemoryState(path);
        return BNN_OK;
    });
}

bnn_status bnn_flip_load_memory_state(bnn_flip* h, const char* path, uint32_t frame) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(!path) throw bnn::Error(BNN_ERR_INVALID_ARG, "null path");
        s.loadMemoryState(path, frame);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_object_index(bnn_flip* h, uint16_t* dst, uint32_t* n_objects) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(dst) {
            const std::vector<uint16_t> o = s.objectIndices();
            if(!o.empty()) std::memcpy(dst, o.data(), o.size() * 2);
        }
        if(n_objects) *n_objects = s.nObjects;
        return BNN_OK;
    });
}

void bnn_io_default_params(bnn_io_params* io) {
    if(!io) return;
    std::memset(io, 0, sizeof(*io));
    std::strcpy(io->data_path, "./SimData");  // ParticleSolverData.h:56-64
    std::strcpy(io->memory_state_folder, "MemoryState");
    std::strcpy(io->frame_data_folder, "FrameData");
    io->load_memory_state = 1;
    io->frame_per_state   = 1;
}

bnn_status bnn_io_from_json(const char* json_text, bnn_io_params* io) {
    if(!json_text || !io) return BNN_ERR_INVALID_ARG;
    bnn_io_default_params(io);
    try {
        bnn::io_from_json(bnn::json::Parser(json_text).parse(), *io);
    } catch(const bnn::SceneLoadError& e) {
        return e.code;
    } catch(const std::exception&) {
        return BNN_ERR_INVALID_ARG;
    }
    return BNN_OK;
}

bnn_status bnn_flip_set_timing(bnn_flip* h, int32_t timing) {
    return guard(h, [&](bnn::FlipSolver& s) {
        if(timing < 0 || timing > 2) throw bnn::Error(BNN_ERR_INVALID_ARG, "timing must be 0, 1 or 2");
        s.prm.timing = timing;
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_bins(bnn_flip* h, uint32_t* perm, uint32_t* cell_start) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.getBins(perm, cell_start);
        return BNN_OK;
    });
}

bnn_status bnn_flip_get_stats(bnn_flip* h, bnn_flip_stats* out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    return guard(h, [&](bnn::FlipSolver& s) {
        *out = s.st;
        return BNN_OK;
    });
}

bnn_status bnn_flip_set_grid(bnn_flip* h, int32_t field, const void* src) {
    return guard(h, [&](bnn::FlipSolver& s) {
        s.setGrid(field, src);
        return BNN_OK;
    });
}

bnn_status bnn_flip_synchronize(bnn_flip* h) {
    return guard(h, [&](bnn::FlipSolver& s) {
        HIP_OK(hipStreamSynchronize(s.stream));
        return BNN_OK;
    });
}

bnn_status bnn_flip_footprint(const bnn_flip_params* p, uint64_t n_particles, uint64_t* bytes) {
    if(!p || !bytes) return BNN_ERR_INVALID_ARG;
    *bytes = bnn::scene_footprint(*p, n_particles);
    return BNN_OK;
}

}  // extern "C"

//--------------------------------------------------------------------------------------------------
// batches: scenes stepped concurrently by a worker pool, each on its own stream
//--------------------------------------------------------------------------------------------------
struct bnn_flip_batch {
    std::unique_ptr<bnn::BatchRunner> runner;
    bnn::BatchTuner                   tuner;
    bool                              pinned = false;  // shape set by the caller: no tuning
    bnn::BatchRunner::Shape           pinnedShape;
    std::string                       err;
    int32_t                           timing = -1;  // < 0: each scene keeps its own setting
};

struct bnn_server {
    std::unique_ptr<bnn::Server> srv;
    std::string                  err;
};

namespace {
// one fused batch call: make(scene, i) gives scene i's task; every scene's status goes to status[i]
template<class F>
bnn_status batch_call(bnn_flip_batch* b, bnn_flip* const* scenes, uint32_t n, bnn_status* status, F&& make) {
    if(!b || (n && !scenes)) return BNN_ERR_INVALID_ARG;
    b->err.clear();
    std::vector<bnn::FlipSolver*> sc(n);
    for(uint32_t i = 0; i < n; ++i) {
        if(!scenes[i] || !scenes[i]->s) {
            b->err = "null scene handle";
            return BNN_ERR_INVALID_ARG;
        }
        for(uint32_t j = 0; j < i; ++j)
            if(scenes[j] == scenes[i]) {
                b->err = "a scene is listed twice";
                return BNN_ERR_INVALID_ARG;
            }
        sc[i] = scenes[i]->s.get();
        if(b->timing >= 0) sc[i]->prm.timing = b->timing;
    }
    try {
        bnn::BatchRunner&             R   = *b->runner;
        bnn::Recorder&                rec = R.recorder();
        const uint64_t                sig = bnn::batch_signature(sc);
        const bool                    tuning = !b->pinned && !b->tuner.fixed(sig);
        const bnn::BatchRunner::Shape sh     = b->pinned ? b->pinnedShape : b->tuner.next(sig);
        R.setShape(sh);
        if(tuning) {
            rec.timedFns = {bnn::p2g_launch_fn(sh.p2gSeg, false), bnn::p2g_launch_fn(sh.p2gSeg, true)};
            rec.timedMs  = 0;
        }
        std::vector<bnn::BatchRunner::Task> tasks(n);
        std::vector<bnn::BatchRunner::Task*> live;
        for(uint32_t i = 0; i < n; ++i) {
            try {
                R.start(tasks[i], sc[i], make(*sc[i], i));
                live.push_back(&tasks[i]);
            } catch(const bnn::Error& e) {
                tasks[i].done   = true;
                tasks[i].status = e.code;
                tasks[i].err    = e.what();
            }
        }
        const auto t0 = std::chrono::steady_clock::now();
        R.runAll(live);
        const double wallMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if(tuning) {
            b->tuner.report(sig, sh, rec.timedMs, wallMs);
            rec.timedFns.clear();
        }
        bnn_status worst = BNN_OK;
        for(uint32_t i = 0; i < n; ++i) {
            const bnn::BatchRunner::Task& t = tasks[i];
            bnn_status                    st = t.status;
            if(st == BNN_OK) st = bnn_solve_status_of(*sc[i], t.ok);
            else {
                scenes[i]->err = t.err;
                if(b->err.empty()) b->err = t.err;
            }
            if(status) status[i] = st;
            if(st != BNN_OK && (worst == BNN_OK || worst == BNN_ERR_NOT_CONVERGED)) worst = st;
        }
        return worst;
    } catch(const bnn::Error& e) {
        b->err = e.what();
        return e.code;
    } catch(const std::exception& e) {
        b->err = e.what();
        return BNN_ERR_DEVICE;
    }
}
}  // namespace

extern "C" {

bnn_status bnn_flip_batch_create(bnn_flip_batch** out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    try {
        auto b    = std::make_unique<bnn_flip_batch>();
        b->runner = std::make_unique<bnn::BatchRunner>();
        *out      = b.release();
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
    b->timing = timing;
    return BNN_OK;
}

bnn_status bnn_flip_batch_step(bnn_flip_batch* b, bnn_flip* const* scenes, const float* dts, uint32_t n, bnn_status* status) {
    if(n && !dts) return BNN_ERR_INVALID_ARG;
    return batch_call(b, scenes, n, status, [&](bnn::FlipSolver& s, uint32_t i) { return s.step(dts[i]); });
}

bnn_status bnn_flip_batch_advance_frame(bnn_flip_batch* b, bnn_flip* const* scenes, uint32_t n, bnn_status* status) {
    return batch_call(b, scenes, n, status, [&](bnn::FlipSolver& s, uint32_t) { return s.advanceFrame(); });
}

bnn_status bnn_flip_batch_pressure_projection(bnn_flip_batch* b, bnn_flip* const* scenes, const float* dts, uint32_t n,
                                              bnn_status* status) {
    if(n && !dts) return BNN_ERR_INVALID_ARG;
    return batch_call(b, scenes, n, status, [&](bnn::FlipSolver& s, uint32_t i) { return s.pressureProjection(dts[i]); });
}

bnn_status bnn_flip_batch_set_shape(bnn_flip_batch* b, const bnn_batch_shape* shape) {
    if(!b) return BNN_ERR_INVALID_ARG;
    if(!shape) {
        b->pinned = false;
        return BNN_OK;
    }
    if((shape->p2g_segment != 4 && shape->p2g_segment != 8 && shape->p2g_segment != 16) || shape->pcg_chunk_first < 1 ||
       shape->pcg_chunk_max < 1 || shape->pcg_chunk_max > 1024 || shape->pcg_chunk_first > 1024 || shape->items_per_group < 1 ||
       shape->items_per_group > 256)
        return BNN_ERR_INVALID_ARG;
    b->pinned                 = true;
    b->pinnedShape.p2gSeg     = shape->p2g_segment;
    b->pinnedShape.chunkFirst = shape->pcg_chunk_first;
    b->pinnedShape.chunkMax   = shape->pcg_chunk_max;
    b->pinnedShape.maxPer     = shape->items_per_group;
    b->pinnedShape.minGroups  = shape->min_groups ? shape->min_groups : 16384u;
    b->pinnedShape.graphs     = shape->graphs != 0;
    return BNN_OK;
}

bnn_status bnn_flip_batch_get_info(const bnn_flip_batch* b, bnn_batch_info* out) {
    if(!b || !out) return BNN_ERR_INVALID_ARG;
    const bnn::Recorder&           rec = const_cast<bnn_flip_batch*>(b)->runner->recorder();
    const bnn::BatchRunner::Shape& sh  = b->runner->shape();
    *out                               = bnn_batch_info{};
    out->shape.p2g_segment             = sh.p2gSeg;
    out->shape.pcg_chunk_first         = sh.chunkFirst;
    out->shape.pcg_chunk_max           = sh.chunkMax;
    out->shape.items_per_group         = sh.maxPer;
    out->shape.min_groups              = sh.minGroups;
    out->shape.graphs                  = sh.graphs ? 1 : 0;
    out->flushes                       = rec.flushes;
    out->launches                      = rec.kernelLaunches;
    out->scene_launches                = rec.mergedOps;
    out->graph_replays                 = rec.graphReplays;
    out->graph_captures                = rec.graphCaptures;
    out->mean_scenes_per_flush         = rec.flushes ? (double)rec.laneFlushes / (double)rec.flushes : 0.0;
    return BNN_OK;
}

//--------------------------------------------------------------------------------------------------
// serving
//--------------------------------------------------------------------------------------------------
bnn_status bnn_server_create(const bnn_server_config* cfg, bnn_server** out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    *out = nullptr;
    try {
        bnn_server_config c{};
        c.batching = 1;
        if(cfg) c = *cfg;
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

}  // extern "C"

namespace {
template<class F> bnn_status server_guard(bnn_server* s, F&& f) {
    if(!s || !s->srv) return BNN_ERR_INVALID_ARG;
    try {
        s->err.clear();
        return f(*s->srv);
    } catch(const bnn::Error& e) {
        s->err = e.what();
        return e.code;
    } catch(const std::exception& e) {
        s->err = e.what();
        return BNN_ERR_DEVICE;
    }
}
}  // namespace

extern "C" {

bnn_status bnn_server_add_scene(bnn_server* s, const bnn_flip_params* p, const bnn_scene_spec* spec, uint64_t seed, uint64_t* id) {
    if(!p || !spec || !id) return BNN_ERR_INVALID_ARG;
    return server_guard(s, [&](bnn::Server& v) {
        *id = v.addScene(*p, *spec, seed);
        return BNN_OK;
    });
}

bnn_status bnn_server_release_scene(bnn_server* s, uint64_t id) {
    return server_guard(s, [&](bnn::Server& v) {
        v.releaseScene(id);
        return BNN_OK;
    });
}

bnn_status bnn_server_submit(bnn_server* s, uint64_t scene, int32_t kind, float dt, uint64_t* id) {
    if(!id) return BNN_ERR_INVALID_ARG;
    return server_guard(s, [&](bnn::Server& v) {
        *id = v.submit(scene, kind, dt);
        return BNN_OK;
    });
}

bnn_status bnn_server_wait(bnn_server* s, uint64_t id, double timeout, bnn_request_result* out) {
    return server_guard(s, [&](bnn::Server& v) { return v.wait(id, timeout, out) ? BNN_OK : BNN_ERR_STATE; });
}

bnn_status bnn_server_drain(bnn_server* s, double timeout) {
    return server_guard(s, [&](bnn::Server& v) { return v.drain(timeout) ? BNN_OK : BNN_ERR_STATE; });
}

bnn_status bnn_server_get_particles(bnn_server* s, uint64_t scene, float* pos, float* vel, uint64_t* n) {
    return server_guard(s, [&](bnn::Server& v) {
        v.getParticles(scene, pos, vel, n);
        return BNN_OK;
    });
}

bnn_status bnn_server_get_stats(bnn_server* s, bnn_server_stats* out) {
    if(!out) return BNN_ERR_INVALID_ARG;
    return server_guard(s, [&](bnn::Server& v) {
        *out = v.stats();
        return BNN_OK;
    });
}

bnn_status bnn_server_reset_stats(bnn_server* s) {
    return server_guard(s, [&](bnn::Server& v) {
        v.resetStats();
        return BNN_OK;
    });
}

}  // extern "C"
