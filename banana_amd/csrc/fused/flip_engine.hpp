// flip_engine.hpp -- host side of the MI355X FLIP substep: per-scene device state (Scene) and the
// Runner that queues the stage kernels of FLIP_Solver<3,float>::advanceVelocity
// (FLIP_Solver.cpp:42-53) for any set of scenes at once.
//
// A Runner owns one HIP stream (default priority). Every stage is launched once for a selection of
// scenes: the kernels read per-scene descriptors (SceneDesc, one array per Runner) and a flat work
// list of (scene, local block) items built on the host and cached in a device arena. A single
// scene is a selection of one; a scene's results do not depend on the other scenes selected with it.
#pragma once

#include <atomic>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "flip_host.hpp"
#include "flip_kernels.hpp"

namespace bnn {

struct ParticleBufs {
    DevBuf px, py, pz, vx, vy, vz, gx, gy, gz, id;
    size_t reserve(size_t n) {
        size_t b = 0;
        for(DevBuf* d : {&px, &py, &pz, &vx, &vy, &vz, &gx, &gy, &gz, &id}) b += d->reserve(n * 4);
        return b;
    }
    Particles view() const {
        return Particles{px.as<float>(), py.as<float>(), pz.as<float>(), vx.as<float>(), vy.as<float>(),
                         vz.as<float>(), gx.as<float>(), gy.as<float>(), gz.as<float>(), id.as<uint32_t>()};
    }
};

// storage of one pressure level in the brick layout (level 0: the fine system, whose x/b/r are the
// PCG's z/r/q)
struct LevelStore {
    int      nx = 0, ny = 0, nz = 0, nbx = 0, nby = 0, nbz = 0;
    uint32_t nb = 0;
    DevBuf   flag, scan, brickSlot, slotBrick, nbr, diag, cx, cy, cz, inv, x, b, r;
};

static inline uint64_t particle_capacity(uint64_t n) { return n + n / 8 + 1024; }

// process-wide unique stamps: a Runner notices a replaced or refilled scene even at a reused address
static inline uint64_t next_generation() {
    static std::atomic<uint64_t> g{0};
    return ++g;
}

//==================================================================================================
// Scene: everything one simulation keeps resident in HBM. Pressure storage is sized for the worst
// case (every brick active), so stepping never allocates and the footprint is known up front.
//==================================================================================================
class Scene {
public:
    bnn_flip_params prm;
    GridGeom        g{};
    BoxDev          dom{};
    float           grav[3] = {0, -9.81f, 0};
    size_t          nU = 0, nV = 0, nW = 0, nC = 0, nN = 0;
    uint32_t        nb = 0;

    // particles (P[pcur] is current; binning permutes into the other buffer)
    uint64_t     np = 0, cap = 0;
    ParticleBufs P[2];
    uint32_t     pcur = 0;
    DevBuf       key, perm;
    bool         binned     = false;  // cell_start / grid coordinates valid for the current positions
    uint64_t     generation = next_generation();  // renewed when the particle count or storage changes

    // grid
    DevBuf   u, v, w, ut, vt, wt, uo, vo, wo, wu, wv, ww;
    DevBuf   vu, vv, vw, vtmp, eu, ev, ew;
    DevBuf   phi, bsdf, cellStart, cellCount, ctrl;
    bool     weightsDirty    = true;   // boundary SDF changed since the weights were computed
    bool     weightsComputed = false;  // computeCellWeights has run at least once
    bool     bsdfNonzero     = false;  // the boundary SDF has been computed or set
    bool     constrainOn     = false;
    uint32_t solidFaces      = 0;  // faces with weight < FLT_MIN in the current weights

    // pressure system and multigrid levels
    DevBuf     x, r, z, s0, s1, q, rhs32, rhs64, diag64, cx64, cy64, cz64, x64;
    LevelStore lv[kMaxLevels];
    int        nLevelsMax = 1, nLevels = 1;
    uint32_t   slots[kMaxLevels] = {};
    uint32_t   nFluid            = 0;
    bool       systemValid = false, lastRefined = false, useX64 = false;

    size_t deviceBytes = 0;

    // per-step state, statistics, frame bookkeeping (ParticleSolverData.h:43-50)
    float          curDt = 0;
    bool           lastOk = true;
    bnn_flip_stats st{};
    float          frameTime     = 0;
    uint32_t       finishedFrame = 0;

    Scene(const bnn_flip_params& p, hipStream_t stream) : prm(p) {
        validate(prm);
        const GridDims G = grid_dims(prm);
        g.nx             = G.n[0];
        g.ny             = G.n[1];
        g.nz             = G.n[2];
        g.bminx          = G.bmin[0];
        g.bminy          = G.bmin[1];
        g.bminz          = G.bmin[2];
        g.dx             = G.dx;
        g.halfdx         = 0.5f * G.dx;
        g.sdfRadius      = G.dx * (float)(1.01 * std::sqrt(3) / 2.0);  // PIC_Data.cpp:43
        g.nbx            = (g.nx + 7) / 8;
        g.nby            = (g.ny + 7) / 8;
        g.nbz            = (g.nz + 7) / 8;
        for(int d = 0; d < 3; ++d) {
            dom.lo[d] = prm.box_min[d];
            dom.hi[d] = prm.box_max[d];
        }
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
        allocate(stream);
    }

    static void validate(const bnn_flip_params& prm) {
        if(!(prm.cell_size > 0)) throw Error(BNN_ERR_INVALID_ARG, "CellSize must be > 0");
        for(int d = 0; d < 3; ++d)
            if(!(prm.box_max[d] > prm.box_min[d])) throw Error(BNN_ERR_INVALID_ARG, "SimulationDomainBox: BoxMax must exceed BoxMin");
        if(prm.advection_steps < 1) throw Error(BNN_ERR_INVALID_ARG, "AdvectionSteps must be >= 1");
        if(prm.correct_position) throw Error(BNN_ERR_UNSUPPORTED, "CorrectPosition (PIC_Solver.cpp:377-485) is not implemented");
        if(prm.gravity_type < 0 || prm.gravity_type > 3) throw Error(BNN_ERR_INVALID_ARG, "GravityType must be 0..3");
        if(prm.preconditioner != BNN_PRECOND_JACOBI && prm.preconditioner != BNN_PRECOND_MG)
            throw Error(BNN_ERR_INVALID_ARG, "unknown preconditioner");
        if(prm.pressure_mode != BNN_PRESSURE_FP32 && prm.pressure_mode != BNN_PRESSURE_REFINED)
            throw Error(BNN_ERR_INVALID_ARG, "unknown pressure mode");
        if(prm.preconditioner == BNN_PRECOND_MG && (prm.mg_sweeps < 1 || prm.mg_sweeps > 8 || prm.mg_coarsest_sweeps < 1 ||
                                                    prm.mg_coarsest_sweeps > 512 || !(prm.mg_omega > 0.f) || prm.mg_omega >= 2.f))
            throw Error(BNN_ERR_INVALID_ARG, "multigrid settings out of range");
        const GridDims G = grid_dims(prm);
        const double   C = (double)G.n[0] * G.n[1] * G.n[2];
        if(C > 1.0e9 || G.n[0] < 3 || G.n[1] < 3 || G.n[2] < 3) throw Error(BNN_ERR_INVALID_ARG, "grid size out of range");
        // the multigrid hierarchy must end in a single brick within kMaxLevels levels
        const int maxDim = 8 << (kMaxLevels - 1);
        if(prm.preconditioner == BNN_PRECOND_MG && (G.n[0] > maxDim || G.n[1] > maxDim || G.n[2] > maxDim))
            throw Error(BNN_ERR_INVALID_ARG, "grid axis too long for the multigrid preconditioner");
    }

    // multigrid level dimensions: level l+1 halves level l while any axis exceeds one brick
    static int level_dims(const bnn_flip_params& prm, int dims[kMaxLevels][3]) {
        const GridDims G = grid_dims(prm);
        int            n = 1;
        dims[0][0] = G.n[0], dims[0][1] = G.n[1], dims[0][2] = G.n[2];
        if(prm.preconditioner != BNN_PRECOND_MG) return 1;
        while(n < kMaxLevels && (dims[n - 1][0] > 8 || dims[n - 1][1] > 8 || dims[n - 1][2] > 8)) {
            for(int d = 0; d < 3; ++d) dims[n][d] = (dims[n - 1][d] + 1) / 2;
            ++n;
        }
        return n;
    }

    // device bytes a scene with these parameters and n particles holds (what allocate() reserves)
    static size_t footprint(const bnn_flip_params& prm, uint64_t n) {
        validate(prm);
        const GridDims G = grid_dims(prm);
        const size_t   nx = G.n[0], ny = G.n[1], nz = G.n[2];
        const size_t   C = nx * ny * nz, N = (nx + 1) * (ny + 1) * (nz + 1);
        const size_t   F = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
        const bool     refined = prm.pressure_mode == BNN_PRESSURE_REFINED;
        size_t         b       = 16 * F + 3 * F + F + 4 * C + 4 * N + 4 * (C + 1) + 4 * C + 256;
        int            dims[kMaxLevels][3];
        const int      L = level_dims(prm, dims);
        for(int l = 0; l < L; ++l) {
            const size_t nbl = (size_t)((dims[l][0] + 7) / 8) * ((dims[l][1] + 7) / 8) * ((dims[l][2] + 7) / 8);
            const size_t cells = nbl * 512;
            b += 4 * nbl * 4 + 24 * nbl + 5 * 4 * cells;  // tables + coefficients
            if(l == 0) b += 7 * 4 * cells + (refined ? 6 : 1) * 8 * cells;
            else b += 3 * 4 * cells;
        }
        if(n) b += (2 * 10 * 4 + 2 * 4) * particle_capacity(n);
        return b;
    }

    void allocate(hipStream_t stream) {
        nU = (size_t)(g.nx + 1) * g.ny * g.nz;
        nV = (size_t)g.nx * (g.ny + 1) * g.nz;
        nW = (size_t)g.nx * g.ny * (g.nz + 1);
        nC = (size_t)g.nx * g.ny * g.nz;
        nN = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
        nb = (uint32_t)(g.nbx * g.nby * g.nbz);
        size_t& B = deviceBytes;
        for(auto* b : {&u, &ut, &uo, &wu}) B += b->reserve(nU * 4);
        for(auto* b : {&v, &vt, &vo, &wv}) B += b->reserve(nV * 4);
        for(auto* b : {&w, &wt, &wo, &ww}) B += b->reserve(nW * 4);
        B += vu.reserve(nU) + vv.reserve(nV) + vw.reserve(nW);
        B += eu.reserve(nU) + ev.reserve(nV) + ew.reserve(nW);
        B += vtmp.reserve(nU + nV + nW);
        B += phi.reserve(nC * 4) + bsdf.reserve(nN * 4) + cellStart.reserve((nC + 1) * 4) + cellCount.reserve(nC * 4);
        B += ctrl.reserve(sizeof(SceneCtrl));
        int dims[kMaxLevels][3];
        nLevelsMax       = level_dims(prm, dims);
        const bool refined = prm.pressure_mode == BNN_PRESSURE_REFINED;
        for(int l = 0; l < nLevelsMax; ++l) {
            LevelStore& L = lv[l];
            L.nx = dims[l][0], L.ny = dims[l][1], L.nz = dims[l][2];
            L.nbx = (L.nx + 7) / 8, L.nby = (L.ny + 7) / 8, L.nbz = (L.nz + 7) / 8;
            L.nb  = (uint32_t)(L.nbx * L.nby * L.nbz);
            const size_t cells = (size_t)L.nb * 512;
            B += L.flag.reserve(L.nb * 4) + L.scan.reserve(L.nb * 4) + L.brickSlot.reserve(L.nb * 4) + L.slotBrick.reserve(L.nb * 4);
            B += L.nbr.reserve((size_t)L.nb * 24);
            for(DevBuf* b : {&L.diag, &L.cx, &L.cy, &L.cz, &L.inv}) B += b->reserve(cells * 4);
            if(l == 0) {
                for(DevBuf* b : {&x, &r, &z, &s0, &s1, &q, &rhs32}) B += b->reserve(cells * 4);
                B += rhs64.reserve(cells * 8);
                if(refined)
                    for(DevBuf* b : {&diag64, &cx64, &cy64, &cz64, &x64}) B += b->reserve(cells * 8);
            } else {
                for(DevBuf* b : {&L.x, &L.b, &L.r}) B += b->reserve(cells * 4);
            }
        }
        // zero everything the reference zero-fills at resize (PIC_Data.cpp:101-117), the counting
        // sort's counts and the coarse-level flags (both are returned to zero after every use)
        for(auto* b : {&u, &v, &w, &ut, &vt, &wt, &uo, &vo, &wo, &wu, &wv, &ww, &vu, &vv, &vw, &eu, &ev, &ew, &vtmp, &phi, &bsdf,
                       &cellStart, &cellCount, &ctrl})
            HIP_OK(hipMemsetAsync(b->p, 0, b->bytes, stream));
        for(int l = 1; l < nLevelsMax; ++l) HIP_OK(hipMemsetAsync(lv[l].flag.p, 0, lv[l].flag.bytes, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    void ensureParticles(uint64_t n) {
        if(n <= cap) return;
        const uint64_t c = particle_capacity(n);
        deviceBytes += P[0].reserve(c) + P[1].reserve(c) + key.reserve(c * 4) + perm.reserve(c * 4);
        cap = c;
    }

    // descriptor the kernels read (redBase: this scene's offset in the reduction partials)
    void fillDesc(SceneDesc& D, uint32_t redBase) const {
        std::memset(&D, 0, sizeof(D));
        D.g        = g;
        D.dom      = dom;
        D.dt       = curDt;
        D.ratio    = prm.pic_flip_ratio;
        D.reflMult = prm.boundary_reflection_multiplier;
        if(prm.gravity_type == 0) D.gravAdd[1] = 9.81f * curDt;
        if(prm.gravity_type == 1)
            for(int d = 0; d < 3; ++d) D.gravAdd[d] = grav[d] * curDt;
        for(int d = 0; d < 3; ++d) D.gravCenter[d] = prm.gravity_center[d];
        D.gravityType    = prm.gravity_type;
        D.applyGravity   = prm.apply_gravity;
        D.reflect        = prm.reflect_velocity_at_boundary;
        D.advSteps       = (uint32_t)prm.advection_steps;
        D.constrainOn    = constrainOn;
        D.useX64         = useX64;
        D.refined        = prm.pressure_mode == BNN_PRESSURE_REFINED;
        D.jacobi         = prm.preconditioner == BNN_PRECOND_JACOBI;
        D.pcgRelTol      = D.refined ? 1e-4f : prm.cg_relative_tolerance;  // refined: inner FP32 solves to 1e-4
        D.mgOmega        = prm.mg_omega;
        D.maxIter        = prm.max_cg_iteration;
        D.maxOuter       = 30;
        D.refineTol      = prm.refine_tolerance > 0 ? prm.refine_tolerance : 1e-10;
        D.mgSweeps       = prm.mg_sweeps;
        D.coarsestSweeps = prm.mg_coarsest_sweeps;
        D.redBase        = redBase;
        D.nLevels        = nLevels;
        D.np             = (uint32_t)np;
        D.pcur           = pcur;
        D.P[0]           = P[0].view();
        D.P[1]           = P[1].view();
        D.key            = key.as<uint32_t>();
        D.perm           = perm.as<uint32_t>();
        D.cellStart      = cellStart.as<uint32_t>();
        D.cellCount      = cellCount.as<uint32_t>();
        D.u = u.as<float>(), D.v = v.as<float>(), D.w = w.as<float>();
        D.ut = ut.as<float>(), D.vt = vt.as<float>(), D.wt = wt.as<float>();
        D.uo = uo.as<float>(), D.vo = vo.as<float>(), D.wo = wo.as<float>();
        D.wu = wu.as<float>(), D.wv = wv.as<float>(), D.ww = ww.as<float>();
        D.phi = phi.as<float>(), D.bsdf = bsdf.as<float>();
        D.vu = vu.as<uint8_t>(), D.vv = vv.as<uint8_t>(), D.vw = vw.as<uint8_t>(), D.vtmp = vtmp.as<uint8_t>();
        D.eu = eu.as<uint8_t>(), D.ev = ev.as<uint8_t>(), D.ew = ew.as<uint8_t>();
        D.ctrl  = ctrl.as<SceneCtrl>();
        D.rhs64 = rhs64.as<double>(), D.diag64 = diag64.as<double>(), D.cx64 = cx64.as<double>();
        D.cy64 = cy64.as<double>(), D.cz64 = cz64.as<double>(), D.x64 = x64.as<double>();
        D.rhs32 = rhs32.as<float>(), D.x = x.as<float>(), D.r = r.as<float>(), D.z = z.as<float>();
        D.s0 = s0.as<float>(), D.s1 = s1.as<float>(), D.q = q.as<float>();
        for(int l = 0; l < nLevelsMax; ++l) {
            const LevelStore& L = lv[l];
            LevelDesc&        M = D.mg[l];
            M.nx = L.nx, M.ny = L.ny, M.nz = L.nz, M.nbx = L.nbx, M.nby = L.nby, M.nbz = L.nbz;
            M.nb = L.nb, M.nSlots = slots[l];
            M.flag = L.flag.as<uint32_t>(), M.scan = L.scan.as<uint32_t>();
            M.brickSlot = L.brickSlot.as<int32_t>(), M.slotBrick = L.slotBrick.as<int32_t>(), M.nbr = L.nbr.as<int32_t>();
            M.diag = L.diag.as<float>(), M.cx = L.cx.as<float>(), M.cy = L.cy.as<float>(), M.cz = L.cz.as<float>();
            M.invdiag = L.inv.as<float>();
            if(l == 0) {
                M.x = z.as<float>(), M.b = r.as<float>(), M.r = q.as<float>();
            } else {
                M.x = L.x.as<float>(), M.b = L.b.as<float>(), M.r = L.r.as<float>();
            }
        }
    }

    //----------------------------------------------------------------------------------------------
    // host transfers (synchronous on the given stream)
    //----------------------------------------------------------------------------------------------
    void setParticles(const float* pos, const float* vel, uint64_t n, hipStream_t stream) {
        if(n > 0xFFFFFFF0ull) throw Error(BNN_ERR_INVALID_ARG, "too many particles for 32-bit particle ids");
        if(n && (!pos || !vel)) throw Error(BNN_ERR_INVALID_ARG, "null particle array");
        ensureParticles(std::max<uint64_t>(n, 1));
        std::vector<float>    tmp(n);
        std::vector<uint32_t> ids(n);
        ParticleBufs&         B      = P[pcur];
        DevBuf*               dst[6] = {&B.px, &B.py, &B.pz, &B.vx, &B.vy, &B.vz};
        for(int a = 0; a < 6; ++a) {
            const float* src = a < 3 ? pos : vel;
            for(uint64_t i = 0; i < n; ++i) tmp[i] = src[3 * i + (a % 3)];
            if(n) HIP_OK(hipMemcpyAsync(dst[a]->p, tmp.data(), n * 4, hipMemcpyHostToDevice, stream));
            HIP_OK(hipStreamSynchronize(stream));
        }
        for(uint64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
        if(n) HIP_OK(hipMemcpyAsync(B.id.p, ids.data(), n * 4, hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemsetAsync(cellCount.p, 0, cellCount.bytes, stream));
        HIP_OK(hipStreamSynchronize(stream));
        np     = n;
        binned     = false;
        generation = next_generation();
    }

    void getParticles(float* pos, float* vel, hipStream_t stream) {
        if(!np) return;
        std::vector<float>    a(np);
        std::vector<uint32_t> ids(np);
        const ParticleBufs&   B = P[pcur];
        HIP_OK(hipMemcpyAsync(ids.data(), B.id.p, np * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        const DevBuf* src[6] = {&B.px, &B.py, &B.pz, &B.vx, &B.vy, &B.vz};
        for(int k = 0; k < 6; ++k) {
            float* out = k < 3 ? pos : vel;
            if(!out) continue;
            HIP_OK(hipMemcpyAsync(a.data(), src[k]->p, np * 4, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            for(uint64_t i = 0; i < np; ++i) out[3 * (uint64_t)ids[i] + (k % 3)] = a[i];
        }
    }

    // a per-cell vector of the fine pressure system in the reference's activeCellIdx order
    // (serial numbering of phi < 0 cells, PIC_Solver.cpp:706-712)
    template<class T> void gatherActive(const DevBuf& vec, double* dst, uint64_t* nActive, hipStream_t stream) {
        std::vector<float>   ph(nC);
        std::vector<int32_t> bs(nb);
        HIP_OK(hipMemcpyAsync(ph.data(), phi.p, nC * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(bs.data(), lv[0].brickSlot.p, nb * 4, hipMemcpyDeviceToHost, stream));
        std::vector<T> xv((size_t)slots[0] * 512);
        if(slots[0]) HIP_OK(hipMemcpyAsync(xv.data(), vec.p, xv.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
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

    DevBuf* field(int f, size_t& bytes, bool forWrite) {
        switch(f) {
            case BNN_FIELD_U: bytes = nU * 4; return &u;
            case BNN_FIELD_V: bytes = nV * 4; return &v;
            case BNN_FIELD_W: bytes = nW * 4; return &w;
            case BNN_FIELD_WEIGHT_U: bytes = nU * 4; return forWrite ? nullptr : &wu;
            case BNN_FIELD_WEIGHT_V: bytes = nV * 4; return forWrite ? nullptr : &wv;
            case BNN_FIELD_WEIGHT_W: bytes = nW * 4; return forWrite ? nullptr : &ww;
            case BNN_FIELD_VALID_U: bytes = nU; return &vu;
            case BNN_FIELD_VALID_V: bytes = nV; return &vv;
            case BNN_FIELD_VALID_W: bytes = nW; return &vw;
            case BNN_FIELD_EXTRAP_U: bytes = nU; return &eu;
            case BNN_FIELD_EXTRAP_V: bytes = nV; return &ev;
            case BNN_FIELD_EXTRAP_W: bytes = nW; return &ew;
            case BNN_FIELD_OLD_U: bytes = nU * 4; return &uo;
            case BNN_FIELD_OLD_V: bytes = nV * 4; return &vo;
            case BNN_FIELD_OLD_W: bytes = nW * 4; return &wo;
            case BNN_FIELD_PHI: bytes = nC * 4; return &phi;
            case BNN_FIELD_BOUNDARY_SDF: bytes = nN * 4; return &bsdf;
            default: return nullptr;
        }
    }

    void getGrid(int f, void* dst, hipStream_t stream) {
        if(!dst) throw Error(BNN_ERR_INVALID_ARG, "null destination");
        size_t  bytes = 0;
        DevBuf* src   = field(f, bytes, false);
        if(!src) throw Error(BNN_ERR_INVALID_ARG, "unknown grid field");
        HIP_OK(hipMemcpyAsync(dst, src->p, bytes, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    void setGrid(int f, const void* src, hipStream_t stream) {
        if(!src) throw Error(BNN_ERR_INVALID_ARG, "null source");
        size_t  bytes = 0;
        DevBuf* dst   = field(f, bytes, true);
        if(!dst) throw Error(BNN_ERR_INVALID_ARG, "field cannot be set");
        if(f == BNN_FIELD_BOUNDARY_SDF) {
            weightsDirty = true;
            bsdfNonzero  = true;
        }
        HIP_OK(hipMemcpyAsync(dst->p, src, bytes, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    void getBins(uint32_t* permOut, uint32_t* startOut, hipStream_t stream) {
        if(!binned) throw Error(BNN_ERR_STATE, "no binning for the current particle positions");
        if(permOut && np) HIP_OK(hipMemcpyAsync(permOut, perm.p, np * 4, hipMemcpyDeviceToHost, stream));
        if(startOut) HIP_OK(hipMemcpyAsync(startOut, cellStart.p, (nC + 1) * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
};

//==================================================================================================
// Runner: queues the stages for a selection of scenes on one stream
//==================================================================================================
enum ListKind : int {
    LK_PART,       // blocks of 256 particles
    LK_CELL,       // blocks of 256 cells
    LK_NODE,       // blocks of 256 nodes ((nx+1)(ny+1)(nz+1))
    LK_FACE0,      // blocks of 256 u / v / w faces
    LK_FACE1,
    LK_FACE2,
    LK_FACEALL,    // blocks of u, then v, then w faces
    LK_BRICK,      // one item per brick of the fine brick grid
    LK_LB256,      // blocks of 256 bricks of level l's brick grid
    LK_SLOT,       // one item per active brick of level l
    LK_SLOT256,    // blocks of 256 active bricks of level l
    LK_SCENES,     // scene indices
    LK_SCAN_CELL,  // scan tiles over cell counts (+ segment table)
    LK_SCAN_BRICK  // scan tiles over level l's brick flags (+ segment table)
};

struct ListRef {
    const void*    items = nullptr;
    uint32_t       n     = 0;
    const ScanSeg* segs  = nullptr;
    uint32_t       nSegs = 0, parts = 0;
};

class Runner {
public:
    hipStream_t          stream = nullptr;
    std::vector<Scene*>  sc;
    std::vector<uint64_t> scGen;
    std::vector<uint32_t> redBase;
    uint32_t             redTotal = 0;
    int                  timing   = 0;

    Runner() { HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); }
    ~Runner() {
        if(stream) (void)hipStreamSynchronize(stream);
        for(hipEvent_t e : evPool) (void)hipEventDestroy(e);
        for(auto& s : stages) {
            (void)hipEventDestroy(s.ev);
            (void)hipHostFree(s.p);
        }
        if(stream) (void)hipStreamDestroy(stream);
    }
    Runner(const Runner&) = delete;
    Runner& operator=(const Runner&) = delete;

    void sync() { HIP_OK(hipStreamSynchronize(stream)); }

    //----------------------------------------------------------------------------------------------
    // binding, selections
    //----------------------------------------------------------------------------------------------
    void bind(const std::vector<Scene*>& scenes) {
        if(scenes != sc) {
            sync();
            sc = scenes;
            scGen.assign(sc.size(), 0);
            redBase.assign(sc.size(), 0);
            sels.clear();
            std::vector<uint32_t> all(sc.size());
            for(uint32_t i = 0; i < all.size(); ++i) all[i] = i;
            sels.push_back(all);
            resetLists();
        }
        refresh();
    }
    // particle counts or storage of a bound scene changed: static lists are stale
    void refresh() {
        for(size_t i = 0; i < sc.size(); ++i)
            if(scGen[i] != sc[i]->generation) {
                scGen[i] = sc[i]->generation;
                resetLists();
            }
        descDirty = true;
        if(sels.size() > 4096) {  // keep only the selection of all scenes
            sels.resize(1);
            resetLists();
        }
    }
    static constexpr int kAll = 0;
    int sel(const std::vector<uint32_t>& ids) {
        for(size_t i = 0; i < sels.size(); ++i)
            if(sels[i] == ids) return (int)i;
        sels.push_back(ids);
        return (int)sels.size() - 1;
    }
    template<class Pred> int selWhere(int base, Pred p) {
        const std::vector<uint32_t> b = sels[base];
        std::vector<uint32_t>       v;
        for(uint32_t i : b)
            if(p(*sc[i])) v.push_back(i);
        return sel(v);
    }
    const std::vector<uint32_t>& ids(int s) const { return sels[s]; }
    bool                         empty(int s) const { return sels[s].empty(); }

    //----------------------------------------------------------------------------------------------
    // descriptors, work lists, uploads, readbacks
    //----------------------------------------------------------------------------------------------
    const SceneDesc* sd() {
        if(descDirty) {
            hDesc.resize(sc.size());
            for(size_t i = 0; i < sc.size(); ++i) sc[i]->fillDesc(hDesc[i], redBase[i]);
            growSync(dDesc, hDesc.size() * sizeof(SceneDesc));
            upload(dDesc.p, hDesc.data(), hDesc.size() * sizeof(SceneDesc));
            descDirty = false;
        }
        return dDesc.as<SceneDesc>();
    }
    void markDesc() { descDirty = true; }
    void slotsChanged() {
        ++slotEpoch;
        descDirty = true;
    }

    ListRef list(int kind, int level, int s) {
        const bool     dyn = kind == LK_SLOT || kind == LK_SLOT256;
        const auto     key = std::make_tuple(kind, level, s, dyn ? slotEpoch : (uint64_t)0);
        const auto     it  = cache.find(key);
        if(it != cache.end()) return it->second;
        ListRef ref = build(kind, level, s);
        cache[key]  = ref;
        return ref;
    }

    // one item list per launch, fetched right before the launch (an arena reset invalidates earlier refs)
    template<class K, class... A> void items(K kern, int kind, int level, int s, int threads, A... args) {
        const SceneDesc* D = sd();
        const ListRef    L = list(kind, level, s);
        if(!L.n) return;
        hipLaunchKernelGGL(kern, dim3(L.n), dim3(threads), 0, stream, D, (const WorkItem*)L.items, args...);
    }
    // one thread per selected scene: kern(SD, scenes, n, args...)
    template<class K, class... A> void perScene(K kern, int s, A... args) {
        const SceneDesc* D = sd();
        const ListRef    L = list(LK_SCENES, 0, s);
        if(!L.n) return;
        hipLaunchKernelGGL(kern, dim3(blocks(L.n, 256)), dim3(256), 0, stream, D, (const uint32_t*)L.items, L.n, args...);
    }
    // one block per selected scene: kern(SD, scenes, args...)
    template<class K, class... A> void blockPerScene(K kern, int s, int threads, A... args) {
        const SceneDesc* D = sd();
        const ListRef    L = list(LK_SCENES, 0, s);
        if(!L.n) return;
        hipLaunchKernelGGL(kern, dim3(L.n), dim3(threads), 0, stream, D, (const uint32_t*)L.items, args...);
    }
    void zeroCounters(int s, int first, int count) {
        const SceneDesc* D = sd();
        const ListRef    L = list(LK_SCENES, 0, s);
        if(!L.n) return;
        hipLaunchKernelGGL(k_ctrl_zero, dim3(blocks((size_t)L.n * count, 256)), dim3(256), 0, stream, D, (const uint32_t*)L.items, L.n,
                           first, count);
    }
    // multi-segment exclusive scan (cells: counts -> cell_start; bricks: flags -> scan)
    void scan(int kind, int level, int s) {
        const ListRef L0 = list(kind, level, s);
        if(!L0.n) return;
        if(scanPart.bytes < (size_t)L0.parts * 4) growSync(scanPart, (size_t)L0.parts * 4 * 2);
        const ListRef L = list(kind, level, s);  // (growing synchronises; the list stays cached)
        uint32_t*     part = scanPart.as<uint32_t>();
        hipLaunchKernelGGL(k_scan_reduce, dim3(L.n), dim3(kBlock), 0, stream, L.segs, (const WorkItem*)L.items, part);
        hipLaunchKernelGGL(k_scan_partials, dim3(L.nSegs), dim3(kBlock), 0, stream, L.segs, part);
        hipLaunchKernelGGL(k_scan_down, dim3(L.n), dim3(kBlock), 0, stream, L.segs, (const WorkItem*)L.items, (const uint32_t*)part);
    }
    // control blocks of the selected scenes, in selection order (synchronises)
    const SceneCtrl* readCtrl(int s) {
        const uint32_t n = (uint32_t)sels[s].size();
        if(!n) return nullptr;
        const size_t bytes = (size_t)n * sizeof(SceneCtrl);
        if(ctrlOut.bytes < bytes) growSync(ctrlOut, bytes * 2);
        ctrlHost.reserve(ctrlOut.bytes);
        const SceneDesc* D = sd();
        const ListRef    L = list(LK_SCENES, 0, s);
        const uint32_t   words = (uint32_t)(bytes / 4);
        hipLaunchKernelGGL(k_ctrl_gather, dim3(blocks(words, 256)), dim3(256), 0, stream, D, (const uint32_t*)L.items, n,
                           ctrlOut.as<uint32_t>());
        HIP_OK(hipMemcpyAsync(ctrlHost.p, ctrlOut.p, bytes, hipMemcpyDeviceToHost, stream));
        sync();
        return ctrlHost.as<SceneCtrl>();
    }

    //----------------------------------------------------------------------------------------------
    // timers: HIP event pairs on the stream, read once per step (tag < BNN_NUM_STAGES: a stage;
    // tag >= 100: kernel timer tag - 100)
    //----------------------------------------------------------------------------------------------
    double   stageAcc[BNN_NUM_STAGES]     = {};
    double   ktAcc[BNN_NUM_KERNEL_TIMERS] = {};
    uint64_t ktCnt[BNN_NUM_KERNEL_TIMERS] = {};
    void     stageBegin(int s) {
        if(timing) openTimer(s);
    }
    void stageEnd(int s) {
        if(timing) closeTimer(s);
    }
    void ktBegin(int k) {
        if(timing >= 2) openTimer(100 + k);
    }
    void ktEnd(int k) {
        if(timing >= 2) closeTimer(100 + k);
    }
    void beginStats() {
        for(double& a : stageAcc) a = 0;
        for(double& a : ktAcc) a = 0;
        for(uint64_t& c : ktCnt) c = 0;
    }
    void collectTimers() {
        if(timerRecs.empty()) return;
        sync();
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

    //==============================================================================================
    // stages (each over a selection of the bound scenes)
    //==============================================================================================
    // computeBoundarySDF with the domain box (PIC_Solver.cpp:1158-1168)
    void boundarySDF(int s) {
        items(k_boundary_sdf, LK_NODE, 0, s, kBlock);
        for(uint32_t i : ids(s)) {
            sc[i]->weightsDirty = true;
            sc[i]->bsdfNonzero  = true;
        }
    }

    // PIC_Solver.cpp:353-373
    void move(int s) {
        items(k_move, LK_PART, 0, s, kBlock);
        for(uint32_t i : ids(s)) sc[i]->binned = false;
    }

    // cell-sorted binning: key + count, scan, scatter, per-cell ordering, permute
    void bin(int s) {
        items(k_cell_key, LK_PART, 0, s, kBlock);
        scan(LK_SCAN_CELL, 0, s);
        items(k_scatter, LK_PART, 0, s, kBlock);
        items(k_segsort, LK_CELL, 0, s, kBlock);
        items(k_permute, LK_PART, 0, s, kBlock);
        for(uint32_t i : ids(s)) {
            sc[i]->pcur ^= 1u;
            sc[i]->binned = true;
        }
        markDesc();
    }
    void requireBinned(int s) {
        const int nb = selWhere(s, [](const Scene& x) { return !x.binned; });
        if(!empty(nb)) bin(nb);
    }

    // FLIP_Solver.cpp:57-95
    void p2g(int s) {
        requireBinned(s);
        ktBegin(BNN_KT_P2G);
        items(k_p2g, LK_NODE, 0, s, kBlock);
        ktEnd(BNN_KT_P2G);
    }

    // PIC_Solver.cpp:985-1090: 4 layers, all components per launch
    void extrapolate(int s) {
        for(int layer = 0; layer < 4; ++layer) items(k_extrap_layer, LK_FACEALL, 0, s, kBlock, layer);
    }

    // PIC_Solver.cpp:1094-1115. Uses the weights as they stand: before the first pressure projection
    // they are the zero fill of PIC_Data.cpp:109 (FLIP_Solver.cpp:45 constrains before
    // computeCellWeights runs), so every extrapolated face qualifies then. A no-op while the boundary
    // SDF is identically zero (zero gradient) or no face has weight < FLT_MIN.
    void constrain(int s) {
        bool any = false;
        for(uint32_t i : ids(s)) {
            Scene&     x  = *sc[i];
            const bool on = x.bsdfNonzero && !(x.weightsComputed && x.solidFaces == 0);
            if(on != x.constrainOn) {
                x.constrainOn = on;
                markDesc();
            }
            any |= on;
        }
        if(!any) return;
        const int on = selWhere(s, [](const Scene& x) { return x.constrainOn; });
        for(int d = 0; d < 3; ++d) {
            items(k_constrain, LK_FACE0 + d, 0, on, kBlock, d);
            items(k_face_restore, LK_FACE0 + d, 0, on, kBlock, d);
        }
    }

    void backup(int s) { items(k_backup, LK_FACEALL, 0, s, kBlock); }
    void gravity(int s) { items(k_gravity, LK_FACEALL, 0, s, kBlock); }

    // computeCellWeights for the scenes whose boundary SDF changed (PIC_Solver.cpp:543-597)
    void weights(int s) {
        const int dirty = selWhere(s, [](const Scene& x) { return x.weightsDirty; });
        if(empty(dirty)) return;
        zeroCounters(dirty, kCtrSolid, 1);
        items(k_weights, LK_NODE, 0, dirty, kBlock);
        const SceneCtrl* c = readCtrl(dirty);
        const auto&      v = ids(dirty);
        for(size_t k = 0; k < v.size(); ++k) {
            Scene& x          = *sc[v[k]];
            x.solidFaces      = c[k].counters[kCtrSolid];
            x.weightsDirty    = false;
            x.weightsComputed = true;
        }
    }

    void fluidSDF(int s) {  // PIC_Solver.cpp:601-689
        requireBinned(s);
        ktBegin(BNN_KT_FLUID_SDF);
        items(k_fluid_sdf, LK_CELL, 0, s, kBlock);
        ktEnd(BNN_KT_FLUID_SDF);
    }

    // active bricks, slot tables, matrix-free system (PIC_Solver.cpp:693-893) and the multigrid
    // hierarchy; one readback per level for the whole selection
    void system(int s) {
        zeroCounters(s, kCtrFluid, 1);
        items(k_brick_flags, LK_BRICK, 0, s, 512);
        scan(LK_SCAN_BRICK, 0, s);
        const SceneCtrl* c = readCtrl(s);
        const auto       v = ids(s);
        for(size_t k = 0; k < v.size(); ++k) {
            Scene& x = *sc[v[k]];
            x.nFluid = c[k].counters[kCtrFluid];
            x.slots[0] = c[k].counters[kCtrSlots + 0];
            if(x.slots[0] > x.nb) throw Error(BNN_ERR_DEVICE, "active brick count exceeds the brick grid");
            x.nLevels = 1;
        }
        // reduction offsets over all bound scenes (partials of unselected scenes keep their place)
        redTotal = 0;
        for(size_t i = 0; i < sc.size(); ++i) {
            redBase[i] = redTotal;
            redTotal += std::max<uint32_t>(sc[i]->slots[0], 1u);
        }
        if(pa.bytes < (size_t)redTotal * 8) {
            growSync(pa, (size_t)redTotal * 8 * 2);
            growSync(pb, (size_t)redTotal * 8 * 2);
        }
        slotsChanged();
        items(k_level_slots, LK_LB256, 0, s, kBlock, 0);
        items(k_level_nbrs, LK_SLOT256, 0, s, kBlock, 0);
        items(k_assemble, LK_SLOT, 0, s, 512);
        // coarse levels while a level has more than one active brick and the grid more than one brick
        int cur = selWhere(s, [](const Scene& x) { return x.nLevelsMax > 1 && x.slots[0] > 1; });
        for(int l = 1; !empty(cur); ++l) {
            items(k_mg_flag_parents, LK_SLOT256, l - 1, cur, kBlock, l);
            scan(LK_SCAN_BRICK, l, cur);
            const SceneCtrl* cc = readCtrl(cur);
            const auto       cv = ids(cur);
            for(size_t k = 0; k < cv.size(); ++k) {
                Scene& x   = *sc[cv[k]];
                x.slots[l] = cc[k].counters[kCtrSlots + l];
                if(x.slots[l] == 0 || x.slots[l] > x.lv[l].nb) throw Error(BNN_ERR_DEVICE, "multigrid level has an invalid brick count");
                x.nLevels = l + 1;
            }
            slotsChanged();
            items(k_level_slots, LK_LB256, l, cur, kBlock, l);
            items(k_level_nbrs, LK_SLOT256, l, cur, kBlock, l);
            items(k_mg_coarsen, LK_SLOT, l, cur, 512, l);
            cur = selWhere(cur, [l](const Scene& x) { return x.slots[l] > 1 && l + 1 < x.nLevelsMax; });
        }
        for(uint32_t i : v) sc[i]->systemValid = true;
    }

    // z = M^-1 r on the fine level of the selected (multigrid) scenes: symmetric V-cycle, levels
    // processed in lock-step, each scene's coarsest level solved in LDS
    void vcycle(int s) {
        int maxL = 1, maxSw = 1;
        for(uint32_t i : ids(s)) {
            maxL  = std::max(maxL, sc[i]->nLevels);
            maxSw = std::max(maxSw, sc[i]->prm.mg_sweeps);
        }
        std::vector<int> down(maxL, -1);
        for(int l = 0; l + 1 < maxL; ++l) down[l] = selWhere(s, [l](const Scene& x) { return x.nLevels - 1 > l; });
        for(int l = 0; l + 1 < maxL; ++l) {
            const int D = down[l];
            if(empty(D)) break;
            for(int sw = 0; sw < maxSw; ++sw) {
                items(k_mg_smooth, LK_SLOT, l, D, 512, l, 0, sw == 0 ? 1 : 0, sw);
                items(k_mg_smooth, LK_SLOT, l, D, 512, l, 1, 0, sw);
            }
            items(k_mg_residual, LK_SLOT, l, D, 512, l);
            items(k_mg_restrict, LK_SLOT, l + 1, D, 512, l);
        }
        blockPerScene(k_mg_coarsest, s, 512);
        for(int l = maxL - 2; l >= 0; --l) {
            const int D = down[l];
            if(empty(D)) continue;
            items(k_mg_prolong, LK_SLOT, l, D, 512, l);
            for(int sw = 0; sw < maxSw; ++sw) {
                items(k_mg_smooth, LK_SLOT, l, D, 512, l, 1, 0, sw);
                items(k_mg_smooth, LK_SLOT, l, D, 512, l, 0, 0, sw);
            }
        }
    }

    // FP32 PCG on rhs32 -> x for the selected scenes (PCGSolver.cpp:143-199)
    void pcg(int s) {
        double*   dpa = pa.as<double>();
        double*   dpb = pb.as<double>();
        perScene(k_pcg_reset, s);
        items(k_pcg_init, LK_SLOT, 0, s, 512, dpa, dpb);
        const int mg0 = selWhere(s, [](const Scene& x) { return x.prm.preconditioner == BNN_PRECOND_MG; });
        if(!empty(mg0)) {
            vcycle(mg0);
            items(k_pcg_dot_zr, LK_SLOT, 0, mg0, 512, dpa);
        }
        blockPerScene(k_pcg_init_fin, s, 1024, (const double*)dpa, (const double*)dpb);
        uint32_t chunk = 4, launched = 0, parity = 0, maxIter = 0;
        for(uint32_t i : ids(s)) maxIter = std::max<uint32_t>(maxIter, sc[i]->prm.max_cg_iteration);
        int active = s;
        for(;;) {
            const SceneCtrl* c = readCtrl(active);
            const auto       v = ids(active);
            std::vector<uint32_t> still;
            uint32_t              left = 1;
            for(size_t k = 0; k < v.size(); ++k)
                if(!c[k].pcg.done) {
                    still.push_back(v[k]);
                    left = std::max(left, c[k].pcg.maxIter > c[k].pcg.iter ? c[k].pcg.maxIter - c[k].pcg.iter : 1u);
                }
            if(still.empty()) break;
            active       = sel(still);
            const int mg = selWhere(active, [](const Scene& x) { return x.prm.preconditioner == BNN_PRECOND_MG; });
            const uint32_t n = std::min(chunk, left);  // never queue more iterations than can remain
            for(uint32_t it = 0; it < n; ++it) {
                ktBegin(BNN_KT_PCG_SPMV);
                items(k_pcg_spmv, LK_SLOT, 0, active, 512, (int)parity, dpa);
                ktEnd(BNN_KT_PCG_SPMV);
                blockPerScene(k_pcg_alpha, active, 1024, (const double*)dpa);
                ktBegin(BNN_KT_PCG_UPDATE);
                items(k_pcg_update, LK_SLOT, 0, active, 512, (int)parity, dpa, dpb);
                ktEnd(BNN_KT_PCG_UPDATE);
                if(!empty(mg)) {
                    vcycle(mg);
                    items(k_pcg_dot_zr, LK_SLOT, 0, mg, 512, dpa);
                }
                blockPerScene(k_pcg_beta, active, 1024, (const double*)dpa, (const double*)dpb);
                parity ^= 1u;
            }
            launched += n;
            chunk = std::min<uint32_t>(chunk * 2, 64);
            if(launched > maxIter + 256) throw Error(BNN_ERR_DEVICE, "PCG device state stopped advancing");
        }
    }

    // PIC_Solver.cpp:897-907. FP32 scenes: one PCG. Refined scenes: FP64 defect correction around
    // FP32 PCG (r = b - A x in FP64, solve A d = r in FP32, x += d) until max|r| <= tol max|b|.
    void solve(int s) {
        const auto v = ids(s);
        for(uint32_t i : v) {
            Scene& x      = *sc[i];
            x.lastRefined = x.prm.pressure_mode == BNN_PRESSURE_REFINED;
            if(x.slots[0] == 0) {
                x.st.pcg_iterations   = 0;
                x.st.outer_iterations = 0;
                x.st.pcg_residual     = 0;
                x.st.rhs_max          = 0;
                x.st.pcg_success      = 1;
                x.lastOk              = true;
            }
        }
        const int live = selWhere(s, [](const Scene& x) { return x.slots[0] > 0; });
        if(empty(live)) return;
        const int ref = selWhere(live, [](const Scene& x) { return x.lastRefined; });
        const int f32 = selWhere(live, [](const Scene& x) { return !x.lastRefined; });
        double*   dpa = pa.as<double>();
        double*   dpb = pb.as<double>();
        perScene(k_solve_begin, live);
        int pcgSel = f32;
        if(!empty(ref)) {
            items(k_residual64, LK_SLOT, 0, ref, 512, 1, dpa, dpb);
            blockPerScene(k_residual64_fin, ref, 1024, (const double*)dpa, (const double*)dpb, 1);
            const SceneCtrl* c = readCtrl(ref);
            std::vector<uint32_t> go = ids(f32);
            const auto            rv = ids(ref);
            for(size_t k = 0; k < rv.size(); ++k)
                if(!c[k].outer.done) go.push_back(rv[k]);
            std::sort(go.begin(), go.end());
            pcgSel = sel(go);
        }
        while(!empty(pcgSel)) {
            pcg(pcgSel);
            const int refActive = selWhere(pcgSel, [](const Scene& x) { return x.lastRefined; });
            if(empty(refActive)) break;
            items(k_accumulate64, LK_SLOT, 0, refActive, 512);
            perScene(k_outer_next, refActive);
            items(k_residual64, LK_SLOT, 0, refActive, 512, 0, dpa, dpb);
            blockPerScene(k_residual64_fin, refActive, 1024, (const double*)dpa, (const double*)dpb, 0);
            const SceneCtrl*      c  = readCtrl(refActive);
            const auto            rv = ids(refActive);
            std::vector<uint32_t> go;
            for(size_t k = 0; k < rv.size(); ++k)
                if(!c[k].outer.done) go.push_back(rv[k]);
            pcgSel = sel(go);
        }
        const SceneCtrl* c  = readCtrl(live);
        const auto       lv = ids(live);
        for(size_t k = 0; k < lv.size(); ++k) {
            Scene& x = *sc[lv[k]];
            if(x.lastRefined) {
                x.st.pcg_iterations   = c[k].outer.totalIter;
                x.st.outer_iterations = c[k].outer.passes;
                x.st.pcg_residual     = c[k].outer.res;
                x.st.rhs_max          = c[k].outer.bmax;
                x.st.pcg_success      = c[k].outer.success;
            } else {
                x.st.pcg_iterations   = c[k].pcg.iter;
                x.st.outer_iterations = 1;
                x.st.pcg_residual     = c[k].pcg.res;
                x.st.rhs_max          = c[k].pcg.bmax;
                x.st.pcg_success      = c[k].pcg.success;
            }
            x.lastOk = x.st.pcg_success != 0;
        }
    }

    void updateVelocity(int s) {  // PIC_Solver.cpp:911-981
        for(uint32_t i : ids(s)) {
            Scene&     x  = *sc[i];
            const bool u6 = x.lastRefined && x.slots[0] > 0;
            if(u6 != x.useX64) {
                x.useX64 = u6;
                markDesc();
            }
        }
        items(k_update_velocity, LK_FACEALL, 0, s, kBlock);
    }

    void pressureProjection(int s) {  // PIC_Solver.cpp:533-539
        stageBegin(BNN_STAGE_WEIGHTS);
        weights(s);
        stageEnd(BNN_STAGE_WEIGHTS);
        stageBegin(BNN_STAGE_PHI);
        fluidSDF(s);
        stageEnd(BNN_STAGE_PHI);
        stageBegin(BNN_STAGE_ASSEMBLY);
        system(s);
        stageEnd(BNN_STAGE_ASSEMBLY);
        stageBegin(BNN_STAGE_PCG);
        solve(s);
        stageEnd(BNN_STAGE_PCG);
        stageBegin(BNN_STAGE_UPDATE);
        updateVelocity(s);
        stageEnd(BNN_STAGE_UPDATE);
    }

    void g2p(int s) {  // FLIP_Solver.cpp:99-133
        requireBinned(s);
        items(k_g2p, LK_PART, 0, s, kBlock);
    }

    void advanceVelocity(int s) {  // FLIP_Solver.cpp:42-53
        stageBegin(BNN_STAGE_BIN);
        requireBinned(s);
        stageEnd(BNN_STAGE_BIN);
        stageBegin(BNN_STAGE_P2G);
        p2g(s);
        stageEnd(BNN_STAGE_P2G);
        stageBegin(BNN_STAGE_EXTRAPOLATE);
        extrapolate(s);
        stageEnd(BNN_STAGE_EXTRAPOLATE);
        stageBegin(BNN_STAGE_CONSTRAIN);
        constrain(s);
        backup(s);
        gravity(s);
        stageEnd(BNN_STAGE_CONSTRAIN);
        pressureProjection(s);
        stageBegin(BNN_STAGE_EXTRAPOLATE);
        extrapolate(s);
        stageEnd(BNN_STAGE_EXTRAPOLATE);
        stageBegin(BNN_STAGE_CONSTRAIN);
        constrain(s);
        stageEnd(BNN_STAGE_CONSTRAIN);
        stageBegin(BNN_STAGE_G2P);
        g2p(s);
        stageEnd(BNN_STAGE_G2P);
    }

    void endStats(int s) {
        sync();
        collectTimers();
        for(uint32_t i : ids(s)) {
            Scene&          x  = *sc[i];
            bnn_flip_stats& st = x.st;
            for(int k = 0; k < BNN_NUM_STAGES; ++k) st.stage_ms[k] = timing ? stageAcc[k] : 0.0;
            for(int k = 0; k < BNN_NUM_KERNEL_TIMERS; ++k) {
                st.kernel_ms[k]       = ktAcc[k];
                st.kernel_launches[k] = ktCnt[k];
            }
            // algorithmic (compulsory) bytes per stage
            const double Np = (double)x.np, C = (double)x.nC, F = (double)(x.nU + x.nV + x.nW), V = (double)x.nN,
                         A = (double)x.nFluid;
            st.stage_bytes[BNN_STAGE_MOVE]        = 24 * Np + 4 * F;
            st.stage_bytes[BNN_STAGE_BIN]         = 72 * Np + 12 * C;
            st.stage_bytes[BNN_STAGE_P2G]         = 24 * Np + 5 * F;
            st.stage_bytes[BNN_STAGE_EXTRAPOLATE] = 2 * 11 * F;
            st.stage_bytes[BNN_STAGE_CONSTRAIN]   = 8 * F + 8 * F / 3;
            st.stage_bytes[BNN_STAGE_WEIGHTS]     = 0;
            st.stage_bytes[BNN_STAGE_PHI]         = 12 * Np + 13 * C + 4 * V;
            st.stage_bytes[BNN_STAGE_ASSEMBLY]    = 9 * C + 4 * C + 8 * F + 24 * A;
            st.stage_bytes[BNN_STAGE_PCG]         = 64 * A * st.pcg_iterations;
            st.stage_bytes[BNN_STAGE_UPDATE]      = 4 * C + 4 * A + 13 * F;
            st.stage_bytes[BNN_STAGE_G2P]         = 36 * Np + 8 * F + 12 * F;
            st.stage_bytes[BNN_STAGE_OTHER]       = 0;
            st.num_particles                      = x.np;
            st.num_cells                          = x.nC;
            st.num_active_cells                   = x.nFluid;
            st.num_active_bricks                  = x.slots[0];
            st.last_substep                       = x.curDt;
        }
    }

    // one substep of every selected scene, each with its own curDt (PIC_Solver.cpp:292-294)
    void step(int s) {
        for(uint32_t i : ids(s))
            if(!(sc[i]->curDt > 0)) throw Error(BNN_ERR_INVALID_ARG, "dt must be > 0");
        markDesc();
        beginStats();
        stageBegin(BNN_STAGE_MOVE);
        move(s);
        stageEnd(BNN_STAGE_MOVE);
        stageBegin(BNN_STAGE_BIN);
        bin(s);
        stageEnd(BNN_STAGE_BIN);
        advanceVelocity(s);
        endStats(s);
        for(uint32_t i : ids(s)) ++sc[i]->st.steps;
    }

    // timestepCFL (PIC_Solver.cpp:342-349) of every selected scene; result in out[k] (selection order)
    void timestepCFL(int s, std::vector<float>& out) {
        zeroCounters(s, kCtrMaxVel, 1);
        items(k_maxabs, LK_FACEALL, 0, s, kBlock);
        const SceneCtrl* c = readCtrl(s);
        const auto       v = ids(s);
        out.resize(v.size());
        for(size_t k = 0; k < v.size(); ++k) {
            const Scene& x = *sc[v[k]];
            float        m;
            std::memcpy(&m, &c[k].counters[kCtrMaxVel], 4);
            const float maxVel = std::max(kFltMin, m);
            const float ts     = maxVel > kFltMin ? x.prm.cfl_factor * (x.g.dx / maxVel) : 3.40282347e+38f;
            out[k]             = std::min(std::max(ts, x.prm.min_timestep), x.prm.max_timestep);
        }
    }

    // advanceFrame of every selected scene (PIC_Solver.cpp:267-311): CFL substeps fitted to each
    // scene's frame; scenes drop out of the lock-step loop as their frame completes
    void advanceFrame(int s) {
        const auto                 v = ids(s);
        std::map<uint32_t, uint32_t> count;
        std::map<uint32_t, bool>     ok;
        for(uint32_t i : v) {
            sc[i]->frameTime = 0;
            count[i]         = 0;
            ok[i]            = true;
        }
        int                active = selWhere(s, [](const Scene& x) { return x.frameTime < x.prm.frame_duration; });
        std::vector<float> ts;
        while(!empty(active)) {
            timestepCFL(active, ts);
            const auto av = ids(active);
            for(size_t k = 0; k < av.size(); ++k) {
                Scene&      x   = *sc[av[k]];
                const float fd  = x.prm.frame_duration;
                float       sub = ts[k];
                const float rem = fd - x.frameTime;
                if(x.frameTime + sub >= fd) sub = rem;
                else if(x.frameTime + 1.5f * sub >= fd) sub = rem * 0.5f;
                x.curDt = sub;
            }
            step(active);
            for(uint32_t i : av) {
                Scene& x = *sc[i];
                ok[i]    = ok[i] && x.lastOk;
                x.frameTime += x.curDt;
                ++count[i];
            }
            active = selWhere(active, [](const Scene& x) { return x.frameTime < x.prm.frame_duration; });
        }
        for(uint32_t i : v) {
            Scene& x = *sc[i];
            ++x.finishedFrame;
            x.st.finished_frame = x.finishedFrame;
            x.st.substep_count  = count[i];
            x.lastOk            = ok[i];
        }
    }

private:
    std::vector<std::vector<uint32_t>> sels;
    std::vector<SceneDesc>             hDesc;
    DevBuf                             dDesc, arena, scanPart, pa, pb, ctrlOut;
    PinnedBuf                          ctrlHost;
    size_t                             arenaUsed = 0;
    bool                               descDirty = true;
    uint64_t                           slotEpoch = 1;
    std::map<std::tuple<int, int, int, uint64_t>, ListRef> cache;

    // pinned staging blocks for host->device uploads, reused once their copy has completed
    struct Stage {
        void*      p     = nullptr;
        size_t     bytes = 0;
        hipEvent_t ev    = nullptr;
        bool       used  = false;
    };
    std::vector<Stage> stages;

    struct TimerRec {
        int    tag;
        size_t b, e;
    };
    std::vector<hipEvent_t> evPool;
    size_t                  evUsed = 0;
    std::vector<TimerRec>   timerRecs;

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

    // grow a runner-owned buffer after the queued work that may use it has drained
    void growSync(DevBuf& b, size_t bytes) {
        if(bytes <= b.bytes && b.p) return;
        sync();
        b.reserve(bytes);
    }

    void upload(void* dst, const void* src, size_t bytes) {
        Stage* st = nullptr;
        for(auto& s : stages)
            if(s.bytes >= bytes && (!s.used || hipEventQuery(s.ev) == hipSuccess)) {
                st = &s;
                break;
            }
        if(!st) {
            Stage s;
            s.bytes = std::max<size_t>(bytes, 1 << 16);
            HIP_OK(hipHostMalloc(&s.p, s.bytes, hipHostMallocDefault));
            HIP_OK(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
            stages.push_back(s);
            st = &stages.back();
        }
        std::memcpy(st->p, src, bytes);
        HIP_OK(hipMemcpyAsync(dst, st->p, bytes, hipMemcpyHostToDevice, stream));
        HIP_OK(hipEventRecord(st->ev, stream));
        st->used = true;
    }

    void resetLists() {
        cache.clear();
        arenaUsed = 0;
    }

    // device space for one list; an exhausted arena is reused from the start (the uploads that
    // overwrite it are ordered after every queued kernel on this stream)
    void* arenaAlloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        if(bytes > arena.bytes) {
            growSync(arena, std::max<size_t>(bytes * 2, (size_t)32 << 20));
            resetLists();
        } else if(arenaUsed + bytes > arena.bytes) {
            resetLists();
        }
        void* p = static_cast<char*>(arena.p) + arenaUsed;
        arenaUsed += bytes;
        return p;
    }

    uint32_t count(int kind, int level, const Scene& x) const {
        switch(kind) {
            case LK_PART: return blocks(x.np, kBlock);
            case LK_CELL: return blocks(x.nC, kBlock);
            case LK_NODE: return blocks(x.nN, kBlock);
            case LK_FACE0: return blocks(x.nU, kBlock);
            case LK_FACE1: return blocks(x.nV, kBlock);
            case LK_FACE2: return blocks(x.nW, kBlock);
            case LK_FACEALL: return blocks(x.nU, kBlock) + blocks(x.nV, kBlock) + blocks(x.nW, kBlock);
            case LK_BRICK: return x.nb;
            case LK_LB256: return level < x.nLevelsMax ? blocks(x.lv[level].nb, kBlock) : 0;
            case LK_SLOT: return level < x.nLevels ? x.slots[level] : 0;
            case LK_SLOT256: return level < x.nLevels ? blocks(x.slots[level], kBlock) : 0;
            default: return 0;
        }
    }

    ListRef build(int kind, int level, int s) {
        const std::vector<uint32_t>& v = sels[s];
        ListRef                      ref;
        if(kind == LK_SCENES) {
            ref.n     = (uint32_t)v.size();
            void* dst = arenaAlloc(std::max<size_t>(v.size(), 1) * 4);
            if(!v.empty()) upload(dst, v.data(), v.size() * 4);
            ref.items = dst;
            return ref;
        }
        if(kind == LK_SCAN_CELL || kind == LK_SCAN_BRICK) {
            std::vector<ScanSeg>  segs;
            std::vector<WorkItem> tiles;
            uint32_t              parts = 0;
            for(uint32_t pos = 0; pos < v.size(); ++pos) {
                const Scene& x = *sc[v[pos]];
                ScanSeg      sg{};
                if(kind == LK_SCAN_CELL) {
                    if(level != 0) continue;
                    sg.in    = x.cellCount.as<uint32_t>();
                    sg.out   = x.cellStart.as<uint32_t>();
                    sg.total = x.cellStart.as<uint32_t>() + x.nC;
                    sg.n     = (uint32_t)x.nC;
                } else {
                    if(level >= x.nLevelsMax) continue;
                    sg.in    = x.lv[level].flag.as<uint32_t>();
                    sg.out   = x.lv[level].scan.as<uint32_t>();
                    sg.total = &x.ctrl.as<SceneCtrl>()->counters[kCtrSlots + level];
                    sg.n     = x.lv[level].nb;
                }
                sg.partBase        = parts;
                const uint32_t nt  = blocks(sg.n, kScanTile);
                const uint32_t idx = (uint32_t)segs.size();
                for(uint32_t t = 0; t < nt; ++t) tiles.push_back(WorkItem{idx, t});
                parts += nt;
                segs.push_back(sg);
            }
            const size_t segBytes = ((segs.size() * sizeof(ScanSeg)) + 255) & ~(size_t)255;
            const size_t bytes    = segBytes + tiles.size() * sizeof(WorkItem);
            std::vector<char> host(std::max<size_t>(bytes, 1));
            if(!segs.empty()) std::memcpy(host.data(), segs.data(), segs.size() * sizeof(ScanSeg));
            if(!tiles.empty()) std::memcpy(host.data() + segBytes, tiles.data(), tiles.size() * sizeof(WorkItem));
            char* dst = static_cast<char*>(arenaAlloc(host.size()));
            if(bytes) upload(dst, host.data(), bytes);
            ref.segs  = reinterpret_cast<const ScanSeg*>(dst);
            ref.nSegs = (uint32_t)segs.size();
            ref.items = dst + segBytes;
            ref.n     = (uint32_t)tiles.size();
            ref.parts = parts;
            return ref;
        }
        std::vector<WorkItem> it;
        size_t                total = 0;
        for(uint32_t i : v) total += count(kind, level, *sc[i]);
        it.reserve(total);
        for(uint32_t i : v) {
            const uint32_t c = count(kind, level, *sc[i]);
            for(uint32_t t = 0; t < c; ++t) it.push_back(WorkItem{i, t});
        }
        ref.n     = (uint32_t)it.size();
        void* dst = arenaAlloc(std::max<size_t>(it.size(), 1) * sizeof(WorkItem));
        if(!it.empty()) upload(dst, it.data(), it.size() * sizeof(WorkItem));
        ref.items = dst;
        return ref;
    }
};

}  // namespace bnn
