// flip_scene.hpp -- per-scene device descriptor and the flat work list that every kernel runs over.
//
// Every kernel takes the array of scene descriptors and a work list of (scene, local block) items:
// one launch covers any set of scenes of any sizes (a single scene is a list of one). A block only
// ever works on one scene, and everything a scene computes (including its reductions, whose
// partials sit at fixed per-scene offsets and are folded per scene in index order) is independent
// of which other scenes share the launch.
#pragma once

#include "../flip_device.hpp"

namespace bnn {

// PCG scalars and stop flag of one scene (PCGSolver.cpp:143-199), FP64 scalars
struct PcgState {
    double   rho, alpha, beta, tmp;
    float    tol, res, bmax, relTol;
    uint32_t iter, maxIter, done, success;
};

// FP64 defect-correction loop of one scene
struct OuterState {
    double   res, bmax, tol;
    uint32_t done, success, passes, totalIter;
};

// small per-scene control block in device memory, read back in one gather per batch
struct SceneCtrl {
    uint32_t   counters[16];  // [0] solid faces, [1] fluid cells, [3] max|u| bits, [4 + l] active bricks of level l
    PcgState   pcg;
    OuterState outer;
};
constexpr int kCtrSolid = 0, kCtrFluid = 1, kCtrMaxVel = 3, kCtrSlots = 4;

// one level of the pressure system in the brick layout; level 0 is the fine system
struct LevelDesc {
    int       nx, ny, nz, nbx, nby, nbz;
    uint32_t  nb, nSlots;
    uint32_t *flag, *scan;
    int32_t * brickSlot, *slotBrick, *nbr;
    float *   diag, *cx, *cy, *cz, *invdiag;
    float *   x, *b, *r;  // level 0: x = z, b = r, r = q of the PCG (the V-cycle computes z = M^-1 r)
};

constexpr int kMaxLevels = 10;

struct SceneDesc {
    GridGeom g;
    BoxDev   dom;
    // substep parameters
    float    dt, ratio, reflMult, gravAdd[3];
    float    gravCenter[3];
    int32_t  gravityType, applyGravity, reflect;
    uint32_t advSteps;
    int32_t  constrainOn, useX64;
    // solve settings
    int32_t  refined, jacobi;
    float    pcgRelTol, mgOmega;
    uint32_t maxIter, maxOuter;
    double   refineTol;
    int32_t  mgSweeps, coarsestSweeps;
    uint32_t redBase;  // this scene's first slot in the reduction partial arrays
    int32_t  nLevels;
    // particles (ping-pong: current = P[pcur])
    uint32_t  np, pcur;
    Particles P[2];
    uint32_t *key, *perm, *cellStart, *cellCount;
    // MAC grid
    float *   u, *v, *w, *ut, *vt, *wt, *uo, *vo, *wo, *wu, *wv, *ww, *phi, *bsdf;
    uint8_t * vu, *vv, *vw, *vtmp, *eu, *ev, *ew;
    SceneCtrl* ctrl;
    // pressure system (fine level = mg[0])
    double *  rhs64, *diag64, *cx64, *cy64, *cz64, *x64;
    float *   rhs32, *x, *r, *z, *s0, *s1, *q;
    LevelDesc mg[kMaxLevels];
};

struct WorkItem {
    uint32_t scene, local;
};

// segment of a multi-segment exclusive scan: in[0..n) -> out[0..n), sum -> *total
struct ScanSeg {
    const uint32_t* in;
    uint32_t*       out;
    uint32_t*       total;
    uint32_t        n, partBase;
};

}  // namespace bnn
