/*
 * bnn_flip.h -- C ABI of the MI355X (gfx950) FLIP substep: a drop-in for the hot path of
 * Banana's FLIP_Solver<3,float> (Simulation/ParticleSolvers/PICFluid/FLIP_Solver.h:30-60,
 * PIC_Solver.h:40-100). Plain C, opaque handles, plain pointers and sizes.
 *
 * Conventions
 *  - Every call returns a bnn_status; BNN_OK == 0. On error the handle keeps a message readable
 *    with bnn_flip_last_error(). Nothing here exits the process (the reference calls exit() from
 *    __BNN_REQUIRE, Macros.h:225-236, and on a failed solve when ExitIfPressureProjectionFailed,
 *    PIC_Solver.cpp:901-906; this ABI reports BNN_ERR_NOT_CONVERGED instead).
 *  - Host arrays are AoS float triples (12 B per particle), like ParticleSimulationData::positions
 *    (ParticleSolverData.h:197). Particle readback is always in the order of the last
 *    bnn_flip_set_particles / bnn_flip_generate_scene call ("user order"), whatever order the
 *    device keeps internally, unless bnn_flip_sort_particles reordered it (reference z-sort).
 *  - A handle is used by one host thread at a time; different handles may be used concurrently.
 *  - Grid arrays are x-fastest (Array.h:276): index ((k*ny + j)*nx + i) with the array's own dims:
 *    u (nx+1, ny, nz), v (nx, ny+1, nz), w (nx, ny, nz+1), cell fields (nx, ny, nz),
 *    boundary SDF nodes (nx+1, ny+1, nz+1).
 */
#ifndef BNN_FLIP_H
#define BNN_FLIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bnn_status {
    BNN_OK                = 0,
    BNN_ERR_INVALID_ARG   = 1,
    BNN_ERR_NOT_CONVERGED = 2, /* PCG hit MaxCGIteration or broke down (PCGSolver.cpp:164-199) */
    BNN_ERR_DEVICE        = 3, /* HIP runtime error */
    BNN_ERR_OUT_OF_MEMORY = 4,
    BNN_ERR_UNSUPPORTED   = 5, /* a reference option this build does not implement */
    BNN_ERR_STATE         = 6, /* call out of order (e.g. step before particles exist) */
    BNN_ERR_REJECTED      = 7  /* serving: admission refused (HBM budget) */
} bnn_status;

/* Pressure solve precision (extension; the reference is FP32 throughout, Setup.h:58). */
typedef enum bnn_pressure_mode {
    BNN_PRESSURE_FP32    = 0, /* FP32 vectors, FP64 dot products and alpha/beta, stop on
                                 max|r| < CGRelativeTolerance * max|rhs| (PCGSolver.cpp:154-186) */
    BNN_PRESSURE_REFINED = 1  /* FP32 PCG inner solves under an FP64 defect-correction loop, stop
                                 on max|b - A x| <= refine_tolerance * max|b| (FP64 residual) */
} bnn_pressure_mode;

/* Particle-grid transfer scheme (the reference's solver class: FLIP_3DSolver, APIC_3DSolver,
 * AFLIP_3DSolver; FLIP_Solver.cpp:57-133, APIC_Solver.cpp:42-125, AFLIP_Solver.cpp:43-140). */
typedef enum bnn_transfer_scheme {
    BNN_TRANSFER_FLIP  = 0, /* FLIP blend of PIC and the grid change */
    BNN_TRANSFER_APIC  = 1, /* affine PIC: each particle carries C (the velocity gradient), splats
                               w (v + C (x_node - x_p)), gathers v and C; no FLIP blend and no
                               extrapolation / constrain before the solve */
    BNN_TRANSFER_AFLIP = 2  /* APIC's transfers with FLIP's step and blend, for v and C */
} bnn_transfer_scheme;

typedef enum bnn_preconditioner {
    BNN_PRECOND_JACOBI = 0, /* PCGSolver::JACOBI (PCGSolver.cpp:301-313) */
    BNN_PRECOND_MG     = 1, /* geometric multigrid V-cycle (extension, see the design notes) */
    BNN_PRECOND_IC0    = 2  /* PCGSolver::MICCL0_SYMMETRIC, the reference's default: natural-order
                               IC(0) with the 0.25 pivot safeguard (PCGSolver.cpp:398-466, 243-269),
                               level-scheduled over the 3n-2 hyperplanes i+j+k; latency-bound, for
                               comparing iterates with the reference, not for speed */
} bnn_preconditioner;

/* Parameters: field names follow the reference JSON keys (ParticleSolverData.cpp:33-78, 148-211,
 * PIC_Data.cpp:33-38); defaults from bnn_flip_default_params() equal Constants.h:31-45 and the
 * member initialisers of ParticleSolverData.h:38-175. */
typedef struct bnn_flip_params {
    float    box_min[3], box_max[3];        /* SimulationDomainBox.BoxMin/BoxMax (required) */
    float    cell_size;                     /* CellSize (1/64) */
    float    ratio_cell_size_pradius;       /* RatioCellSizePRadius (2) */
    uint32_t n_expand_cells;                /* NExpandCells (2) */
    float    min_timestep, max_timestep;    /* MinTimestep (1e-6), MaxTimestep (1/30) */
    float    cfl_factor;                    /* CFLFactor (1) */
    float    frame_duration;                /* FrameDuration (1/30) */
    float    cg_relative_tolerance;         /* CGRelativeTolerance (1e-15) */
    uint32_t max_cg_iteration;              /* MaxCGIteration (5000) */
    float    pic_flip_ratio;                /* PIC_FLIP_ratio (0.97) */
    uint32_t advection_steps;               /* AdvectionSteps (1) */
    int32_t  gravity_type;                  /* GravityType: 0 Earth, 1 Directional, 2 ToCenter, 3 FromCenter */
    float    gravity_direction[3];          /* GravityDirection */
    float    gravity_center[3];             /* GravityCenter */
    int32_t  apply_gravity;                 /* ApplyGravity (true) */
    int32_t  reflect_velocity_at_boundary;  /* ReflectVelocityAtBoundary (false) */
    float    boundary_reflection_multiplier;/* BoundaryReflectionMultiplier (1) */
    int32_t  boundary_sdf_at_init;          /* extension: compute the boundary SDF once at scene
                                               creation (false = reference behaviour, walls act
                                               only through particle projection) */
    int32_t  correct_position;              /* CorrectPosition (false): correctParticlePositions after
                                               each move (PIC_Solver.cpp:377-485) */
    int32_t  exit_if_pressure_projection_failed; /* ExitIfPressureProjectionFailed: when set, a
                                               failed solve makes step() return
                                               BNN_ERR_NOT_CONVERGED instead of BNN_OK */
    int32_t  enable_sort_particle;          /* EnableSortParticle (false) */
    uint32_t sort_frequency;                /* SortFrequency (10) */
    /* extensions */
    int32_t  pressure_mode;                 /* bnn_pressure_mode (FP32) */
    double   refine_tolerance;              /* REFINED outer tolerance (1e-10) */
    int32_t  preconditioner;                /* bnn_preconditioner (MG) */
    int32_t  timing;                        /* 1: record per-stage HIP events (stats.stage_ms),
                                               2: also per-launch events of the timed kernels */
    int32_t  mg_sweeps;                     /* MG: red-black sweeps inside each 8^3 brick (neighbour
                                               bricks held) before and after the coarse correction
                                               on each level (4) */
    float    mg_omega;                      /* MG: scale of the coarse-grid correction (1.8; the
                                               piecewise-constant coarse space under-corrects) */
    int32_t  mg_coarsest_sweeps;            /* MG: symmetric sweeps on the coarsest level (8) */
    uint32_t start_frame, final_frame;      /* GlobalParameters StartFrame / FinalFrame (1, 1) */
    float    repulsive_force_stiffness;     /* RepulsiveForceStiffness (50) */
    uint64_t random_seed;                   /* extension: seed of correctParticlePositions' random
                                               push for coincident pairs (the reference's generator is
                                               unseeded); step s uses a hash of (seed, s) */
    float    refine_inner_tolerance;        /* extension, REFINED: relative residual each inner FP32
                                               PCG solve reaches before the next FP64 correction */
    int32_t  transfer_scheme;               /* bnn_transfer_scheme (FLIP); set from GlobalParameters
                                               Solver by bnn_scene_from_json */
    int32_t  warm_start;                    /* extension, REFINED (0): start the FP64 defect correction
                                               from the previous substep's pressure (same cells, where
                                               still fluid) instead of 0 (the reference's PCG starts at
                                               0, PCGSolver.cpp:147-149); same stopping rule */
} bnn_flip_params;

/* --- boundary objects ----------------------------------------------------------------------
 * The scene's AdditionalBoundaryObjects (ParticleSolver.cpp:242-298): box or sphere geometry (a
 * sphere is the unit sphere, a box spans box_min .. box_max, both in object space) under
 * translate * rotate(axis, angle) * uniform scale, optionally animated by key frames (Animation.h:
 * cubic splines with zero end slopes, or linear, per channel; periodic from start_frame). The fluid
 * stays inside every boundary object (signed distance positive inside). Static objects and the
 * domain box act as one CSG union, each animated object on its own; animated objects move at every
 * substep of bnn_flip_advance_frame after the first frame (PIC_Solver.cpp:279-281), and the
 * boundary SDF is then recomputed from all objects (PIC_Solver.cpp:103-115,1158-1168). */
#define BNN_GEOMETRY_BOX 0
#define BNN_GEOMETRY_SPHERE 1
#define BNN_MAX_KEYFRAMES 16
#define BNN_MAX_BOUNDARY_OBJECTS 7
typedef struct bnn_keyframe {
    uint32_t frame;
    float    translation[3];
    float    rotation_axis_angle[4]; /* axis, angle in degrees (RotationAxisAngle) */
    float    scale;
    int32_t  has_rotation, has_scale; /* the key frame sets RotationAxisAngle / Scale */
} bnn_keyframe;
typedef struct bnn_boundary_object {
    int32_t      geometry;               /* BNN_GEOMETRY_BOX or BNN_GEOMETRY_SPHERE */
    float        box_min[3], box_max[3]; /* BoxMin / BoxMax (box) */
    float        translation[3];         /* Translation */
    float        rotation_axis_angle[4]; /* RotationAxisAngle, degrees */
    float        scale;                  /* Scale */
    int32_t      has_translation, has_rotation, has_scale;
    uint32_t     n_keyframes;            /* Animation.KeyFrames (0: static object) */
    bnn_keyframe keyframes[BNN_MAX_KEYFRAMES];
    int32_t      periodic;               /* Animation.Periodic */
    uint32_t     start_frame;            /* Animation.StartFrame */
    int32_t      cubic_translation, cubic_rotation, cubic_scale; /* Animation.CubicInterpolation* (1) */
} bnn_boundary_object;

/* Seeded scene generator input (replaces ParticleGenerator, ParticleGenerator.hpp:87-136): one
 * particle per sub-cell of the grid's sub-cell lattice (ppa^3 per cell) whose centre lies inside an
 * object, jittered uniformly (fraction `jitter`) inside its sub-cell; objects in order, a sub-cell
 * claimed by an earlier object is skipped; optional N(0, vel_sigma^2) velocity per component. */
typedef struct bnn_scene_spec {
    int32_t n_boxes;
    float   box_min[8][3], box_max[8][3], box_vel[8][3];
    int32_t n_spheres;
    float   sphere_center[8][3], sphere_radius[8], sphere_vel[8][3];
    float   vel_sigma;
    int32_t ppa;
    float   jitter;
} bnn_scene_spec;

/* Grid fields for bnn_flip_get_grid (element type in brackets). */
typedef enum bnn_field {
    BNN_FIELD_U = 0, BNN_FIELD_V = 1, BNN_FIELD_W = 2,                     /* [float] velocities */
    BNN_FIELD_WEIGHT_U = 3, BNN_FIELD_WEIGHT_V = 4, BNN_FIELD_WEIGHT_W = 5, /* [float] open fraction */
    BNN_FIELD_VALID_U = 6, BNN_FIELD_VALID_V = 7, BNN_FIELD_VALID_W = 8,   /* [int8] */
    BNN_FIELD_EXTRAP_U = 9, BNN_FIELD_EXTRAP_V = 10, BNN_FIELD_EXTRAP_W = 11, /* [int8] */
    BNN_FIELD_OLD_U = 12, BNN_FIELD_OLD_V = 13, BNN_FIELD_OLD_W = 14,     /* [float] backup */
    BNN_FIELD_PHI = 15,          /* [float] fluid SDF; label fluid <=> phi < 0 */
    BNN_FIELD_BOUNDARY_SDF = 16  /* [float] node SDF of the solid boundary */
} bnn_field;

#define BNN_NUM_STAGES 12
/* stage order in bnn_flip_stats.stage_ms / stage_bytes */
typedef enum bnn_stage {
    BNN_STAGE_MOVE = 0,        /* moveParticles (RK2 + boundary projection) */
    BNN_STAGE_BIN = 1,         /* cell keys, counting sort, payload permute */
    BNN_STAGE_P2G = 2,         /* mapParticles2Grid */
    BNN_STAGE_EXTRAPOLATE = 3, /* both extrapolateVelocity calls */
    BNN_STAGE_CONSTRAIN = 4,   /* both constrainGridVelocity calls + backup + gravity */
    BNN_STAGE_WEIGHTS = 5,     /* computeCellWeights (only when the boundary changed) */
    BNN_STAGE_PHI = 6,         /* computeFluidSDF + labels */
    BNN_STAGE_ASSEMBLY = 7,    /* active bricks + computeSystem */
    BNN_STAGE_PCG = 8,         /* solveSystem */
    BNN_STAGE_UPDATE = 9,      /* updateProjectedVelocity */
    BNN_STAGE_G2P = 10,        /* computeChangesGridVelocity + mapGrid2Particles */
    BNN_STAGE_OTHER = 11
} bnn_stage;

/* per-kernel timers (timing >= 2): a HIP event pair on the solver stream around every launch */
#define BNN_NUM_KERNEL_TIMERS 13
typedef enum bnn_kernel_timer {
    BNN_KT_PCG_SPMV    = 0,  /* s = z + beta s; q = A s; s.q */
    BNN_KT_PCG_UPDATE  = 1,  /* x, r updates; max|r| (Jacobi: also z = M^-1 r and z.r; with the
                                multigrid preconditioner over several levels, the same launch also
                                runs the finest pre-smoother, and z.r comes from the V-cycle) */
    BNN_KT_P2G         = 2,
    BNN_KT_FLUID_SDF   = 3,
    BNN_KT_G2P         = 4,
    BNN_KT_MOVE        = 5,
    BNN_KT_MG_PRE      = 6,  /* multigrid pre-smoother, finest level (the first V-cycle of a solve) */
    BNN_KT_MG_RESTRICT = 7,  /* residual + restriction, finest level */
    BNN_KT_MG_PROLONG  = 8,  /* coarse-grid correction, finest level */
    BNN_KT_MG_POST     = 9,  /* post-smoother (and z.r partials), finest level */
    BNN_KT_MG_COARSE   = 10, /* every V-cycle launch on the coarser levels */
    BNN_KT_EXTRAP      = 11, /* extrapolation, one launch per face component (all four layers) */
    BNN_KT_PERMUTE     = 12  /* binning payload permute */
} bnn_kernel_timer;

typedef struct bnn_flip_stats {
    uint64_t num_particles;
    uint64_t num_cells;            /* nx*ny*nz */
    uint64_t num_active_cells;     /* fluid cells (rows of the pressure system) */
    uint64_t num_active_bricks;    /* 8^3 bricks holding at least one fluid cell */
    uint32_t pcg_iterations;       /* total inner iterations of the last solve */
    uint32_t outer_iterations;     /* REFINED: defect-correction passes of the last solve */
    double   pcg_residual;         /* last max|r| (FP32 mode) or max|b - A x| (REFINED) */
    double   rhs_max;              /* max|rhs| of the last solve */
    int32_t  pcg_success;
    float    last_substep;         /* substep of the last step / frame */
    uint32_t substep_count;        /* substeps of the last advance_frame */
    double   stage_ms[BNN_NUM_STAGES];    /* last step, HIP events on the solver stream (timing=1) */
    double   stage_bytes[BNN_NUM_STAGES]; /* algorithmic bytes of the last step per stage */
    uint64_t steps;                /* steps since creation */
    uint32_t finished_frame;       /* GlobalParameters::finishedFrame (ParticleSolverData.h:50) */
    double   kernel_ms[BNN_NUM_KERNEL_TIMERS];       /* summed launch durations, last step */
    uint64_t kernel_launches[BNN_NUM_KERNEL_TIMERS]; /* launches timed, last step */
} bnn_flip_stats;

typedef struct bnn_flip bnn_flip;

/* replaces the scene's additional boundary objects (at most BNN_MAX_BOUNDARY_OBJECTS); with
 * boundary_sdf_at_init the boundary SDF is recomputed from the domain box and the objects */
bnn_status bnn_flip_set_boundary_objects(bnn_flip* h, const bnn_boundary_object* objs, uint32_t n);
/* sets GlobalParameters::finishedFrame (ParticleSolverData.h:50), for a host that loaded a memory
 * state with its own readers and handed the arrays over: the frame count the sort rule
 * (PIC_Solver.cpp:315-328) and the animated boundary objects (PIC_Solver.cpp:70-118) follow */
bnn_status bnn_flip_set_finished_frame(bnn_flip* h, uint32_t frame);
/* the AdditionalBoundaryObjects of a scene file (GeometryType Box / Sphere, Translation,
 * RotationAxisAngle, Scale, Animation); *n receives their count, at most cap are written.
 * bnn_flip_create_from_json applies them. */
bnn_status bnn_boundaries_from_json(const char* json_text, bnn_boundary_object* objs, uint32_t cap, uint32_t* n);

/* --- lifecycle (ParticleSolver::loadScene, ParticleSolver.cpp:30-168) ------------------------ */
void        bnn_flip_default_params(bnn_flip_params* p);
bnn_status  bnn_flip_create(const bnn_flip_params* p, bnn_flip** out);
/* scene JSON in the reference format (GlobalParameters / SimulationParameters /
 * ParticleGenerators with Box or Sphere geometry), ParticleSolver.cpp:30-168: creates the solver
 * and generates the particles with the seeded generator */
bnn_status  bnn_flip_create_from_json(const char* json_text, uint64_t seed, bnn_flip** out);
/* host-only: parse a scene JSON into parameters (defaults first) and a generator spec; warnings
 * (unknown keys, ignored IO options) are written to `warnings` (may be NULL) */
bnn_status  bnn_scene_from_json(const char* json_text, bnn_flip_params* p, bnn_scene_spec* spec, char* warnings,
                                size_t warnings_cap);
bnn_status  bnn_flip_destroy(bnn_flip* h);
const char* bnn_flip_last_error(const bnn_flip* h);
bnn_status  bnn_flip_get_params(const bnn_flip* h, bnn_flip_params* p);
bnn_status  bnn_flip_grid_dims(const bnn_flip* h, int32_t n[3]);
/* grid origin bMin (domain grown by NExpandCells) and cell size */
bnn_status  bnn_flip_grid_origin(const bnn_flip* h, float bmin[3], float* cell_size);

/* --- particles (PIC_Data::addParticles, PIC_Data.cpp:60-72) ------------------------------------ */
bnn_status  bnn_flip_set_particles(bnn_flip* h, const float* pos_xyz, const float* vel_xyz, uint64_t n);
bnn_status  bnn_flip_generate_scene(bnn_flip* h, const bnn_scene_spec* spec, uint64_t seed);
/* host-only generator: writes up to `cap` particles, returns the count in *n (pos/vel may be NULL) */
bnn_status  bnn_generate_scene_host(const bnn_flip_params* p, const bnn_scene_spec* spec, uint64_t seed,
                                    float* pos_xyz, float* vel_xyz, uint64_t cap, uint64_t* n);
uint64_t    bnn_flip_num_particles(const bnn_flip* h);
bnn_status  bnn_flip_get_particles(bnn_flip* h, float* pos_xyz, float* vel_xyz);

/* --- stepping (PIC_Solver::advanceFrame, PIC_Solver.cpp:267-311) ------------------------------ */
/* one body of the substep loop with substep = dt: moveParticles -> advanceVelocity */
bnn_status  bnn_flip_step(bnn_flip* h, float dt);
/* the full advanceFrame: CFL timestep, fit-to-frame, substeps, finishedFrame += 1 */
bnn_status  bnn_flip_advance_frame(bnn_flip* h);
/* ParticleSolver::doFrameSimulation (ParticleSolver.cpp:200-219): sortParticles + advanceFrame */
bnn_status  bnn_flip_do_frame_simulation(bnn_flip* h, uint32_t frame);
/* PIC_Solver::timestepCFL (PIC_Solver.cpp:342-349) */
bnn_status  bnn_flip_timestep_cfl(bnn_flip* h, float* dt);

/* --- override points of the reference (PIC_Solver.h:64-70, FLIP_Solver.h:49-52) ------------- */
bnn_status  bnn_flip_move_particles(bnn_flip* h, float dt);          /* PIC_Solver.cpp:353-373 */
/* PIC_Solver::correctParticlePositions (PIC_Solver.cpp:377-485) with the handle's seed for this step
 * (random_seed and the step count); runs only when CorrectPosition is set, like the reference */
bnn_status  bnn_flip_correct_particle_positions(bnn_flip* h, float dt);
/* APIC / AFLIP: the particles' affine matrices C (ParticleData::C, APIC_Data.h), 9 floats per particle
 * in the caller's order, row d = the gradient of velocity component d; new particles start at 0 */
bnn_status  bnn_flip_get_affine(bnn_flip* h, float* c);
bnn_status  bnn_flip_set_affine(bnn_flip* h, const float* c);
/* PIC_Solver::sortParticles (PIC_Solver.cpp:315-328): with EnableSortParticle, at finishedFrame 0
 * and whenever (finishedFrame + 1) % SortFrequency == 0, reorders the particles' user order by the
 * Morton key of NeighborSearch::z_sort (NeighborSearch.cpp:45-86), equal keys keeping their previous
 * order; otherwise does nothing. bnn_flip_generate_scene sorts new particles the same way
 * (PIC_Solver.cpp:57); bnn_flip_set_particles does not. Only the readback order changes. */
bnn_status  bnn_flip_sort_particles(bnn_flip* h);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the same z-sort as
 * bnn_flip_sort_particles (same rule, same result), with its scratch (digit histograms, key
 * accumulators, scan partials) taken from a static per-device array of BNN_NOALLOC_SCRATCH_BYTES
 * instead of buffers the handle allocates; calls no hipMalloc / hipFree. Limit: 32 MiB of scratch,
 * i.e. at most 67,088,384 particles; beyond that it returns BNN_ERR_INVALID_ARG. All calls on one
 * device share the array, so calls that overlap in time on one device (different handles, different
 * threads) race: serialise them. */
#define BNN_NOALLOC_SCRATCH_BYTES (32u << 20)
bnn_status  bnn_flip_sort_particles_noalloc(bnn_flip* h);
bnn_status  bnn_flip_advance_velocity(bnn_flip* h, float dt);      /* FLIP_Solver.cpp:42-53 */
bnn_status  bnn_flip_map_particles_to_grid(bnn_flip* h);              /* FLIP_Solver.cpp:57-95 */
bnn_status  bnn_flip_map_grid_to_particles(bnn_flip* h);              /* FLIP_Solver.cpp:99-111 */
bnn_status  bnn_flip_extrapolate_velocity(bnn_flip* h);               /* PIC_Solver.cpp:985-1090 */
bnn_status  bnn_flip_constrain_grid_velocity(bnn_flip* h);            /* PIC_Solver.cpp:1094-1115 */
bnn_status  bnn_flip_pressure_projection(bnn_flip* h, float dt);     /* PIC_Solver.cpp:533-539 */
bnn_status  bnn_flip_compute_fluid_sdf(bnn_flip* h);                  /* PIC_Solver.cpp:601-689 */
bnn_status  bnn_flip_compute_cell_weights(bnn_flip* h);               /* PIC_Solver.cpp:543-597 */
/* cell-sorted binning of the current particles (Grid::collectIndexToCells +
 * getParticleIdxSortedByCell, Grid.cpp:92-114, 283-309); permutes the device particle arrays */
bnn_status  bnn_flip_bin_particles(bnn_flip* h);

/* --- readback ------------------------------------------------------------------------------- */
bnn_status  bnn_flip_get_grid(bnn_flip* h, int32_t field, void* dst);
/* pressure per fluid cell in the reference's activeCellIdx order (linear x-fastest numbering of
 * phi < 0 cells, PIC_Solver.cpp:706-712); *n_active receives the count; dst may be NULL */
bnn_status  bnn_flip_get_pressure(bnn_flip* h, double* dst, uint64_t* n_active);
/* right-hand side of the last assembled system, same order */
bnn_status  bnn_flip_get_rhs(bnn_flip* h, double* dst, uint64_t* n_active);
/* z = M^-1 r with the handle's preconditioner on the last assembled system (PCGSolver.cpp:221-269),
 * r and z in the order of bnn_flip_get_pressure (n_active entries, rounded to FP32 on the way in).
 * Uses the solver's work vectors; the pressure is left as it is. */
bnn_status  bnn_flip_apply_preconditioner(bnn_flip* h, const double* r, double* z);
/* BNN_PRECOND_IC0: 1/sqrt of the modified pivots of the last factorisation (the reference's
 * icInvDiag after formPreconditioner), same order */
bnn_status  bnn_flip_get_ic0_factor(bnn_flip* h, double* dst, uint64_t* n_active);
/* last binning: perm[i] = index (in the array order before the sort) of sorted slot i,
 * cell_start[c] = first sorted slot of cell c (ncells+1 entries). Either may be NULL. */
bnn_status  bnn_flip_get_bins(bnn_flip* h, uint32_t* perm, uint32_t* cell_start);
bnn_status  bnn_flip_get_stats(bnn_flip* h, bnn_flip_stats* s);
/* overwrite a grid field (same element types as bnn_flip_get_grid) -- the device-side analogue of
 * restoring grid velocities from a memory state (PIC_Solver.cpp:157-201); setting
 * BNN_FIELD_BOUNDARY_SDF recomputes the face weights before their next use */
bnn_status  bnn_flip_set_grid(bnn_flip* h, int32_t field, const void* src);
/* wait for all queued device work of this handle */
bnn_status  bnn_flip_synchronize(bnn_flip* h);
/* --- frame output and memory states (BananaParticleData files, ParticleSerialization.cpp:533-830) --
 * bnn_data_file_path: DataIO::getFilePath (DataIO.h:119-124), "<data_path>/<folder>/frame.%04d.bnn".
 * bnn_flip_save_frame_data: PIC_Solver::saveFrameData (PIC_Solver.cpp:232-262): particle_radius and
 *   the positions compressed to uint16 per component (ParticleHelpers.cpp:47-79); with
 *   save_velocity also the compressed velocities, with save_object_index NObjects and object_index
 *   (the "ParticleVelocity" / "ObjectIndex" entries of OptionalSavingData). Missing folders are created.
 * bnn_flip_save_memory_state / bnn_flip_load_memory_state: PIC_Solver::saveMemoryState /
 *   loadMemoryState (PIC_Solver.cpp:158-229): grid resolution, the u/v/w grids, particle radius,
 *   object count, and per particle the object index, position and velocity, uncompressed. Loading
 *   checks the grid resolution and particle radius against the scene, replaces particles and face
 *   velocities, and sets the finished frame to `frame`.
 * bnn_latest_state_index: the state index loadMemoryState picks: start_frame - 1 when that file
 *   exists and start_frame - 1 > 1, else the highest index in [1, final_frame] with a file, else -1. */
bnn_status  bnn_data_file_path(const char* data_path, const char* folder, int32_t frame, char* out, size_t cap);
int32_t     bnn_latest_state_index(const char* data_path, const char* folder, uint32_t start_frame, uint32_t final_frame);
bnn_status  bnn_flip_save_frame_data(bnn_flip* h, const char* path, int32_t save_velocity, int32_t save_object_index);
bnn_status  bnn_flip_save_memory_state(bnn_flip* h, const char* path);
bnn_status  bnn_flip_load_memory_state(bnn_flip* h, const char* path, uint32_t frame);
bnn_status  bnn_flip_get_object_index(bnn_flip* h, uint16_t* dst, uint32_t* n_objects);

/* File IO keys of GlobalParameters (ParticleSolverData.h:56-64, ParticleSolverData.cpp:48-56) with
 * the reference's defaults; bnn_io_from_json reads them from a scene. */
typedef struct bnn_io_params {
    char     data_path[512];              /* DataPath ("./SimData") */
    char     memory_state_folder[128];    /* MemoryStateDataFolder ("MemoryState") */
    char     frame_data_folder[128];      /* FrameDataFolder ("FrameData") */
    int32_t  load_memory_state;           /* LoadMemoryState (true) */
    int32_t  save_memory_state;           /* SaveMemoryState (false) */
    int32_t  save_frame_data;             /* SaveFrameData (false) */
    uint32_t frame_per_state;             /* FramePerState (1) */
    int32_t  save_velocity;               /* OptionalSavingData has "ParticleVelocity" */
    int32_t  save_object_index;           /* OptionalSavingData has "ObjectIndex" */
} bnn_io_params;
void        bnn_io_default_params(bnn_io_params* io);
bnn_status  bnn_io_from_json(const char* json_text, bnn_io_params* io);

/* changes bnn_flip_params.timing of a live handle (0: no events, 1: stage events, 2: also per-kernel
 * events); takes effect at the next step */
bnn_status  bnn_flip_set_timing(bnn_flip* h, int32_t timing);
/* device bytes a scene with these parameters and n_particles keeps resident in HBM (what
 * bnn_flip_create + bnn_flip_set_particles allocate; stepping allocates nothing per scene) */
bnn_status  bnn_flip_footprint(const bnn_flip_params* p, uint64_t n_particles, uint64_t* bytes);

/* --- batching --------------------------------------------------------------------------------
 * Extension with no counterpart in the reference (FLIP_Solver steps one scene per solver object;
 * ParticleSolver.cpp:182-219). For callers that step many independent scenes of any sizes: every
 * stage kernel of the listed scenes runs as one launch over a flat work list of (scene, block)
 * items, all on one stream; the host reads device values (brick counts, stop flags) for all scenes
 * with one copy and one wait; launch sequences that repeat (PCG iterations) are replayed from
 * hipGraphs. A scene's blocks compute exactly what they compute alone, so each scene's results are
 * bit-identical to stepping it alone, in any batch. Launch shapes that cannot change results (the
 * APIC / AFLIP P2G column segment, PCG iterations per stop-flag poll, scene blocks per workgroup,
 * graphs on or off)
 * are chosen by an autotuner over the first calls of each batch size, or pinned with
 * bnn_flip_batch_set_shape. The scenes stay owned by their handles; a handle must not be used
 * elsewhere while a batch call runs on it. */
typedef struct bnn_flip_batch bnn_flip_batch;
bnn_status  bnn_flip_batch_create(bnn_flip_batch** out);
bnn_status  bnn_flip_batch_destroy(bnn_flip_batch* b);
const char* bnn_flip_batch_last_error(const bnn_flip_batch* b);
/* applies bnn_flip_params.timing to every scene of later batch calls (< 0: keep each scene's own) */
bnn_status  bnn_flip_batch_set_timing(bnn_flip_batch* b, int32_t timing);
/* one substep of scenes[i] with dts[i], i < n; status[i] (may be NULL) receives each scene's own
 * status (BNN_ERR_NOT_CONVERGED as bnn_flip_step would return it) */
bnn_status  bnn_flip_batch_step(bnn_flip_batch* b, bnn_flip* const* scenes, const float* dts, uint32_t n, bnn_status* status);
/* advanceFrame of every scene: CFL substeps per scene, scenes leave the batch as their frame ends */
bnn_status  bnn_flip_batch_advance_frame(bnn_flip_batch* b, bnn_flip* const* scenes, uint32_t n, bnn_status* status);
/* pressureProjection (PIC_Solver.cpp:533-539) of every scene on its current grid velocities */
bnn_status  bnn_flip_batch_pressure_projection(bnn_flip_batch* b, bnn_flip* const* scenes, const float* dts, uint32_t n,
                                               bnn_status* status);
typedef struct bnn_batch_shape {
    int32_t  p2g_segment;     /* z nodes per thread of the APIC / AFLIP P2G gather: 4, 8 or 16 */
    uint32_t pcg_chunk_first; /* PCG iterations queued before the first stop-flag poll */
    uint32_t pcg_chunk_max;   /* ... and at most per poll after it (later chunks: the iterations a geometric
                                 fit of max|r| so far says remain) */
    uint32_t items_per_group; /* scene blocks one workgroup of a batched launch runs in turn, at most */
    uint32_t min_groups;      /* ... while keeping at least this many workgroups (0: 16384) */
    int32_t  graphs;          /* replay repeated launch sequences from hipGraphs */
} bnn_batch_shape;
/* pins the launch shape of later calls (NULL: back to the autotuner) */
bnn_status bnn_flip_batch_set_shape(bnn_flip_batch* b, const bnn_batch_shape* shape);
typedef struct bnn_batch_info {
    bnn_batch_shape shape;                 /* the shape of the last call */
    uint64_t        flushes;               /* host waits (each reads back every scene's values at once) */
    uint64_t        launches;              /* kernel launches issued */
    uint64_t        scene_launches;        /* per-scene launches they stand for */
    uint64_t        graph_replays, graph_captures;
    double          mean_scenes_per_flush; /* scenes with work in a flush, on average */
} bnn_batch_info;
/* counters since the batch was created */
bnn_status bnn_flip_batch_get_info(const bnn_flip_batch* b, bnn_batch_info* out);

/* --- serving ---------------------------------------------------------------------------------
 * Extension with no counterpart in the reference (its Simulator drives one solver in a loop,
 * Simulator.cpp:42-69). A request loop for step and pressure-solve requests on resident scenes:
 * scenes are admitted while the sum of their footprints (bnn_flip_footprint) fits the HBM budget
 * (later scenes wait in arrival order, smaller ones may pass). With batching, one engine thread runs
 * every request in flight as part of one fused batch (see batching above): each round it admits the
 * oldest ready requests (a scene's requests run one at a time, in submission order), advances every
 * request to its next device readback with merged launches and one wait, and completes the finished
 * ones, so a request that arrives while others run joins them at the next round. Each result is the
 * same bits as the request run alone. */
typedef struct bnn_server bnn_server;
typedef enum bnn_request_kind {
    BNN_REQ_STEP     = 0, /* moveParticles + advanceVelocity with the request's dt */
    BNN_REQ_PRESSURE = 1  /* pressureProjection with the request's dt */
} bnn_request_kind;
typedef struct bnn_server_config {
    uint64_t hbm_budget_bytes; /* admission cap on the summed resident footprint (0: 200e9) */
    uint32_t max_batch;        /* batching: requests in flight at once, at most (0: 1024) */
    uint32_t workers;          /* threads that create admitted scenes (0: 4); without batching 1 */
    int32_t  batching;         /* 1 (default with a NULL config): continuous batching of all requests in
                                  flight into fused launches; 0: one request at a time, alone */
    int32_t  barrier_rounds;   /* batching: engine rounds a request waits at a phase boundary for the
                                  others to reach theirs, at most (0: none; aligned phases merge
                                  better, unaligned ones finish sooner) */
    uint32_t max_batch_mcells; /* batching: admit no further request while the requests in flight hold
                                  this many million grid cells (0: no limit); the oldest ready
                                  request is always admitted when none is in flight */
    int32_t  reserved[3];
} bnn_server_config;
typedef struct bnn_request_result {
    uint64_t   request_id, scene_id;
    int32_t    kind;
    bnn_status status;
    double     t_submit, t_start, t_done; /* seconds since the server was created */
    uint32_t   batch_size;                /* requests running on the device when it started */
    uint32_t   pcg_iterations;
    int32_t    pcg_success;
    int32_t    pad;
    double     pcg_residual;
} bnn_request_result;
typedef struct bnn_server_stats {
    uint64_t submitted, completed, failed, batches; /* batches: requests run */
    uint64_t scenes_resident, scenes_waiting, resident_bytes, peak_resident_bytes;
    double   latency_p50_ms, latency_p99_ms, latency_mean_ms, latency_max_ms; /* submit -> done */
    double   throughput_rps; /* completed requests / (last completion - first submission) */
    double   mean_batch;     /* requests in flight per engine round, on average (1 without batching) */
    double   busy_s;         /* summed wall time of requests, start to done */
} bnn_server_stats;
bnn_status  bnn_server_create(const bnn_server_config* cfg, bnn_server** out);
bnn_status  bnn_server_destroy(bnn_server* s); /* finishes queued work first */
const char* bnn_server_last_error(const bnn_server* s);
double      bnn_server_now(const bnn_server* s);
/* queue a scene (parameters + seeded generator); BNN_ERR_REJECTED when its footprint alone exceeds
 * the budget. Requests may be submitted at once; they run after the scene is resident. */
bnn_status  bnn_server_add_scene(bnn_server* s, const bnn_flip_params* p, const bnn_scene_spec* spec, uint64_t seed,
                                 uint64_t* scene_id);
/* free a scene after its queued requests */
bnn_status  bnn_server_release_scene(bnn_server* s, uint64_t scene_id);
bnn_status  bnn_server_submit(bnn_server* s, uint64_t scene_id, int32_t kind, float dt, uint64_t* request_id);
/* wait for one request (timeout_s < 0: no limit); BNN_ERR_STATE on timeout */
bnn_status  bnn_server_wait(bnn_server* s, uint64_t request_id, double timeout_s, bnn_request_result* out);
/* wait until no request is queued or running */
bnn_status  bnn_server_drain(bnn_server* s, double timeout_s);
/* particles of an idle resident scene in generation order (*n receives the count; arrays may be NULL) */
bnn_status  bnn_server_get_particles(bnn_server* s, uint64_t scene_id, float* pos_xyz, float* vel_xyz, uint64_t* n);
bnn_status  bnn_server_get_stats(bnn_server* s, bnn_server_stats* out);
bnn_status  bnn_server_reset_stats(bnn_server* s);

#ifdef __cplusplus
}
#endif
#endif /* BNN_FLIP_H */
