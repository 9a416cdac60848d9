"""Scene descriptions for the FLIP dam-break workloads.

Each scene is (params, spec): `params` holds the solver parameters under the reference's JSON key
meanings (SimulationParameters / GlobalParameters, ParticleSolverData.h:38-175) and `spec` the
seeded particle generator input (objects filled at ppa^3 particles per cell, jittered in sub-cells).
"""
import math

import numpy as np

DT_600 = 1.0 / 600.0


def base_params(n, boundary_sdf_at_init=False, cg_tol=1e-6):
    """Unit-cube domain at cell size 1/n with the reference defaults (Constants.h:31-45)."""
    return dict(
        box_min=(0.0, 0.0, 0.0), box_max=(1.0, 1.0, 1.0),
        cell_size=1.0 / n, ratio_cell_size_pradius=2.0, n_expand_cells=2,
        min_timestep=DT_600, max_timestep=DT_600, cfl_factor=1.0, frame_duration=DT_600,
        cg_relative_tolerance=cg_tol, max_cg_iteration=5000,
        pic_flip_ratio=0.97, advection_steps=1, gravity_type=0,
        gravity_direction=(0.0, 0.0, 0.0), gravity_center=(0.0, 0.0, 0.0), apply_gravity=1,
        reflect_velocity_at_boundary=0, boundary_reflection_multiplier=1.0,
        boundary_sdf_at_init=1 if boundary_sdf_at_init else 0,
    )


def dam_break_spec(sphere=False):
    spec = dict(ppa=2, jitter=1.0, vel_sigma=0.0,
                boxes=[dict(min=(0.0, 0.0, 0.0), max=(0.4, 0.6, 1.0), vel=(0.0, 0.0, 0.0))], spheres=[])
    if sphere:
        spec["spheres"] = [dict(center=(0.7, 0.75, 0.5), radius=0.12, vel=(0.0, -2.0, 0.0))]
    return spec


def dam_break(n, sphere=False, boundary_sdf_at_init=False, cg_tol=1e-6):
    """Dam break [0,0.4]x[0,0.6]x[0,1] on the unit cube at n^3 (+2-cell band), 8 particles/cell."""
    return base_params(n, boundary_sdf_at_init, cg_tol), dam_break_spec(sphere)


def random_scene(rng, n_lo=24, n_hi=96, vel_sigma=0.5, boundary_sdf_at_init=True, cg_tol=1e-6):
    """One scene of the uneven batch: edge length log-uniform in [n_lo, n_hi] cells, fluid-box
    extents uniform in [0.2, 0.8] per axis, N(0, vel_sigma^2) initial velocities."""
    n = int(round(n_lo * (n_hi / n_lo) ** rng.uniform()))
    ext = rng.uniform(0.2, 0.8, size=3)
    params = base_params(n, boundary_sdf_at_init, cg_tol)
    spec = dict(ppa=2, jitter=1.0, vel_sigma=vel_sigma,
                boxes=[dict(min=(0.0, 0.0, 0.0), max=tuple(float(e) for e in ext), vel=(0.0, 0.0, 0.0))], spheres=[])
    return params, spec


def grid_dims(params):
    """Cell counts n = ceil((bMax - bMin) / dx) in float32, bMin/bMax grown by NExpandCells
    (ParticleSolverData.cpp:247-253, Grid.cpp:47)."""
    dx = np.float32(params["cell_size"])
    ex = dx * np.float32(params["n_expand_cells"])
    n = []
    bmin = []
    for d in range(3):
        lo = np.float32(params["box_min"][d]) - ex
        hi = np.float32(params["box_max"][d]) + ex
        n.append(int(math.ceil(np.float32((hi - lo) / dx))))
        bmin.append(lo)
    return tuple(n), np.array(bmin, np.float32), dx
