"""Additional and animated boundary objects on the device (ParticleSolver.cpp:242-298,
PIC_Solver.cpp:70-118,1158-1168), against the oracle's restatement on the same inputs: the boundary
SDF of the domain box with a sphere container and a turned box (static, one CSG union) is equal bit
for bit; particles moved through them are projected to the same bits; an animated box (the
reference's rotating-box scene, linear rotation key frames) moves at every substep after the first
frame, and the boundary SDF after three frames is again equal bit for bit.

The shapes' inverse transforms are computed on the host in FP64 and stored in FP32 by both sides
(by different formulas: R^T / s on the device side, a cofactor inverse in the oracle), so bit
equality also checks that the two restatements agree.
"""
import numpy as np
import pytest

from banana_amd import scenes

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 600.0)

STATIC = [dict(geometry="sphere", translation=(0.02, -0.05, 0.03), scale=0.85),
          dict(geometry="box", box_min=(-0.6, -0.7, -0.5), box_max=(0.6, 0.7, 0.5), translation=(0.05, 0.0, 0.0),
               rotation_axis_angle=(0.3, 1.0, 0.2, 25.0))]
ROTATING = [dict(geometry="box", box_min=(-0.8, -0.4, -0.8), box_max=(0.8, 0.4, 0.8),
                 animation=dict(periodic=True, start_frame=1, cubic_rotation=False,
                                keyframes=[dict(frame=1, rotation_axis_angle=(1, 0, 0, 0)),
                                           dict(frame=3, rotation_axis_angle=(1, 0, 0, 30)),
                                           dict(frame=5, rotation_axis_angle=(1, 0, 0, 60))]))]


def _params(n, **kw):
    params = scenes.base_params(n, boundary_sdf_at_init=True)
    params.update(box_min=(-1.0, -1.0, -1.0), box_max=(1.0, 1.0, 1.0), cell_size=2.0 / n, **kw)
    return params


def _pair(oracle, bnn, n, objs, seed=4, **kw):
    from banana_amd import boundary_objects
    params = _params(n, **kw)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.7, 0.7, (6000, 3)).astype(np.float32)
    vel = rng.normal(0.0, 1.0, (6000, 3)).astype(np.float32)
    g = bnn.FLIPSolver(params)
    g.set_boundary_objects(objs)
    g.set_particles(pos, vel)
    o = oracle.Oracle(params)
    o.set_boundary_objects(boundary_objects(objs), len(objs))
    o.set_particles(pos, vel)
    return params, pos, vel, g, o


def test_static_objects_boundary_sdf_bit_exact(oracle, bnn):
    params, pos, vel, g, o = _pair(oracle, bnn, 32, STATIC)
    bg = g.grid("boundary_sdf")
    bo = o.field(oracle.F_BOUNDARY_SDF).astype(np.float32)
    assert np.array_equal(bg, bo), np.abs(bg - bo).max()
    # the union is tighter than the domain box alone somewhere (the objects act)
    dom = bnn.FLIPSolver(_params(32))
    assert np.count_nonzero(bg < dom.grid("boundary_sdf")) > 1000


def test_move_through_static_objects_bit_exact(oracle, bnn):
    params, pos, vel, g, o = _pair(oracle, bnn, 32, STATIC, reflect_velocity_at_boundary=1, boundary_reflection_multiplier=0.5)
    o.map_particles_to_grid()
    o.extrapolate_velocity()
    for c, f in (("u", oracle.F_U), ("v", oracle.F_V), ("w", oracle.F_W)):
        g.set_grid(c, o.field(f))
    big = np.float32(0.08)  # particles leave the sphere / box and are projected back, reflecting
    g.move_particles(big)
    o.move_particles(big)
    pg, vg = g.particles()
    po, vo = o.particles()
    assert np.array_equal(pg, po), np.abs(pg - po).max()
    assert np.array_equal(vg, vo), np.abs(vg - vo).max()
    # one projection along the union's gradient per step brings most particles back inside (where the
    # sphere and the box are both violated, a single projection can leave a particle short of the
    # other surface, in the reference as here)
    before, after = np.mean(o.boundary_sd(0, pos) < -1e-3), np.mean(o.boundary_sd(0, pg) < -1e-3)
    print(f"outside the boundary objects: {100 * before:.1f}% of the particles before the move, {100 * after:.1f}% after")
    assert before > 0.2 and after < 0.25 * before


def test_rotating_box_frames(oracle, bnn):
    """Fixed substeps (two per frame) so that both sides evaluate the animation at the same frame
    fractions; after three frames the boundary SDF (the box turned to its frame-2.5 angle) is equal
    bit for bit and the particles agree to the pressure solves' tolerance (device: FP64-refined
    pressure; oracle: FP32 IC(0)-PCG at rel-tol 1e-6)."""
    fd = np.float32(1.0 / 60.0)
    kw = dict(frame_duration=fd, min_timestep=fd / 2, max_timestep=fd / 2, pressure_mode=1)
    params, pos, vel, g, o = _pair(oracle, bnn, 24, ROTATING, **kw)
    for _ in range(3):
        g.advance_frame()
        o.advance_frame()
    bg = g.grid("boundary_sdf")
    bo = o.field(oracle.F_BOUNDARY_SDF).astype(np.float32)
    assert np.array_equal(bg, bo), np.abs(bg - bo).max()
    # the box has turned: its SDF differs from the unturned box's
    still = [dict(ROTATING[0])]
    still[0].pop("animation")
    s = bnn.FLIPSolver(_params(24, **kw))
    s.set_boundary_objects(still)
    assert np.count_nonzero(np.abs(bg - s.grid("boundary_sdf")) > 1e-3) > 100
    pg, _ = g.particles()
    po, _ = o.particles()
    e = np.linalg.norm(pg.astype(np.float64) - po) / np.linalg.norm(po.astype(np.float64))
    print(f"positions after 3 frames with a rotating box: rel L2 {e:.2e}")
    assert e <= 1e-5


def test_resumed_frame_counter(oracle, bnn):
    """A host that loads a memory state itself hands the frame count over
    (bnn_flip_set_finished_frame): the animated box is then placed by the same frames as in the
    oracle resumed at that frame, and differently from a device left at frame 0."""
    fd = np.float32(1.0 / 60.0)
    kw = dict(frame_duration=fd, min_timestep=fd / 2, max_timestep=fd / 2, pressure_mode=1)
    params, pos, vel, g, o = _pair(oracle, bnn, 24, ROTATING, **kw)
    g.set_finished_frame(3)
    o.set_finished_frame(3)
    g.advance_frame()
    o.advance_frame()
    assert g.stats()["finished_frame"] == 4
    bg = g.grid("boundary_sdf")
    assert np.array_equal(bg, o.field(oracle.F_BOUNDARY_SDF).astype(np.float32))
    fresh = bnn.FLIPSolver(params)
    fresh.set_boundary_objects(ROTATING)
    fresh.set_particles(pos, vel)
    fresh.advance_frame()  # frame 0: the animation does not run yet
    assert np.count_nonzero(np.abs(bg - fresh.grid("boundary_sdf")) > 1e-3) > 100
