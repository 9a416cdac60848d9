"""The reference's preconditioner on the device: natural-order IC(0) with the 0.25 pivot safeguard
(PCGSolver.cpp:398-466) and its triangular solves (PCGSolver.cpp:243-269), level-scheduled over the
hyperplanes i+j+k, compared with the oracle's sequential restatement on the same grid state:
the factor, one application, and the PCG iterates (iteration count and converged pressure).

The device assembles in FP64 and rounds the coefficients to FP32, the oracle assembles in FP32, so
the two FP32 matrices can differ in the last bit of an entry; the bounds below are that rounding
carried through the recurrences, not a looser algorithm.
"""
import numpy as np
import pytest

from banana_amd import scenes

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 600.0)
IC0 = 2  # BNN_PRECOND_IC0


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _pair(oracle, bnn, n, seed=5):
    """Oracle (FP32, IC(0)) advanced to the assembled system; device with the same grid velocities,
    projected with IC(0)-PCG in FP32 mode."""
    params, spec = scenes.dam_break(n, sphere=True, boundary_sdf_at_init=True, cg_tol=1e-6)
    spec["vel_sigma"] = 0.5
    pos, vel = bnn.generate_scene_host(params, spec, seed)
    o = oracle.Oracle(params)
    o.set_particles(pos, vel)
    o.compute_cell_weights()
    o.map_particles_to_grid()
    o.extrapolate_velocity()
    o.constrain_grid_velocity()
    o.backup_grid_velocity()
    o.add_gravity(DT)
    o.compute_fluid_sdf()
    state = {f: o.field(f) for f in (oracle.F_U, oracle.F_V, oracle.F_W, oracle.F_WEIGHT_U, oracle.F_WEIGHT_V,
                                     oracle.F_WEIGHT_W, oracle.F_PHI, oracle.F_BOUNDARY_SDF)}
    o.compute_system(DT)
    g = bnn.FLIPSolver(params, pressure_mode=0, preconditioner=IC0)
    g.set_particles(pos, vel)
    for c, f in (("u", oracle.F_U), ("v", oracle.F_V), ("w", oracle.F_W)):
        g.set_grid(c, state[f])
    return params, state, o, g, pos, vel


@pytest.mark.parametrize("n", [24, 40])
def test_ic0_factor_and_apply_match_oracle(oracle, bnn, n):
    params, state, o, g, _, _ = _pair(oracle, bnn, n)
    g.pressure_projection(DT)
    assert np.array_equal(g.grid("phi"), state[oracle.F_PHI].astype(np.float32))
    inv_o = o.ic0_factor()
    inv_g = g.ic0_factor()
    assert inv_g.shape == inv_o.shape
    assert np.array_equal(inv_g == 0, inv_o == 0)  # the same rows skipped
    d = np.abs(inv_g - inv_o) / np.maximum(np.abs(inv_o), 1e-30)
    same = np.mean(inv_g == inv_o)
    print(f"n={n}: IC(0) factor max rel diff {d.max():.2e}, {100 * same:.1f}% bit-equal")
    assert d.max() <= 2e-6
    rhs = g.rhs()
    o.set_precond(2)
    z_o = o.apply_precond(rhs)
    z_g = g.apply_preconditioner(rhs)
    e = rel_l2(z_g, z_o)
    print(f"n={n}: one IC(0) application, rel L2 vs the oracle {e:.2e}")
    assert e <= 1e-5


@pytest.mark.parametrize("n", [24, 40])
def test_ic0_pcg_iterates_match_oracle(oracle, bnn, n):
    """Same stopping rule (max|r| < 1e-6 max|b|, PCGSolver.cpp:154-186) and the same preconditioner:
    the iteration counts agree to within rounding noise, and both solutions are equally close to the
    FP64 solve of the system."""
    params, state, o, g, _, _ = _pair(oracle, bnn, n)
    o.set_precond(2)
    assert o.solve_system() == 0
    it_o, res_o, ok_o = o.pcg_stats()
    _, p_o = o.system()
    g.pressure_projection(DT)
    st = g.stats()
    assert st["pcg_success"] == 1
    it_g = st["pcg_iterations"]
    p_g = g.pressure()
    o64 = oracle.Oracle(params, precision=64)
    for f, a in state.items():
        o64.set_field(f, a)
    o64.compute_system(DT)
    o64.set_cg(1e-15, 20000)
    assert o64.solve_system() == 0
    _, p64 = o64.system()
    e_g, e_o = rel_l2(p_g, p64), rel_l2(p_o, p64)
    print(f"n={n}: IC(0)-PCG iterations device {it_g}, oracle {it_o}; pressure vs FP64: device {e_g:.2e}, oracle {e_o:.2e}")
    assert ok_o and abs(it_g - it_o) <= max(2, int(0.05 * it_o))
    assert e_g <= max(1e-5, 1.5 * e_o)


def test_ic0_refined_reaches_fp64(oracle, bnn):
    """IC(0) as the inner preconditioner of the FP64 defect correction reaches the 1e-9 bar."""
    params, state, o, _, pos, vel = _pair(oracle, bnn, 24)
    g = bnn.FLIPSolver(params, pressure_mode=1, refine_tolerance=1e-11, preconditioner=IC0)
    g.set_particles(pos, vel)
    for c, f in (("u", oracle.F_U), ("v", oracle.F_V), ("w", oracle.F_W)):
        g.set_grid(c, state[f])
    g.pressure_projection(DT)
    assert g.stats()["pcg_success"] == 1
    o64 = oracle.Oracle(params, precision=64)
    for f, a in state.items():
        o64.set_field(f, a)
    o64.compute_system(DT)
    o64.set_cg(1e-15, 20000)
    assert o64.solve_system() == 0
    _, p64 = o64.system()
    assert rel_l2(g.pressure(), p64) <= 1e-9
