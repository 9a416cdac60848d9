"""APIC and AFLIP (APIC_Solver.cpp:42-125, AFLIP_Solver.cpp:43-140) on the device against the oracle:
the splat with the affine term (same weights and terms, FP32 sums in another order), the gathers of
v and C bit for bit, the binning permute carrying C, and whole substeps to the velocity bar."""
import numpy as np
import pytest

from banana_amd import scenes

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 600.0)
SCHEMES = {"apic": 1, "aflip": 2}


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _pair(oracle, bnn, scheme, n=32, walls=False, seed=3, shuffle=True):
    params, spec = scenes.dam_break(n, sphere=True, boundary_sdf_at_init=walls)
    spec["vel_sigma"] = 0.5
    params.update(transfer_scheme=scheme)
    pos, vel = bnn.generate_scene_host(params, spec, seed)
    if shuffle:
        idx = np.random.default_rng(seed).permutation(pos.shape[0])
        pos, vel = pos[idx], vel[idx]
    rng = np.random.default_rng(seed + 1)
    C = rng.normal(0, 2.0, (pos.shape[0], 3, 3)).astype(np.float32)
    g = bnn.FLIPSolver(params)
    g.set_particles(pos, vel)
    assert np.array_equal(g.affine(), np.zeros_like(C))  # new particles start with C = 0
    g.set_affine(C)
    o = oracle.Oracle(params)
    o.set_particles(pos, vel)
    o.set_scheme(scheme)
    o.set_affine(C)
    return params, pos, vel, C, g, o


@pytest.mark.parametrize("scheme", SCHEMES.values(), ids=SCHEMES.keys())
def test_affine_p2g(oracle, bnn, scheme):
    params, pos, vel, C, g, o = _pair(oracle, bnn, scheme)
    g.map_particles_to_grid()
    o.map_particles_to_grid()
    for c, f, fv in (("u", oracle.F_U, oracle.F_VALID_U), ("v", oracle.F_V, oracle.F_VALID_V), ("w", oracle.F_W, oracle.F_VALID_W)):
        a, b = g.grid(c), o.field(f)
        assert np.array_equal(g.grid("valid_" + c), o.field(fv)), c
        assert np.max(np.abs(a - b)) <= 1e-5 * max(1.0, np.abs(b).max()), c
    # the binning permute moved C with the particles: the caller's order reads back unchanged
    assert np.array_equal(g.affine(), C)


@pytest.mark.parametrize("scheme", SCHEMES.values(), ids=SCHEMES.keys())
def test_affine_g2p_bit_exact(oracle, bnn, scheme):
    params, pos, vel, C, g, o = _pair(oracle, bnn, scheme)
    o.map_particles_to_grid()
    o.extrapolate_velocity()
    o.backup_grid_velocity()
    o.add_gravity(DT)
    for c, f, fo in (("u", oracle.F_U, oracle.F_OLD_U), ("v", oracle.F_V, oracle.F_OLD_V), ("w", oracle.F_W, oracle.F_OLD_W)):
        g.set_grid(c, o.field(f))
        g.set_grid("old_" + c, o.field(fo))
    g.map_grid_to_particles()
    o.compute_changes_grid_velocity()
    o.map_grid_to_particles()
    _, vg = g.particles()
    _, vo = o.particles()
    assert np.array_equal(vg, vo), np.abs(vg - vo).max()
    cg, co = g.affine(), o.affine()
    assert np.array_equal(cg, co), np.abs(cg - co).max()


@pytest.mark.parametrize("scheme", SCHEMES.values(), ids=SCHEMES.keys())
def test_affine_steps(oracle, bnn, scheme):
    """Whole substeps (APIC: P2G, gravity, projection, extrapolate, constrain, G2P; AFLIP: FLIP's
    sequence), the device refined against the FP32 oracle at its floor and the FP64 oracle."""
    params, pos, vel, C, g0, _ = _pair(oracle, bnn, scheme, walls=True, shuffle=False)
    params.update(cg_relative_tolerance=1e-7)
    g = bnn.FLIPSolver(params, pressure_mode=1, refine_tolerance=1e-10)
    g.set_particles(pos, vel)
    g.set_affine(C)
    outs = {}
    for prec in (32, 64):
        o = oracle.Oracle(params, precision=prec)
        o.set_particles(pos, vel)
        o.set_scheme(scheme)
        o.set_affine(C)
        if prec == 64:
            o.set_cg(1e-14, 20000)
        outs[prec] = o
    for s in range(2):
        g.step(DT)
        assert g.stats()["pcg_success"] == 1
        pg, vg = g.particles()
        for prec, o in outs.items():
            o.step(DT)
            po, vo = o.particles()
            e = rel_l2(vg, vo)
            ec = rel_l2(g.affine(), o.affine())
            print(f"step {s + 1} vs FP{prec} oracle: velocity {e:.2e}, C {ec:.2e}, max |dpos| {np.abs(pg - po).max():.2e}")
            if s == 0:
                assert e <= 1e-5 and ec <= 1e-4
            assert e <= 1e-4
            assert np.abs(pg - po).max() <= 1e-5
    g0.close()
    g.close()
