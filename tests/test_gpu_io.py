"""Frame output and memory states on the device (PIC_Solver.cpp:158-262): the files follow the
reference's BananaParticleData layout byte for byte (checked with an independent reader written from
ParticleSerialization.cpp), compressed positions equal the reference's float formula
(ParticleHelpers.cpp:47-79) bit for bit, and a saved state reloads into a new scene exactly, so that
stepping on from it matches stepping on from the same state set through the ABI, bit for bit.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import bpd
from banana_amd import scenes

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 600.0)


def _scene(bnn, n=24):
    params, spec = scenes.dam_break(n, sphere=True, boundary_sdf_at_init=True)
    spec["vel_sigma"] = 0.5
    g = bnn.FLIPSolver(params)
    g.generate_scene(spec, 21)
    return params, spec, g


def test_frame_data_file(bnn, tmp_path):
    params, spec, g = _scene(bnn)
    path = str(tmp_path / "FrameData" / "frame.0000.bnn")
    g.save_frame_data(path, save_velocity=True, save_object_index=True)
    lines, fixed, particle, n = bpd.read(path)
    pos, vel = g.particles()
    assert n == pos.shape[0]
    # header: fixed attributes by name, then particle attributes by name (std::map order)
    assert lines == ["BananaParticleData",
                     "FixedAttribute NObjects uint 4 1 4",
                     "FixedAttribute particle_radius real 4 1 4",
                     f"ParticleAttribute object_index uint16 2 1 {2 * n}",
                     f"ParticleAttribute particle_position compressed_real 4 3 {24 + 6 * n}",
                     f"ParticleAttribute particle_velocity compressed_real 4 3 {24 + 6 * n}",
                     f"NParticles {n}", "EndHeader."]
    assert bpd.values(fixed["particle_radius"])[0] == np.float32(params["cell_size"]) / np.float32(2.0)
    assert bpd.values(fixed["NObjects"])[0] == 2  # the box and the sphere generators
    for name, v in (("particle_position", pos), ("particle_velocity", vel)):
        mn, mx, q = bpd.compressed(particle[name], n)
        rmn, rmx, rq = bpd.compress_reference(v)
        assert np.array_equal(mn, rmn) and np.array_equal(mx, rmx), name
        assert np.array_equal(q, rq), name
        # decompression error within one quantum (plus the rounding of the decompression itself)
        back = mn.astype(np.float64) + q / 65535.0 * (mx.astype(np.float64) - mn)
        assert np.all(np.abs(back - v) <= (mx.astype(np.float64) - mn) / 65535.0 * 1.01), name
    # object indices: generator order (box 0, then the sphere's particles not already in the box)
    obj = bpd.values(particle["object_index"])
    obj_g, nobj = g.object_index()
    assert np.array_equal(obj, obj_g) and nobj == 2
    assert set(np.unique(obj)) == {0, 1}
    g.close()


def test_memory_state_round_trip(bnn, tmp_path):
    params, spec, a = _scene(bnn)
    for _ in range(3):
        a.step(DT)
    path = str(tmp_path / "MemoryState" / "frame.0003.bnn")
    a.save_memory_state(path)
    lines, fixed, particle, n = bpd.read(path)
    nu = int(np.prod(a.field_shape("u")))
    nv = int(np.prod(a.field_shape("v")))
    nw = int(np.prod(a.field_shape("w")))
    assert lines == ["BananaParticleData", "FixedAttribute NObjects uint 4 1 4", "FixedAttribute grid_resolution uint 4 3 12",
                     f"FixedAttribute grid_velocity_0 real 4 {nu} {4 * nu}", f"FixedAttribute grid_velocity_1 real 4 {nv} {4 * nv}",
                     f"FixedAttribute grid_velocity_2 real 4 {nw} {4 * nw}", "FixedAttribute particle_radius real 4 1 4",
                     f"ParticleAttribute object_index uint16 2 1 {2 * n}", f"ParticleAttribute particle_position real 4 3 {12 * n}",
                     f"ParticleAttribute particle_velocity real 4 3 {12 * n}", f"NParticles {n}", "EndHeader."]
    assert tuple(bpd.values(fixed["grid_resolution"])) == a.n
    pos_a, vel_a = a.particles()
    assert np.array_equal(bpd.values(particle["particle_position"]).reshape(-1, 3), pos_a)
    assert np.array_equal(bpd.values(particle["particle_velocity"]).reshape(-1, 3), vel_a)
    assert np.array_equal(bpd.values(fixed["grid_velocity_0"]), a.grid("u").ravel())

    b = bnn.FLIPSolver(params)
    b.load_memory_state(path, 3)
    pos_b, vel_b = b.particles()
    assert np.array_equal(pos_b, pos_a) and np.array_equal(vel_b, vel_a)
    for c in ("u", "v", "w"):
        assert np.array_equal(b.grid(c), a.grid(c)), c
    assert np.array_equal(b.object_index()[0], a.object_index()[0])
    assert b.stats()["finished_frame"] == 3
    # the same state set through the ABI steps bit-identically. (The original differs: a new scene's
    # first constrain sees the zero-filled weights, FLIP_Solver.cpp:45 before PIC_Solver.cpp:534,
    # in the reference as here.)
    c = bnn.FLIPSolver(params)
    c.set_particles(pos_a, vel_a)
    for f in ("u", "v", "w"):
        c.set_grid(f, a.grid(f))
    for _ in range(2):
        for s in (b, c):
            s.step(DT)
        pb, vb = b.particles()
        pc, vc = c.particles()
        assert np.array_equal(pb, pc) and np.array_equal(vb, vc)
    for s in (a, b, c):
        s.close()


def test_memory_state_rejects_another_grid(bnn, tmp_path):
    params, spec, a = _scene(bnn, 24)
    path = str(tmp_path / "s.bnn")
    a.save_memory_state(path)
    p2, _ = scenes.dam_break(32)
    b = bnn.FLIPSolver(p2)
    with pytest.raises(bnn.BnnError, match="grid resolution"):
        b.load_memory_state(path, 1)
    with pytest.raises(bnn.BnnError):
        b.load_memory_state(str(tmp_path / "missing.bnn"), 1)
    a.close()
    b.close()


def test_cli_saves_and_resumes(bnn, tmp_path):
    """The command-line driver follows the reference's flow: frame 0 is saved when the particles are
    generated, every frame saves its frame data and (FramePerState) its memory state, and a second
    run resumes from the latest state instead of generating (PIC_Solver.cpp:40-66, 307-310)."""
    cli = os.path.join(os.path.dirname(bnn.__file__), "bnn_flip_cli")
    scene = {"GlobalParameters": {"Solver": "FLIP_3DSolver", "FrameDuration": 1.0 / 300.0, "StartFrame": 1, "FinalFrame": 2,
                                  "DataPath": str(tmp_path), "SaveFrameData": True, "SaveMemoryState": True, "FramePerState": 2,
                                  "OptionalSavingData": ["ParticleVelocity"]},
             "SimulationParameters": {"SimulationDomainBox": {"BoxMin": [0, 0, 0], "BoxMax": [1, 1, 1]}, "CellSize": 1.0 / 20,
                                      "MinTimestep": 1.0 / 600, "MaxTimestep": 1.0 / 600, "BoundarySDFAtInit": True},
             "ParticleGenerators": [{"GeometryType": "Box", "BoxMin": [0, 0, 0], "BoxMax": [0.4, 0.6, 1.0]}]}
    sf = tmp_path / "scene.json"
    sf.write_text(json.dumps(scene))
    r = subprocess.run([cli, str(sf), "--seed", "3", "--dump", str(tmp_path / "run1.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for f in (0, 1, 2):
        assert os.path.exists(bnn.data_file_path(str(tmp_path), "FrameData", f)), f
    assert not os.path.exists(bnn.data_file_path(str(tmp_path), "MemoryState", 1))  # FramePerState = 2
    assert os.path.exists(bnn.data_file_path(str(tmp_path), "MemoryState", 2))
    # second run to frame 3: resumes from state 2 and runs one frame
    scene["GlobalParameters"]["FinalFrame"] = 3
    sf.write_text(json.dumps(scene))
    r = subprocess.run([cli, str(sf), "--seed", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    frames = [json.loads(l)["frame"] for l in r.stdout.splitlines() if l.startswith("{\"frame\"")]
    assert frames == [3]
    assert os.path.exists(bnn.data_file_path(str(tmp_path), "FrameData", 3))
