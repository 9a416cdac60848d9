"""Scene JSON in the reference's format -> parameters and generator spec (host-only, CPU)."""
import ctypes as C
import json

import numpy as np
import pytest

import banana_amd
from banana_amd import _lib

DAM = {
    "GlobalParameters": {"Solver": "FLIP_3DSolver", "FrameDuration": 0.03333333333, "StartFrame": 1, "FinalFrame": 4,
                         "SaveFrameData": False, "EnableSortParticle": False, "SortFrequency": 10},
    "SimulationParameters": {
        "SimulationDomainBox": {"BoxMin": [0, 0, 0], "BoxMax": [1, 1, 1]},
        "CellSize": 0.03125, "CGRelativeTolerance": 1e-6, "MaxCGIteration": 1000, "CFLFactor": 0.5,
        "PIC_FLIP_Ratio": 0.5,  # mis-cased in the reference's scene files: the default 0.97 stays
        "BoundarySDFAtInit": True,
    },
    "ParticleGenerators": [
        {"GeometryType": "Box", "UniqueName": "Dam", "BoxMin": [0, 0, 0], "BoxMax": [0.4, 0.6, 1.0]},
        {"GeometryType": "Sphere", "UniqueName": "Ball", "Translation": [0.7, 0.75, 0.5], "Scale": 0.12,
         "InitialVelocity": [0, -2, 0]},
    ],
}


def parse(obj_or_text):
    L = _lib.load()
    text = obj_or_text if isinstance(obj_or_text, str) else json.dumps(obj_or_text)
    p, s = _lib.Params(), _lib.SceneSpec()
    warn = C.create_string_buffer(4096)
    st = L.bnn_scene_from_json(text.encode(), C.byref(p), C.byref(s), warn, 4096)
    return st, p, s, warn.value.decode()


def test_dam_break_scene(bnn):
    st, p, s, warn = parse(DAM)
    assert st == 0, _lib.load().bnn_flip_last_error(None)
    assert np.isclose(p.cell_size, 1 / 32) and np.isclose(p.cg_relative_tolerance, 1e-6) and p.max_cg_iteration == 1000
    assert np.isclose(p.cfl_factor, 0.5) and np.isclose(p.frame_duration, 1 / 30) and p.final_frame == 4
    assert np.isclose(p.pic_flip_ratio, 0.97), "case-sensitive key: PIC_FLIP_Ratio must not be read"
    assert "PIC_FLIP_Ratio" in warn
    assert p.boundary_sdf_at_init == 1
    assert list(p.box_min) == [0, 0, 0] and list(p.box_max) == [1, 1, 1]
    assert s.n_boxes == 1 and s.n_spheres == 1
    assert np.allclose(list(s.box_max[0]), [0.4, 0.6, 1.0])
    assert np.isclose(s.sphere_radius[0], 0.12) and np.allclose(list(s.sphere_center[0]), [0.7, 0.75, 0.5])
    assert np.allclose(list(s.sphere_vel[0]), [0, -2, 0])
    # same particles as the Python scene description of this workload
    from banana_amd import scenes

    params, spec = scenes.dam_break(32, sphere=True, boundary_sdf_at_init=True)
    a = banana_amd.generate_scene_host(params, spec, 9)
    b = banana_amd.generate_scene_host(p, s, 9)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_comments_and_defaults(bnn):
    text = "// a scene\n" + json.dumps({"GlobalParameters": {}, "SimulationParameters": {"SimulationDomainBox": {}},
                                         "ParticleGenerators": [{"GeometryType": "Box"}]})
    st, p, s, warn = parse(text)
    assert st == 0
    assert np.isclose(p.cell_size, 1 / 64) and list(p.box_min) == [-1, -1, -1]  # ParticleSolverData.h:103-104
    assert s.n_boxes == 1


@pytest.mark.parametrize("mutate,status", [
    (lambda d: d["GlobalParameters"].update(Solver="SPH_3DSolver"), 5),
    (lambda d: d.update(AdditionalBoundaryObjects=[{"GeometryType": "Torus"}]), 5),
    (lambda d: d.update(AdditionalBoundaryObjects=[{"GeometryType": "Box", "RotationEulerAngles": [0, 0, 90]}]), 5),
    (lambda d: d["ParticleGenerators"].append({"GeometryType": "Torus"}), 5),
    (lambda d: d["SimulationParameters"].pop("SimulationDomainBox"), 1),
    (lambda d: d.pop("ParticleGenerators"), 1),
    (lambda d: d["SimulationParameters"].update(GravityType="Sideways"), 1),
])
def test_rejections(bnn, mutate, status):
    d = json.loads(json.dumps(DAM))
    mutate(d)
    st, p, s, warn = parse(d)
    assert st == status


def test_malformed_json(bnn):
    st, p, s, warn = parse('{"GlobalParameters": {"Solver": "FLIP_3DSolver",}')
    assert st == 1
    assert b"scene JSON" in _lib.load().bnn_flip_last_error(None)


def test_cpp_adapter_and_cli_build(bnn, tmp_path):
    """The C++ adapter header compiles stand-alone against the C ABI (the CLI is built by build())."""
    import os
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "use.cpp"
    src.write_text('#include "bnn_flip_solver.hpp"\nint main(){ bnn::FLIP_Solver3D s; return s.getSolverName() == "FLIP_3DSolver" ? 0 : 1; }\n')
    subprocess.check_call(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(root, "include"), str(src)])
    assert os.path.exists(os.path.join(root, "banana_amd", "bnn_flip_cli"))


def test_solver_names_select_the_transfer_scheme(bnn):
    """ParticleSolverFactory names of the PIC family (FLIP_Solver.h:37, APIC_Solver.h:37,
    AFLIP_Solver.h:36) select the transfer scheme; RepulsiveForceStiffness and CorrectPosition are read."""
    import copy

    for name, scheme in (("FLIP_3DSolver", 0), ("APIC_3DSolver", 1), ("AFLIP_3DSolver", 2)):
        d = copy.deepcopy(DAM)
        d["GlobalParameters"]["Solver"] = name
        d["SimulationParameters"].update(CorrectPosition=True, RepulsiveForceStiffness=20.0)
        st, p, s, warn = parse(d)
        assert st == 0 and p.transfer_scheme == scheme, name
        assert p.correct_position == 1 and p.repulsive_force_stiffness == 20.0


def test_boundary_objects_from_json(bnn):
    """AdditionalBoundaryObjects as the reference's scene files write them (FLIP3D_BallDrop.json,
    FLIP3D_RotationalBoxBoundary.json): geometry, transform keys, animation key frames (angles in
    degrees), the Periodic / StartFrame pair and the interpolation flags."""
    import banana_amd
    d = json.loads(json.dumps(DAM))
    d["AdditionalBoundaryObjects"] = [
        {"GeometryType": "Sphere", "UniqueName": "SphereBoundary", "CSGOperator": "Intersection", "Translation": [0, 0, 0], "Scale": 1},
        {"GeometryType": "Box", "UniqueName": "BoxBoundary", "CSGOperator": "Intersection", "BoxMin": [-1, -0.5, -1],
         "BoxMax": [1, 0.5, 1], "Animation": {"Periodic": True, "StartFrame": 50, "CubicInterpolationRotation": False,
                                              "KeyFrames": [{"Frame": 50, "RotationAxisAngle": [1, 0, 0, 0]},
                                                            {"Frame": 150, "RotationAxisAngle": [1, 0, 0, 180]},
                                                            {"Frame": 250, "RotationAxisAngle": [1, 0, 0, 360]}]}}]
    arr, n = banana_amd.boundaries_from_json(json.dumps(d))
    assert n == 2
    s, b = arr[0], arr[1]
    assert s.geometry == 1 and s.has_translation == 1 and s.has_scale == 1 and s.scale == 1.0 and s.n_keyframes == 0
    assert b.geometry == 0 and list(b.box_min) == [-1, -0.5, -1] and list(b.box_max) == [1, 0.5, 1]
    assert b.n_keyframes == 3 and b.periodic == 1 and b.start_frame == 50
    assert (b.cubic_translation, b.cubic_rotation, b.cubic_scale) == (1, 0, 1)
    assert [k.frame for k in b.keyframes[:3]] == [50, 150, 250]
    assert [k.rotation_axis_angle[3] for k in b.keyframes[:3]] == [0, 180, 360] and all(k.has_rotation for k in b.keyframes[:3])
    # a scene with them still parses (the boundary objects are not part of the parameters)
    st, p, sp, warn = parse(d)
    assert st == 0
