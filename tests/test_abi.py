"""The C ABI library loads and exports every symbol include/bnn_flip.h declares; host-only logic
(defaults, generator, error reporting without a device). No device compute here."""
import ctypes as C
import os
import re

import numpy as np

import banana_amd
from banana_amd import _lib, scenes

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "bnn_flip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bnn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(bnn):
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    # and the Python view binds exactly these
    assert set(names) == set(_lib.SIGNATURES)


def test_struct_layout_matches_header(bnn, tmp_path):
    """Compile the header with the host C compiler and compare sizes/offsets with the ctypes view."""
    import subprocess

    structs = {"bnn_flip_params": _lib.Params, "bnn_scene_spec": _lib.SceneSpec, "bnn_flip_stats": _lib.Stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-o", str(exe), str(src)])
    out = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, cls in structs.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


def test_default_params_are_reference_defaults(bnn):
    p = banana_amd.make_params()
    assert np.isclose(p.cell_size, 1 / 64) and p.ratio_cell_size_pradius == 2 and p.n_expand_cells == 2
    assert np.isclose(p.min_timestep, 1e-6) and np.isclose(p.max_timestep, 1 / 30) and p.cfl_factor == 1
    assert np.isclose(p.cg_relative_tolerance, 1e-15) and p.max_cg_iteration == 5000
    assert np.isclose(p.pic_flip_ratio, 0.97) and p.advection_steps == 1 and p.apply_gravity == 1
    assert p.boundary_sdf_at_init == 0 and p.correct_position == 0 and p.sort_frequency == 10
    assert p.repulsive_force_stiffness == 50.0  # ParticleSolverData.h:143


def test_host_generator_matches_oracle(bnn, oracle):
    params, spec = scenes.dam_break(24, sphere=True)
    spec["vel_sigma"] = 0.5
    a = banana_amd.generate_scene_host(params, spec, 123)
    b = oracle.generate_scene(params, spec, 123)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_create_reports_errors_without_exiting(bnn):
    L = _lib.load()
    p = banana_amd.make_params(scenes.base_params(16))
    p.cell_size = -1.0
    h = C.c_void_p()
    st = L.bnn_flip_create(C.byref(p), C.byref(h))
    assert st == 1 and not h.value
    assert b"CellSize" in L.bnn_flip_last_error(None)
    p = banana_amd.make_params(scenes.base_params(16))
    p.preconditioner = 7
    assert L.bnn_flip_create(C.byref(p), C.byref(h)) == 1  # BNN_ERR_INVALID_ARG
    assert b"preconditioner" in L.bnn_flip_last_error(None)
    # null arguments
    assert L.bnn_flip_step(None, 0.01) == 1
    assert L.bnn_flip_destroy(None) == 1
