"""FLIP_Solver-shaped Python view of one device-resident scene (a thin wrapper on include/bnn_flip.h).

Method names follow the reference's override points (PIC_Solver.h:64-70, FLIP_Solver.h:49-52,
ParticleSolver.h:51-64); every call goes through libbnn_flip.so to the gfx950 kernels.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import FIELD, BnnError, Params, SceneSpec, Stats


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_params(d=None, **kw):
    """bnn_flip_params from the library defaults (Constants.h:31-45) overridden by d / kw."""
    L = _lib.load()
    p = Params()
    L.bnn_flip_default_params(C.byref(p))
    for k, v in dict(d or {}, **kw).items():
        if isinstance(v, (list, tuple, np.ndarray)):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def make_spec(d):
    s = SceneSpec()
    s.ppa = d.get("ppa", 2)
    s.jitter = d.get("jitter", 1.0)
    s.vel_sigma = d.get("vel_sigma", 0.0)
    boxes = d.get("boxes", [])
    s.n_boxes = len(boxes)
    for i, b in enumerate(boxes):
        for a in range(3):
            s.box_min[i][a] = b["min"][a]
            s.box_max[i][a] = b["max"][a]
            s.box_vel[i][a] = b.get("vel", (0, 0, 0))[a]
    sph = d.get("spheres", [])
    s.n_spheres = len(sph)
    for i, b in enumerate(sph):
        s.sphere_radius[i] = b["radius"]
        for a in range(3):
            s.sphere_center[i][a] = b["center"][a]
            s.sphere_vel[i][a] = b.get("vel", (0, 0, 0))[a]
    return s


def io_from_json(text):
    """The file IO keys of a scene's GlobalParameters (bnn_io_from_json) as a dict."""
    L = _lib.load()
    io = _lib.IoParams()
    st = L.bnn_io_from_json(text.encode() if isinstance(text, str) else text, C.byref(io))
    if st != 0:
        raise BnnError(st, "bnn_io_from_json failed")
    return {k: (getattr(io, k).decode() if isinstance(getattr(io, k), bytes) else getattr(io, k)) for k, _ in _lib.IoParams._fields_}


def boundary_objects(objs):
    """ctypes array of bnn_boundary_object from dicts: geometry ("box" / "sphere"), box_min, box_max,
    translation, rotation_axis_angle (degrees), scale, and an optional animation dict with
    keyframes [{frame, translation, rotation_axis_angle, scale}], periodic, start_frame and
    cubic_translation / cubic_rotation / cubic_scale (default True)."""
    arr = (_lib.BoundaryObject * max(len(objs), 1))()
    for o, d in zip(arr, objs):
        o.geometry = 1 if d.get("geometry", "box").lower() == "sphere" else 0
        for a in range(3):
            o.box_min[a] = d.get("box_min", (-1, -1, -1))[a]
            o.box_max[a] = d.get("box_max", (1, 1, 1))[a]
            o.translation[a] = d.get("translation", (0, 0, 0))[a]
        r = d.get("rotation_axis_angle", (1, 1, 1, 0))
        for a in range(4):
            o.rotation_axis_angle[a] = r[a]
        o.scale = d.get("scale", 1.0)
        o.has_translation = int("translation" in d)
        o.has_rotation = int("rotation_axis_angle" in d)
        o.has_scale = int("scale" in d)
        an = d.get("animation")
        o.cubic_translation = o.cubic_rotation = o.cubic_scale = 1
        if an:
            o.periodic = int(an.get("periodic", False))
            o.start_frame = int(an.get("start_frame", 0))
            o.cubic_translation = int(an.get("cubic_translation", True))
            o.cubic_rotation = int(an.get("cubic_rotation", True))
            o.cubic_scale = int(an.get("cubic_scale", True))
            o.n_keyframes = len(an["keyframes"])
            for k, kd in zip(o.keyframes, an["keyframes"]):
                k.frame = int(kd["frame"])
                for a in range(3):
                    k.translation[a] = kd.get("translation", (0, 0, 0))[a]
                rk = kd.get("rotation_axis_angle", (1, 1, 1, 0))
                for a in range(4):
                    k.rotation_axis_angle[a] = rk[a]
                k.scale = kd.get("scale", 1.0)
                k.has_rotation = int("rotation_axis_angle" in kd)
                k.has_scale = int("scale" in kd)
    return arr


def boundaries_from_json(text):
    """The AdditionalBoundaryObjects of a scene file as a ctypes array (bnn_boundaries_from_json)."""
    L = _lib.load()
    arr = (_lib.BoundaryObject * 8)()
    n = C.c_uint32(0)
    st = L.bnn_boundaries_from_json(text.encode() if isinstance(text, str) else text, C.cast(arr, C.c_void_p), 8, C.byref(n))
    if st != 0:
        raise BnnError(st, (L.bnn_flip_last_error(None) or b"").decode())
    return arr, n.value


def data_file_path(data_path, folder, frame):
    """DataIO::getFilePath: <data_path>/<folder>/frame.%04d.bnn"""
    L = _lib.load()
    buf = C.create_string_buffer(4096)
    st = L.bnn_data_file_path(os.fsencode(data_path), os.fsencode(folder), frame, buf, len(buf))
    if st != 0:
        raise BnnError(st, "bnn_data_file_path failed")
    return buf.value.decode()


def latest_state_index(data_path, folder, start_frame, final_frame):
    return int(_lib.load().bnn_latest_state_index(os.fsencode(data_path), os.fsencode(folder), start_frame, final_frame))


def generate_scene_host(params, spec, seed):
    """Host-side seeded generator (bnn_generate_scene_host); returns (pos, vel) as (n,3) float32."""
    L = _lib.load()
    p = params if isinstance(params, Params) else make_params(params)
    s = spec if isinstance(spec, SceneSpec) else make_spec(spec)
    n = C.c_uint64()
    _check(L.bnn_generate_scene_host(C.byref(p), C.byref(s), seed, None, None, 0, C.byref(n)), None)
    pos = np.zeros((n.value, 3), np.float32)
    vel = np.zeros((n.value, 3), np.float32)
    _check(L.bnn_generate_scene_host(C.byref(p), C.byref(s), seed, _ptr(pos), _ptr(vel), n.value, C.byref(n)), None)
    return pos, vel


def _check(status, h):
    if status != _lib.BNN_OK:
        L = _lib.load()
        msg = L.bnn_flip_last_error(h)
        raise BnnError(status, msg.decode() if msg else "")


class FLIPSolver:
    """One scene resident in HBM. `params` is a dict of bnn_flip_params fields (reference JSON key
    meanings) or a Params struct."""

    def __init__(self, params, **kw):
        self.L = _lib.load()
        self._p = params if isinstance(params, Params) else make_params(params, **kw)
        h = C.c_void_p()
        st = self.L.bnn_flip_create(C.byref(self._p), C.byref(h))
        if st != _lib.BNN_OK:
            raise BnnError(st, (self.L.bnn_flip_last_error(None) or b"").decode())
        self.h = h
        n = np.zeros(3, np.int32)
        self.L.bnn_flip_grid_dims(self.h, _ptr(n))
        self.n = tuple(int(x) for x in n)
        bmin = np.zeros(3, np.float32)
        dx = C.c_float()
        self.L.bnn_flip_grid_origin(self.h, _ptr(bmin), C.byref(dx))
        self.bmin, self.dx = bmin, np.float32(dx.value)

    def close(self):
        if getattr(self, "h", None):
            self.L.bnn_flip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, status):
        _check(status, self.h)
        return status

    # --- scene ------------------------------------------------------------------------------
    def set_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        assert pos.shape == vel.shape and pos.ndim == 2 and pos.shape[1] == 3
        self._c(self.L.bnn_flip_set_particles(self.h, _ptr(pos), _ptr(vel), pos.shape[0]))

    def generate_scene(self, spec, seed):
        s = spec if isinstance(spec, SceneSpec) else make_spec(spec)
        self._c(self.L.bnn_flip_generate_scene(self.h, C.byref(s), seed))

    def num_particles(self):
        return int(self.L.bnn_flip_num_particles(self.h))

    def particles(self):
        n = self.num_particles()
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        self._c(self.L.bnn_flip_get_particles(self.h, _ptr(pos), _ptr(vel)))
        return pos, vel

    # --- stepping (PIC_Solver::advanceFrame and its override points) ---------------------------
    def step(self, dt):
        return self._c(self.L.bnn_flip_step(self.h, dt))

    def advance_frame(self):
        return self._c(self.L.bnn_flip_advance_frame(self.h))

    def do_frame_simulation(self, frame):
        return self._c(self.L.bnn_flip_do_frame_simulation(self.h, frame))

    def timestep_cfl(self):
        dt = C.c_float()
        self._c(self.L.bnn_flip_timestep_cfl(self.h, C.byref(dt)))
        return float(dt.value)

    def affine(self):
        """APIC / AFLIP: the particles' affine matrices, shape (n, 3, 3), row d = the gradient of v_d."""
        c = np.zeros((self.num_particles(), 3, 3), np.float32)
        self._c(self.L.bnn_flip_get_affine(self.h, _ptr(c)))
        return c

    def set_affine(self, c):
        c = np.ascontiguousarray(c, np.float32)
        assert c.shape == (self.num_particles(), 3, 3)
        self._c(self.L.bnn_flip_set_affine(self.h, _ptr(c)))

    def set_finished_frame(self, frame):
        """GlobalParameters::finishedFrame after a host-side memory-state load: the frame count the
        sort rule and the animated boundary objects follow."""
        self._c(self.L.bnn_flip_set_finished_frame(self.h, int(frame)))

    def set_boundary_objects(self, objs):
        """The scene's additional boundary objects: a list of dicts (see boundary_objects) or a
        ctypes array with its count, as (array, n)."""
        if isinstance(objs, tuple):
            arr, n = objs
        else:
            arr, n = boundary_objects(objs), len(objs)
        self._c(self.L.bnn_flip_set_boundary_objects(self.h, C.cast(arr, C.c_void_p), n))

    def correct_particle_positions(self, dt):
        self._c(self.L.bnn_flip_correct_particle_positions(self.h, dt))

    def move_particles(self, dt):
        self._c(self.L.bnn_flip_move_particles(self.h, dt))

    def sort_particles(self):
        self._c(self.L.bnn_flip_sort_particles(self.h))

    def sort_particles_noalloc(self):
        """The same z-sort with its scratch in the library's static per-device array (no allocation)."""
        self._c(self.L.bnn_flip_sort_particles_noalloc(self.h))

    def advance_velocity(self, dt):
        return self._c(self.L.bnn_flip_advance_velocity(self.h, dt))

    def map_particles_to_grid(self):
        self._c(self.L.bnn_flip_map_particles_to_grid(self.h))

    def map_grid_to_particles(self):
        self._c(self.L.bnn_flip_map_grid_to_particles(self.h))

    def extrapolate_velocity(self):
        self._c(self.L.bnn_flip_extrapolate_velocity(self.h))

    def constrain_grid_velocity(self):
        self._c(self.L.bnn_flip_constrain_grid_velocity(self.h))

    def pressure_projection(self, dt):
        return self._c(self.L.bnn_flip_pressure_projection(self.h, dt))

    def compute_fluid_sdf(self):
        self._c(self.L.bnn_flip_compute_fluid_sdf(self.h))

    def compute_cell_weights(self):
        self._c(self.L.bnn_flip_compute_cell_weights(self.h))

    def bin_particles(self):
        self._c(self.L.bnn_flip_bin_particles(self.h))

    def synchronize(self):
        self._c(self.L.bnn_flip_synchronize(self.h))

    # --- readback ---------------------------------------------------------------------------
    def field_shape(self, name):
        nx, ny, nz = self.n
        if name.endswith("u"):
            return (nz, ny, nx + 1)
        if name.endswith("v"):
            return (nz, ny + 1, nx)
        if name.endswith("w"):
            return (nz + 1, ny, nx)
        if name == "boundary_sdf":
            return (nz + 1, ny + 1, nx + 1)
        return (nz, ny, nx)

    def grid(self, name):
        dtype = np.int8 if name.startswith(("valid", "extrap")) else np.float32
        out = np.zeros(self.field_shape(name), dtype)
        self._c(self.L.bnn_flip_get_grid(self.h, FIELD[name], _ptr(out)))
        return out

    def set_grid(self, name, arr):
        dtype = np.int8 if name.startswith(("valid", "extrap")) else np.float32
        a = np.ascontiguousarray(arr, dtype)
        assert a.shape == self.field_shape(name), (a.shape, self.field_shape(name))
        self._c(self.L.bnn_flip_set_grid(self.h, FIELD[name], _ptr(a)))

    def pressure(self):
        n = C.c_uint64()
        self._c(self.L.bnn_flip_get_pressure(self.h, None, C.byref(n)))
        out = np.zeros(n.value, np.float64)
        self._c(self.L.bnn_flip_get_pressure(self.h, _ptr(out), C.byref(n)))
        return out

    def rhs(self):
        n = C.c_uint64()
        self._c(self.L.bnn_flip_get_rhs(self.h, None, C.byref(n)))
        out = np.zeros(n.value, np.float64)
        self._c(self.L.bnn_flip_get_rhs(self.h, _ptr(out), C.byref(n)))
        return out

    def save_frame_data(self, path, save_velocity=False, save_object_index=False):
        """PIC_Solver::saveFrameData: a BananaParticleData frame file (compressed positions)."""
        self._c(self.L.bnn_flip_save_frame_data(self.h, os.fsencode(path), int(save_velocity), int(save_object_index)))

    def save_memory_state(self, path):
        self._c(self.L.bnn_flip_save_memory_state(self.h, os.fsencode(path)))

    def load_memory_state(self, path, frame):
        self._c(self.L.bnn_flip_load_memory_state(self.h, os.fsencode(path), frame))

    def object_index(self):
        n = self.num_particles()
        out = np.zeros(n, np.uint16)
        nobj = C.c_uint32()
        self._c(self.L.bnn_flip_get_object_index(self.h, _ptr(out), C.byref(nobj)))
        return out, int(nobj.value)

    def set_timing(self, timing):
        """0: no events, 1: stage events, 2: also per-kernel events (from the next step on)."""
        self._c(self.L.bnn_flip_set_timing(self.h, timing))

    def apply_preconditioner(self, r):
        """z = M^-1 r on the last assembled system (active-cell order, like pressure())."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        self._c(self.L.bnn_flip_apply_preconditioner(self.h, _ptr(r), _ptr(z)))
        return z

    def ic0_factor(self):
        n = C.c_uint64()
        self._c(self.L.bnn_flip_get_ic0_factor(self.h, None, C.byref(n)))
        out = np.zeros(n.value, np.float64)
        self._c(self.L.bnn_flip_get_ic0_factor(self.h, _ptr(out), C.byref(n)))
        return out

    def bins(self):
        n = self.num_particles()
        perm = np.zeros(n, np.uint32)
        start = np.zeros(self.n[0] * self.n[1] * self.n[2] + 1, np.uint32)
        self._c(self.L.bnn_flip_get_bins(self.h, _ptr(perm), _ptr(start)))
        return perm, start

    def stats(self):
        s = Stats()
        self._c(self.L.bnn_flip_get_stats(self.h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in Stats._fields_ if k not in ("stage_ms", "stage_bytes", "kernel_ms", "kernel_launches")}
        d["kernel_ms"] = {n: s.kernel_ms[i] for i, n in enumerate(_lib.KERNEL_TIMER_NAMES)}
        d["kernel_launches"] = {n: s.kernel_launches[i] for i, n in enumerate(_lib.KERNEL_TIMER_NAMES)}
        d["stage_ms"] = {n: s.stage_ms[i] for i, n in enumerate(_lib.STAGE_NAMES)}
        d["stage_bytes"] = {n: s.stage_bytes[i] for i, n in enumerate(_lib.STAGE_NAMES)}
        return d
