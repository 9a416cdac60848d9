"""ctypes bindings for the CPU oracle (oracle/flip_oracle.cpp).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liborc.so")

# field ids (orc::Field)
F_U, F_V, F_W = 0, 1, 2
F_WEIGHT_U, F_WEIGHT_V, F_WEIGHT_W = 3, 4, 5
F_VALID_U, F_VALID_V, F_VALID_W = 6, 7, 8
F_EXTRAP_U, F_EXTRAP_V, F_EXTRAP_W = 9, 10, 11
F_OLD_U, F_OLD_V, F_OLD_W = 12, 13, 14
F_PHI, F_BOUNDARY_SDF, F_ACTIVE_IDX = 15, 16, 17
F_TMP_U, F_TMP_V, F_TMP_W = 18, 19, 20
FLAG_FIELDS = {F_VALID_U, F_VALID_V, F_VALID_W, F_EXTRAP_U, F_EXTRAP_V, F_EXTRAP_W}


class OrcParams(C.Structure):
    _fields_ = [
        ("box_min", C.c_float * 3), ("box_max", C.c_float * 3),
        ("cell_size", C.c_float), ("ratio_cell_size_pradius", C.c_float), ("n_expand_cells", C.c_uint32),
        ("min_timestep", C.c_float), ("max_timestep", C.c_float), ("cfl_factor", C.c_float),
        ("frame_duration", C.c_float), ("cg_relative_tolerance", C.c_float), ("max_cg_iteration", C.c_uint32),
        ("pic_flip_ratio", C.c_float), ("advection_steps", C.c_uint32), ("gravity_type", C.c_int32),
        ("gravity_direction", C.c_float * 3), ("gravity_center", C.c_float * 3), ("apply_gravity", C.c_int32),
        ("reflect_velocity_at_boundary", C.c_int32), ("boundary_reflection_multiplier", C.c_float),
        ("boundary_sdf_at_init", C.c_int32),
    ]


class OrcSceneSpec(C.Structure):
    _fields_ = [
        ("n_boxes", C.c_int32), ("box_min", (C.c_float * 3) * 8), ("box_max", (C.c_float * 3) * 8),
        ("box_vel", (C.c_float * 3) * 8), ("n_spheres", C.c_int32), ("sphere_center", (C.c_float * 3) * 8),
        ("sphere_radius", C.c_float * 8), ("sphere_vel", (C.c_float * 3) * 8), ("vel_sigma", C.c_float),
        ("ppa", C.c_int32), ("jitter", C.c_float),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        vp, f32, u64, i32, u32 = C.c_void_p, C.c_float, C.c_uint64, C.c_int32, C.c_uint32
        P = C.POINTER
        L.orc_create.restype = vp
        L.orc_create.argtypes = [P(OrcParams), i32, i32]
        L.orc_destroy.argtypes = [vp]
        L.orc_generate_scene.restype = u64
        L.orc_generate_scene.argtypes = [P(OrcParams), P(OrcSceneSpec), u64, vp, vp]
        L.orc_grid_dims.argtypes = [vp, vp]
        L.orc_set_precond.argtypes = [vp, i32]
        L.orc_set_cg.argtypes = [vp, C.c_double, u32]
        L.orc_set_particles.argtypes = [vp, vp, vp, u64]
        L.orc_num_particles.restype = u64
        L.orc_num_particles.argtypes = [vp]
        L.orc_get_particles.argtypes = [vp, vp, vp]
        L.orc_get_field.restype = u64
        L.orc_get_field.argtypes = [vp, i32, vp]
        L.orc_set_field.restype = u64
        L.orc_set_field.argtypes = [vp, i32, vp]
        L.orc_num_active.restype = u32
        L.orc_num_active.argtypes = [vp]
        L.orc_get_system.argtypes = [vp, vp, vp]
        L.orc_get_matrix.argtypes = [vp, vp, vp]
        L.orc_pcg_stats.argtypes = [vp, P(u32), P(C.c_double), P(i32)]
        L.orc_ic0_factor.argtypes = [vp, vp]
        L.orc_correct_particle_positions.argtypes = [vp, f32, f32, u64]
        L.orc_set_scheme.argtypes = [vp, i32]
        L.orc_set_affine.argtypes = [vp, vp]
        L.orc_get_affine.argtypes = [vp, vp]
        L.orc_advance_velocity.restype = i32
        L.orc_advance_velocity.argtypes = [vp, f32]
        L.orc_apply_precond.argtypes = [vp, vp, vp]
        L.orc_step.restype = i32
        L.orc_step.argtypes = [vp, f32]
        L.orc_advance_frame.restype = i32
        L.orc_advance_frame.argtypes = [vp, P(f32), P(u32)]
        for name in ["orc_map_particles_to_grid", "orc_extrapolate_velocity", "orc_constrain_grid_velocity",
                     "orc_backup_grid_velocity", "orc_compute_cell_weights", "orc_compute_fluid_sdf",
                     "orc_compute_changes_grid_velocity", "orc_map_grid_to_particles", "orc_compute_boundary_sdf"]:
            getattr(L, name).argtypes = [vp]
        for name in ["orc_move_particles", "orc_add_gravity", "orc_compute_system", "orc_update_projected_velocity"]:
            getattr(L, name).argtypes = [vp, f32]
        L.orc_solve_system.restype = i32
        L.orc_solve_system.argtypes = [vp]
        L.orc_timestep_cfl.restype = f32
        L.orc_timestep_cfl.argtypes = [vp]
        L.orc_bin_particles.argtypes = [vp, u64, vp, f32, vp, vp, vp, vp]
        L.orc_cell_index.argtypes = [vp, u64, vp, f32, vp]
        L.orc_morton_keys.argtypes = [vp, u64, f32, vp]
        L.orc_morton_sort.argtypes = [vp, u64, f32, vp]
        L.orc_set_boundary_objects.argtypes = [vp, vp, u32]
        L.orc_advance_scene.argtypes = [vp, f32]
        L.orc_set_finished_frame.argtypes = [vp, u32]
        L.orc_boundary_sd.argtypes = [vp, i32, vp, u64, vp]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_params(d):
    p = OrcParams()
    for k, v in d.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def make_spec(d):
    s = OrcSceneSpec()
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


def generate_scene(params, spec, seed):
    L = lib()
    p, s = make_params(params), make_spec(spec)
    n = L.orc_generate_scene(C.byref(p), C.byref(s), seed, None, None)
    pos = np.zeros((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32)
    L.orc_generate_scene(C.byref(p), C.byref(s), seed, ptr(pos), ptr(vel))
    return pos, vel


class Oracle:
    """Handle on one oracle solver instance (precision 32 = reference parity mode, 64 = FP64 oracle)."""

    def __init__(self, params, precision=32, threads=1):
        self.L = lib()
        self.params = params
        self._p = make_params(params)
        self.h = self.L.orc_create(C.byref(self._p), precision, threads)
        n = np.zeros(3, np.int32)
        self.L.orc_grid_dims(self.h, ptr(n))
        self.n = tuple(int(x) for x in n)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def field_shape(self, f):
        nx, ny, nz = self.n
        if f in (F_U, F_WEIGHT_U, F_VALID_U, F_EXTRAP_U, F_OLD_U, F_TMP_U):
            return (nz, ny, nx + 1)
        if f in (F_V, F_WEIGHT_V, F_VALID_V, F_EXTRAP_V, F_OLD_V, F_TMP_V):
            return (nz, ny + 1, nx)
        if f in (F_W, F_WEIGHT_W, F_VALID_W, F_EXTRAP_W, F_OLD_W, F_TMP_W):
            return (nz + 1, ny, nx)
        if f == F_BOUNDARY_SDF:
            return (nz + 1, ny + 1, nx + 1)
        return (nz, ny, nx)

    def set_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        self.L.orc_set_particles(self.h, ptr(pos), ptr(vel), pos.shape[0])

    def particles(self):
        n = self.L.orc_num_particles(self.h)
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        self.L.orc_get_particles(self.h, ptr(pos), ptr(vel))
        return pos, vel

    def field(self, f):
        shape = self.field_shape(f)
        if f in FLAG_FIELDS:
            out = np.zeros(shape, np.int8)
        elif f == F_ACTIVE_IDX:
            out = np.zeros(shape, np.uint32)
        else:
            out = np.zeros(shape, np.float64)
        self.L.orc_get_field(self.h, f, ptr(out))
        return out

    def set_field(self, f, arr):
        if f in FLAG_FIELDS:
            a = np.ascontiguousarray(arr, np.int8)
        elif f == F_ACTIVE_IDX:
            a = np.ascontiguousarray(arr, np.uint32)
        else:
            a = np.ascontiguousarray(arr, np.float64)
        assert a.shape == self.field_shape(f)
        self.L.orc_set_field(self.h, f, ptr(a))

    def num_active(self):
        return int(self.L.orc_num_active(self.h))

    def system(self):
        na = self.num_active()
        rhs = np.zeros(na, np.float64)
        pr = np.zeros(na, np.float64)
        self.L.orc_get_system(self.h, ptr(rhs), ptr(pr))
        return rhs, pr

    def matrix(self):
        na = self.num_active()
        cols = np.zeros((na, 7), np.int32)
        vals = np.zeros((na, 7), np.float64)
        self.L.orc_get_matrix(self.h, ptr(cols), ptr(vals))
        return cols, vals

    def num_particles(self):
        return int(self.L.orc_num_particles(self.h))

    def set_scheme(self, scheme):
        """0 FLIP, 1 APIC, 2 AFLIP"""
        self.L.orc_set_scheme(self.h, scheme)

    def affine(self):
        c = np.zeros((self.num_particles(), 3, 3), np.float32)
        self.L.orc_get_affine(self.h, ptr(c))
        return c

    def set_affine(self, c):
        c = np.ascontiguousarray(c, np.float32)
        assert c.shape == (self.num_particles(), 3, 3)
        self.L.orc_set_affine(self.h, ptr(c))

    def ic0_factor(self):
        inv = np.zeros(self.num_active(), np.float64)
        self.L.orc_ic0_factor(self.h, ptr(inv))
        return inv

    def apply_precond(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        self.L.orc_apply_precond(self.h, ptr(r), ptr(z))
        return z

    def pcg_stats(self):
        it, res, ok = C.c_uint32(), C.c_double(), C.c_int32()
        self.L.orc_pcg_stats(self.h, C.byref(it), C.byref(res), C.byref(ok))
        return int(it.value), float(res.value), bool(ok.value)

    def step(self, dt):
        return self.L.orc_step(self.h, dt)

    def advance_frame(self):
        sub, cnt = C.c_float(), C.c_uint32()
        r = self.L.orc_advance_frame(self.h, C.byref(sub), C.byref(cnt))
        return r, float(sub.value), int(cnt.value)

    def set_precond(self, kind):
        self.L.orc_set_precond(self.h, kind)

    def set_cg(self, tol, max_iters):
        self.L.orc_set_cg(self.h, tol, max_iters)

    def set_boundary_objects(self, arr, n):
        """arr: a ctypes array of the library's bnn_boundary_object (the oracle's struct has the
        same layout)."""
        self.L.orc_set_boundary_objects(self.h, C.cast(arr, C.c_void_p), n)

    def boundary_sd(self, group, pts):
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.zeros(pts.shape[0], np.float64)
        self.L.orc_boundary_sd(self.h, group, ptr(pts), pts.shape[0], ptr(out))
        return out

    def __getattr__(self, name):
        # stage calls: o.map_particles_to_grid(), o.add_gravity(dt), ...
        fn = getattr(self.L, "orc_" + name)
        return lambda *a: fn(self.h, *a)


def bin_particles(pos, bmin, dx, n):
    L = lib()
    pos = np.ascontiguousarray(pos, np.float32)
    np_ = pos.shape[0]
    C_ = n[0] * n[1] * n[2]
    key = np.zeros(np_, np.uint32)
    perm = np.zeros(np_, np.uint32)
    start = np.zeros(C_ + 1, np.uint32)
    bm = np.asarray(bmin, np.float32)
    nn = np.asarray(n, np.int32)
    L.orc_bin_particles(ptr(pos), np_, ptr(bm), dx, ptr(nn), ptr(key), ptr(perm), ptr(start))
    return key, perm, start


def cell_index(pos, bmin, dx):
    L = lib()
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.zeros((pos.shape[0], 3), np.int32)
    bm = np.asarray(bmin, np.float32)
    L.orc_cell_index(ptr(pos), pos.shape[0], ptr(bm), dx, ptr(out))
    return out


def morton_keys(pos, inv_cell_size):
    L = lib()
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.zeros(pos.shape[0], np.uint64)
    L.orc_morton_keys(ptr(pos), pos.shape[0], inv_cell_size, ptr(out))
    return out


def morton_sort(pos, inv_cell_size):
    L = lib()
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.zeros(pos.shape[0], np.uint32)
    L.orc_morton_sort(ptr(pos), pos.shape[0], inv_cell_size, ptr(out))
    return out
