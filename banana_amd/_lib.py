"""ctypes view of the C ABI in include/bnn_flip.h (libbnn_flip.so, built in-tree for gfx950).

Loading fails loudly when the library is missing: there is no CPU fallback for any stage.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbnn_flip.so")

BNN_OK = 0
STATUS = {0: "BNN_OK", 1: "BNN_ERR_INVALID_ARG", 2: "BNN_ERR_NOT_CONVERGED", 3: "BNN_ERR_DEVICE",
          4: "BNN_ERR_OUT_OF_MEMORY", 5: "BNN_ERR_UNSUPPORTED", 6: "BNN_ERR_STATE", 7: "BNN_ERR_REJECTED"}
PRESSURE_FP32, PRESSURE_REFINED = 0, 1
PRECOND_JACOBI, PRECOND_MG, PRECOND_IC0 = 0, 1, 2
TRANSFER_FLIP, TRANSFER_APIC, TRANSFER_AFLIP = 0, 1, 2
NUM_STAGES = 12
NUM_KERNEL_TIMERS = 13
KERNEL_TIMER_NAMES = ["pcg_spmv", "pcg_update", "p2g", "fluid_sdf", "g2p", "move", "mg_pre", "mg_restrict", "mg_prolong",
                      "mg_post", "mg_coarse", "extrap", "permute"]
STAGE_NAMES = ["move", "bin", "p2g", "extrapolate", "constrain", "weights", "phi", "assembly", "pcg", "update",
               "g2p", "other"]

FIELD = dict(u=0, v=1, w=2, weight_u=3, weight_v=4, weight_w=5, valid_u=6, valid_v=7, valid_w=8,
             extrap_u=9, extrap_v=10, extrap_w=11, old_u=12, old_v=13, old_w=14, phi=15, boundary_sdf=16)


class Params(C.Structure):
    _fields_ = [
        ("box_min", C.c_float * 3), ("box_max", C.c_float * 3),
        ("cell_size", C.c_float), ("ratio_cell_size_pradius", C.c_float), ("n_expand_cells", C.c_uint32),
        ("min_timestep", C.c_float), ("max_timestep", C.c_float), ("cfl_factor", C.c_float),
        ("frame_duration", C.c_float), ("cg_relative_tolerance", C.c_float), ("max_cg_iteration", C.c_uint32),
        ("pic_flip_ratio", C.c_float), ("advection_steps", C.c_uint32), ("gravity_type", C.c_int32),
        ("gravity_direction", C.c_float * 3), ("gravity_center", C.c_float * 3), ("apply_gravity", C.c_int32),
        ("reflect_velocity_at_boundary", C.c_int32), ("boundary_reflection_multiplier", C.c_float),
        ("boundary_sdf_at_init", C.c_int32), ("correct_position", C.c_int32),
        ("exit_if_pressure_projection_failed", C.c_int32), ("enable_sort_particle", C.c_int32),
        ("sort_frequency", C.c_uint32), ("pressure_mode", C.c_int32), ("refine_tolerance", C.c_double),
        ("preconditioner", C.c_int32), ("timing", C.c_int32), ("mg_sweeps", C.c_int32), ("mg_omega", C.c_float),
        ("mg_coarsest_sweeps", C.c_int32), ("start_frame", C.c_uint32), ("final_frame", C.c_uint32),
        ("repulsive_force_stiffness", C.c_float), ("random_seed", C.c_uint64), ("refine_inner_tolerance", C.c_float),
        ("transfer_scheme", C.c_int32), ("warm_start", C.c_int32),
    ]


class KeyFrame(C.Structure):
    _fields_ = [("frame", C.c_uint32), ("translation", C.c_float * 3), ("rotation_axis_angle", C.c_float * 4), ("scale", C.c_float),
                ("has_rotation", C.c_int32), ("has_scale", C.c_int32)]


class BoundaryObject(C.Structure):
    _fields_ = [("geometry", C.c_int32), ("box_min", C.c_float * 3), ("box_max", C.c_float * 3), ("translation", C.c_float * 3),
                ("rotation_axis_angle", C.c_float * 4), ("scale", C.c_float), ("has_translation", C.c_int32),
                ("has_rotation", C.c_int32), ("has_scale", C.c_int32), ("n_keyframes", C.c_uint32), ("keyframes", KeyFrame * 16),
                ("periodic", C.c_int32), ("start_frame", C.c_uint32), ("cubic_translation", C.c_int32),
                ("cubic_rotation", C.c_int32), ("cubic_scale", C.c_int32)]


class SceneSpec(C.Structure):
    _fields_ = [
        ("n_boxes", C.c_int32), ("box_min", (C.c_float * 3) * 8), ("box_max", (C.c_float * 3) * 8),
        ("box_vel", (C.c_float * 3) * 8), ("n_spheres", C.c_int32), ("sphere_center", (C.c_float * 3) * 8),
        ("sphere_radius", C.c_float * 8), ("sphere_vel", (C.c_float * 3) * 8), ("vel_sigma", C.c_float),
        ("ppa", C.c_int32), ("jitter", C.c_float),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("num_particles", C.c_uint64), ("num_cells", C.c_uint64), ("num_active_cells", C.c_uint64),
        ("num_active_bricks", C.c_uint64), ("pcg_iterations", C.c_uint32), ("outer_iterations", C.c_uint32),
        ("pcg_residual", C.c_double), ("rhs_max", C.c_double), ("pcg_success", C.c_int32),
        ("last_substep", C.c_float), ("substep_count", C.c_uint32),
        ("stage_ms", C.c_double * NUM_STAGES), ("stage_bytes", C.c_double * NUM_STAGES), ("steps", C.c_uint64),
        ("finished_frame", C.c_uint32),
        ("kernel_ms", C.c_double * NUM_KERNEL_TIMERS), ("kernel_launches", C.c_uint64 * NUM_KERNEL_TIMERS),
    ]


class IoParams(C.Structure):
    _fields_ = [("data_path", C.c_char * 512), ("memory_state_folder", C.c_char * 128), ("frame_data_folder", C.c_char * 128),
                ("load_memory_state", C.c_int32), ("save_memory_state", C.c_int32), ("save_frame_data", C.c_int32),
                ("frame_per_state", C.c_uint32), ("save_velocity", C.c_int32), ("save_object_index", C.c_int32)]


class BatchShape(C.Structure):
    _fields_ = [("p2g_segment", C.c_int32), ("pcg_chunk_first", C.c_uint32), ("pcg_chunk_max", C.c_uint32),
                ("items_per_group", C.c_uint32), ("min_groups", C.c_uint32), ("graphs", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [("shape", BatchShape), ("flushes", C.c_uint64), ("launches", C.c_uint64), ("scene_launches", C.c_uint64),
                ("graph_replays", C.c_uint64), ("graph_captures", C.c_uint64), ("mean_scenes_per_flush", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch", C.c_uint32), ("workers", C.c_uint32),
                ("batching", C.c_int32), ("barrier_rounds", C.c_int32), ("max_batch_mcells", C.c_uint32),
                ("reserved", C.c_int32 * 3)]


class RequestResult(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("scene_id", C.c_uint64), ("kind", C.c_int32), ("status", C.c_int32),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("batch_size", C.c_uint32), ("pcg_iterations", C.c_uint32), ("pcg_success", C.c_int32),
                ("pad", C.c_int32), ("pcg_residual", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64), ("batches", C.c_uint64),
                ("scenes_resident", C.c_uint64), ("scenes_waiting", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("peak_resident_bytes", C.c_uint64), ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("latency_mean_ms", C.c_double), ("latency_max_ms", C.c_double), ("throughput_rps", C.c_double),
                ("mean_batch", C.c_double), ("busy_s", C.c_double)]


REQ_STEP, REQ_PRESSURE = 0, 1

# every exported symbol of include/bnn_flip.h: name -> (restype, argtypes)
_vp, _f32, _u32, _u64, _i32, _P = C.c_void_p, C.c_float, C.c_uint32, C.c_uint64, C.c_int32, C.POINTER
SIGNATURES = {
    "bnn_flip_default_params": (None, [_P(Params)]),
    "bnn_flip_create": (_i32, [_P(Params), _P(_vp)]),
    "bnn_flip_create_from_json": (_i32, [C.c_char_p, _u64, _P(_vp)]),
    "bnn_scene_from_json": (_i32, [C.c_char_p, _P(Params), _P(SceneSpec), C.c_char_p, C.c_size_t]),
    "bnn_flip_destroy": (_i32, [_vp]),
    "bnn_flip_last_error": (C.c_char_p, [_vp]),
    "bnn_flip_get_params": (_i32, [_vp, _P(Params)]),
    "bnn_flip_grid_dims": (_i32, [_vp, _vp]),
    "bnn_flip_grid_origin": (_i32, [_vp, _vp, _P(_f32)]),
    "bnn_flip_set_particles": (_i32, [_vp, _vp, _vp, _u64]),
    "bnn_flip_generate_scene": (_i32, [_vp, _P(SceneSpec), _u64]),
    "bnn_generate_scene_host": (_i32, [_P(Params), _P(SceneSpec), _u64, _vp, _vp, _u64, _P(_u64)]),
    "bnn_flip_num_particles": (_u64, [_vp]),
    "bnn_flip_get_particles": (_i32, [_vp, _vp, _vp]),
    "bnn_flip_step": (_i32, [_vp, _f32]),
    "bnn_flip_advance_frame": (_i32, [_vp]),
    "bnn_flip_do_frame_simulation": (_i32, [_vp, _u32]),
    "bnn_flip_timestep_cfl": (_i32, [_vp, _P(_f32)]),
    "bnn_flip_move_particles": (_i32, [_vp, _f32]),
    "bnn_flip_correct_particle_positions": (_i32, [_vp, _f32]),
    "bnn_flip_get_affine": (_i32, [_vp, _vp]),
    "bnn_flip_set_affine": (_i32, [_vp, _vp]),
    "bnn_flip_sort_particles": (_i32, [_vp]),
    "bnn_flip_sort_particles_noalloc": (_i32, [_vp]),
    "bnn_flip_advance_velocity": (_i32, [_vp, _f32]),
    "bnn_flip_map_particles_to_grid": (_i32, [_vp]),
    "bnn_flip_map_grid_to_particles": (_i32, [_vp]),
    "bnn_flip_extrapolate_velocity": (_i32, [_vp]),
    "bnn_flip_constrain_grid_velocity": (_i32, [_vp]),
    "bnn_flip_pressure_projection": (_i32, [_vp, _f32]),
    "bnn_flip_compute_fluid_sdf": (_i32, [_vp]),
    "bnn_flip_bin_particles": (_i32, [_vp]),
    "bnn_flip_compute_cell_weights": (_i32, [_vp]),
    "bnn_flip_get_grid": (_i32, [_vp, _i32, _vp]),
    "bnn_flip_get_pressure": (_i32, [_vp, _vp, _P(_u64)]),
    "bnn_flip_get_rhs": (_i32, [_vp, _vp, _P(_u64)]),
    "bnn_flip_apply_preconditioner": (_i32, [_vp, _vp, _vp]),
    "bnn_flip_get_ic0_factor": (_i32, [_vp, _vp, _P(_u64)]),
    "bnn_flip_get_bins": (_i32, [_vp, _vp, _vp]),
    "bnn_flip_get_stats": (_i32, [_vp, _P(Stats)]),
    "bnn_flip_set_grid": (_i32, [_vp, _i32, _vp]),
    "bnn_flip_synchronize": (_i32, [_vp]),
    "bnn_flip_set_timing": (_i32, [_vp, _i32]),
    "bnn_data_file_path": (_i32, [C.c_char_p, C.c_char_p, _i32, C.c_char_p, C.c_size_t]),
    "bnn_latest_state_index": (_i32, [C.c_char_p, C.c_char_p, _u32, _u32]),
    "bnn_flip_save_frame_data": (_i32, [_vp, C.c_char_p, _i32, _i32]),
    "bnn_flip_save_memory_state": (_i32, [_vp, C.c_char_p]),
    "bnn_flip_load_memory_state": (_i32, [_vp, C.c_char_p, _u32]),
    "bnn_flip_get_object_index": (_i32, [_vp, _vp, _P(_u32)]),
    "bnn_io_default_params": (None, [_P(IoParams)]),
    "bnn_io_from_json": (_i32, [C.c_char_p, _P(IoParams)]),
    "bnn_flip_footprint": (_i32, [_P(Params), _u64, _P(_u64)]),
    "bnn_flip_batch_create": (_i32, [_P(_vp)]),
    "bnn_flip_batch_destroy": (_i32, [_vp]),
    "bnn_flip_batch_last_error": (C.c_char_p, [_vp]),
    "bnn_flip_batch_set_timing": (_i32, [_vp, _i32]),
    "bnn_flip_batch_step": (_i32, [_vp, _P(_vp), _P(_f32), _u32, _P(_i32)]),
    "bnn_flip_batch_advance_frame": (_i32, [_vp, _P(_vp), _u32, _P(_i32)]),
    "bnn_flip_batch_pressure_projection": (_i32, [_vp, _P(_vp), _P(_f32), _u32, _P(_i32)]),
    "bnn_flip_batch_set_shape": (_i32, [_vp, _vp]),
    "bnn_flip_set_boundary_objects": (_i32, [_vp, _vp, _u32]),
    "bnn_flip_set_finished_frame": (_i32, [_vp, _u32]),
    "bnn_boundaries_from_json": (_i32, [C.c_char_p, _vp, _u32, _P(_u32)]),
    "bnn_flip_batch_get_info": (_i32, [_vp, _vp]),
    "bnn_server_create": (_i32, [_P(ServerConfig), _P(_vp)]),
    "bnn_server_destroy": (_i32, [_vp]),
    "bnn_server_last_error": (C.c_char_p, [_vp]),
    "bnn_server_now": (C.c_double, [_vp]),
    "bnn_server_add_scene": (_i32, [_vp, _P(Params), _P(SceneSpec), _u64, _P(_u64)]),
    "bnn_server_release_scene": (_i32, [_vp, _u64]),
    "bnn_server_submit": (_i32, [_vp, _u64, _i32, _f32, _P(_u64)]),
    "bnn_server_wait": (_i32, [_vp, _u64, C.c_double, _P(RequestResult)]),
    "bnn_server_drain": (_i32, [_vp, C.c_double]),
    "bnn_server_get_particles": (_i32, [_vp, _u64, _vp, _vp, _P(_u64)]),
    "bnn_server_get_stats": (_i32, [_vp, _P(ServerStats)]),
    "bnn_server_reset_stats": (_i32, [_vp]),
}

_lib = None


class BnnError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def load():
    """Loads libbnn_flip.so; raises if it is absent (build with banana_amd.build.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python banana_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
