"""banana_amd: MI355X (gfx950) FLIP substep of Banana's FLIP_Solver<3,float>, behind a C ABI.

The compute path is libbnn_flip.so (hand-written HIP kernels); this package holds the build step,
the ctypes view of the ABI and scene descriptions. There is no CPU fallback.
"""
from ._lib import BnnError, load  # noqa: F401
from .solver import (FLIPSolver, boundaries_from_json, boundary_objects, data_file_path, generate_scene_host, io_from_json, latest_state_index,  # noqa: F401
                     make_params, make_spec)
