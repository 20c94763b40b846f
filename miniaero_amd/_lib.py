"""ctypes binding of include/miniaero.h (libminiaero_amd.so, built in-tree).

The library is the product: there is no Python or CPU fallback.  Loading fails
loudly when the shared object is missing.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libminiaero_amd.so")

MA_OK = 0
MA_ERR_INVALID = 1
MA_ERR_UNSUPPORTED = 2
MA_ERR_NOMEM = 3
MA_ERR_DEVICE = 4
MA_ERR_NONFINITE = 5

MA_BC_EXTRAPOLATE, MA_BC_TANGENT, MA_BC_NOSLIP, MA_BC_INFLOW = 1, 2, 3, 4
BC_NAMES = {MA_BC_EXTRAPOLATE: "Extrapolate", MA_BC_TANGENT: "Tangent",
            MA_BC_NOSLIP: "NoSlip", MA_BC_INFLOW: "Inflow"}
BC_KINDS = {v: k for k, v in BC_NAMES.items()}
SIDES = ("bottom", "top", "front", "back", "right", "left")

# Every entry point declared in include/miniaero.h.
EXPORTS = (
    "ma_last_error", "ma_options_read", "ma_mesh_generate", "ma_mesh_view", "ma_mesh_destroy",
    "ma_solver_create", "ma_solver_initialize", "ma_solver_set_state", "ma_solver_step",
    "ma_solver_synchronize", "ma_solver_get_state", "ma_solver_get_residual",
    "ma_solver_get_gradients", "ma_solver_get_time", "ma_solver_check_finite",
    "ma_solver_stream", "ma_solver_destroy", "ma_solver_profile", "ma_solver_sizes",
    "ma_measure_copy", "ma_mesh_plan", "ma_batch_create", "ma_batch_sizes", "ma_batch_get_case_state",
    "ma_batch_get_case_time", "ma_batch_plan", "ma_solver_set_graph", "ma_solver_set_persistent",
    "ma_solver_tuning", "ma_solver_residual_norms", "ma_solver_residual_norms_noalloc", "ma_server_create", "ma_server_submit", "ma_server_poll",
    "ma_server_get_state", "ma_server_release", "ma_server_stats", "ma_server_now", "ma_server_destroy",
    "ma_server_plan_cache",
)


class ma_server_config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_longlong), ("max_active", ctypes.c_int),
                ("setup_threads", ctypes.c_int), ("steps_in_flight", ctypes.c_int),
                ("plan_cache_bytes", ctypes.c_longlong)]


class ma_case_result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("steps", ctypes.c_int),
                ("cells", ctypes.c_longlong), ("device_bytes", ctypes.c_longlong),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_finish", ctypes.c_double)]


class MiniAeroError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"miniaero status {status}: {msg}")
        self.status = status


class ma_options(ctypes.Structure):
    _fields_ = [
        ("problem_type", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("angle", ctypes.c_double),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("ntimesteps", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("output_results", ctypes.c_int),
        ("output_frequency", ctypes.c_int),
        ("second_order_space", ctypes.c_int),
        ("viscous", ctypes.c_int),
        ("bc_kind", ctypes.c_int * 6),
        ("y_stretch", ctypes.c_double),
        ("limiter", ctypes.c_int),
        ("ic_kind", ctypes.c_int),
        ("ic", ctypes.c_double * 5),
        ("use_inflow", ctypes.c_int),
        ("inflow", ctypes.c_double * 5),
        ("prandtl", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("perturb", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
        ("precision", ctypes.c_int),
        ("tile_shape", ctypes.c_int * 3),
        ("autotune", ctypes.c_int),
        ("workgroups_per_cu", ctypes.c_int),
    ]


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class ma_faces(ctypes.Structure):
    _fields_ = [
        ("nfaces", ctypes.c_int),
        ("coordinates", _D), ("face_normal", _D), ("face_tangent", _D), ("face_binormal", _D),
        ("face_cell_conn", _I), ("cell_flux_index", _I), ("permute_vector", _I),
    ]


class ma_bc_group(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("faces", ma_faces)]


class ma_mesh_data(ctypes.Structure):
    _fields_ = [
        ("num_owned_cells", ctypes.c_int),
        ("num_ghosts", ctypes.c_int),
        ("cell_coordinates", _D),
        ("cell_volumes", _D),
        ("internal_faces", ma_faces),
        ("num_boundary_groups", ctypes.c_int),
        ("boundary_faces", ctypes.POINTER(ma_bc_group)),
    ]


_lib = None


def load() -> ctypes.CDLL:
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C miniaero_amd`; there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    PP = ctypes.POINTER(ctypes.c_void_p)
    L.ma_last_error.restype = ctypes.c_char_p
    L.ma_last_error.argtypes = []
    L.ma_options_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ma_options)]
    L.ma_mesh_generate.argtypes = [ctypes.POINTER(ma_options), PP]
    L.ma_mesh_view.argtypes = [P, ctypes.POINTER(ma_mesh_data)]
    L.ma_mesh_destroy.argtypes = [P]
    L.ma_mesh_destroy.restype = None
    L.ma_solver_create.argtypes = [ctypes.POINTER(ma_mesh_data), ctypes.POINTER(ma_options), PP]
    L.ma_solver_initialize.argtypes = [P]
    L.ma_solver_set_state.argtypes = [P, _D]
    L.ma_solver_step.argtypes = [P, ctypes.c_int]
    L.ma_solver_synchronize.argtypes = [P]
    L.ma_solver_get_state.argtypes = [P, _D]
    L.ma_solver_get_residual.argtypes = [P, _D]
    L.ma_solver_get_gradients.argtypes = [P, _D, _D]
    L.ma_solver_get_time.argtypes = [P, _D, _D]
    L.ma_solver_check_finite.argtypes = [P, ctypes.POINTER(ctypes.c_longlong)]
    L.ma_solver_stream.argtypes = [P, PP]
    L.ma_solver_destroy.argtypes = [P]
    L.ma_solver_destroy.restype = None
    L.ma_solver_profile.restype = ctypes.c_int
    L.ma_solver_profile.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _I, _D, ctypes.c_int]
    L.ma_solver_sizes.argtypes = [P, ctypes.POINTER(ctypes.c_longlong)]
    L.ma_measure_copy.argtypes = [ctypes.c_longlong, ctypes.c_int, _D]
    L.ma_mesh_plan.argtypes = [ctypes.POINTER(ma_mesh_data), ctypes.POINTER(ctypes.c_longlong)]
    L.ma_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.POINTER(ma_mesh_data)),
                                  ctypes.POINTER(ma_options), PP]
    L.ma_batch_plan.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.POINTER(ma_mesh_data)),
                                ctypes.POINTER(ctypes.c_longlong)]
    L.ma_solver_set_graph.argtypes = [P, ctypes.c_int]
    L.ma_solver_set_persistent.argtypes = [P, ctypes.c_int]
    L.ma_solver_tuning.argtypes = [P, _I, _I, _D]
    L.ma_solver_residual_norms.argtypes = [P, _D]
    L.ma_solver_residual_norms_noalloc.argtypes = [P, P]
    L.ma_server_create.argtypes = [ctypes.POINTER(ma_server_config), PP]
    L.ma_server_submit.argtypes = [P, ctypes.POINTER(ma_options), ctypes.POINTER(ctypes.c_longlong)]
    L.ma_server_poll.restype = ctypes.c_int
    L.ma_server_poll.argtypes = [P, ctypes.POINTER(ma_case_result), ctypes.c_int, ctypes.c_double]
    L.ma_server_get_state.argtypes = [P, ctypes.c_longlong, _D]
    L.ma_server_release.argtypes = [P, ctypes.c_longlong]
    L.ma_server_stats.argtypes = [P, ctypes.POINTER(ctypes.c_longlong)]
    L.ma_server_now.restype = ctypes.c_double
    L.ma_server_now.argtypes = [P]
    L.ma_server_plan_cache.argtypes = [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.ma_server_destroy.argtypes = [P]
    L.ma_server_destroy.restype = None
    L.ma_batch_sizes.argtypes = [P, _I, _I]
    L.ma_batch_get_case_state.argtypes = [P, ctypes.c_int, _D]
    L.ma_batch_get_case_time.argtypes = [P, ctypes.c_int, _D, _D]
    _lib = L
    return L


def measure_copy(nbytes: int, reps: int = 20) -> float:
    """Mean milliseconds of one device-to-device copy of nbytes by a plain
    copy kernel (reads and writes nbytes)."""
    ms = ctypes.c_double(0.0)
    check(load().ma_measure_copy(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def check(status: int) -> None:
    if status != MA_OK:
        raise MiniAeroError(status, load().ma_last_error().decode())
