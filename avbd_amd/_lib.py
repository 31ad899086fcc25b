"""ctypes binding of the avbd C ABI (include/avbd.h).

The shared library is built in-tree by ``avbd_amd.build`` (``__graft_entry__.build()``).
There is no fallback: if the library is missing or cannot be loaded, importing
this module raises, so nothing ever runs on a silent CPU path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libavbd_amd.so"

AVBD_OK = 0
AVBD_E_ARG = -1
AVBD_E_CAPACITY = -2
AVBD_E_COLORS = -3
AVBD_E_UNSUPPORTED_CONTACT = -4
AVBD_E_NONFINITE = -5
AVBD_E_HIP = -6
AVBD_E_NOT_BUILT = -7

# EnergyType order of the reference (common/common.hpp:156-163)
NEO_HOOKEAN = 0
COSSERAT_ROD = 1
GROUND_COLLISION = 2
RIGID_BODY_GROUND_COLLISION = 3
TRIANGLE_RIGID_COLLISION = 4
TRIANGLE_ROD_COLLISION = 5
ONE_SIDED_FIXED_JOINT = 6
FIXED_JOINT = 7   # two-sided fixed joint (new)

FLAG_NO_GRAPH = 1
FLAG_SERIAL_BROADPHASE = 2
FLAG_NO_PERSISTENT = 4
FLAG_PERSISTENT = 8
FLAG_SEPARATE_CHEBYSHEV = 16
FLAG_FP64_STATE = 32
FLAG_HOST_COLORING = 64


class SceneDesc(C.Structure):
    _fields_ = [
        ("particle_capacity", C.c_uint32),
        ("energy_capacity", C.c_uint32),
        ("dt", C.c_double),
        ("g_accel", C.c_double),
        ("solver_iters", C.c_uint32),
        ("iter_acceleration", C.c_double),
        ("flags", C.c_uint32),
        ("device", C.c_int32),
        ("lanes_per_vertex", C.c_uint32),
        ("block_size", C.c_uint32),
        ("gather_batch", C.c_uint32),
    ]


class NeoHookeanRecord(C.Structure):
    _fields_ = [
        ("particle_indices", C.c_uint32 * 4),
        ("lambda_", C.c_double),
        ("mu", C.c_double),
        ("kd", C.c_double),
        ("Q", C.c_double * 9),
        ("rest_volume", C.c_double),
        ("E_prev", C.c_double * 9),
    ]


class GroundRecord(C.Structure):
    _fields_ = [
        ("particle_index", C.c_uint32),
        ("use_static", C.c_uint32),
        ("unclamped_lambda_tb", C.c_uint32),
        ("pad_", C.c_uint32),
        ("k_start", C.c_double),
        ("k_n", C.c_double),
        ("k_t", C.c_double),
        ("k_b", C.c_double),
        ("lambda_n", C.c_double),
        ("lambda_t", C.c_double),
        ("lambda_b", C.c_double),
        ("C_n_prev", C.c_double),
        ("mu_s", C.c_double),
        ("mu_k", C.c_double),
        ("cp_x", C.c_double),
        ("cp_z", C.c_double),
    ]


# CollisionGeometryType of the reference (collision/CollisionGeometryType.hpp)
SHAPE_ROD_SEGMENT = 1
SHAPE_SPHERE = 2
SHAPE_BOX = 3
SHAPE_CAPSULE = 4


class CosseratRecord(C.Structure):
    _fields_ = [
        ("particle_indices", C.c_uint32 * 2),
        ("stiffness", C.c_double * 6),
        ("rest_length", C.c_double),
        ("rest_curvature", C.c_double * 3),
    ]


class RigidGroundRecord(C.Structure):
    _fields_ = [
        ("particle_index", C.c_uint32),
        ("shape", C.c_uint32),
        ("use_static", C.c_uint32),
        ("unclamped_lambda_tb", C.c_uint32),
        ("radius", C.c_double),
        ("cp_rb_local", C.c_double * 3),
        ("cp_ground", C.c_double * 3),
        ("k_start", C.c_double),
        ("k_n", C.c_double),
        ("k_t", C.c_double),
        ("k_b", C.c_double),
        ("lambda_n", C.c_double),
        ("lambda_t", C.c_double),
        ("lambda_b", C.c_double),
        ("C_n_prev", C.c_double),
        ("mu_s", C.c_double),
        ("mu_k", C.c_double),
    ]


class JointRecord(C.Structure):
    _fields_ = [
        ("particle_index", C.c_uint32),
        ("pad_", C.c_uint32),
        ("position_offset", C.c_double * 3),
        ("rotation_offset", C.c_double * 4),
        ("ref_position", C.c_double * 3),
        ("ref_rotation", C.c_double * 4),
        ("k", C.c_double * 6),
        ("lambda_", C.c_double * 6),
        ("C_prev", C.c_double * 6),
    ]


class FixedJointRecord(C.Structure):
    """avbd_fixed_joint_record: two-sided fixed joint between two oriented particles."""
    _fields_ = [
        ("particle_index", C.c_uint32 * 2),
        ("position_offset", (C.c_double * 3) * 2),
        ("rotation_offset", (C.c_double * 4) * 2),
        ("k", C.c_double * 6),
        ("lambda_", C.c_double * 6),
        ("C_prev", C.c_double * 6),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("residual_abs", C.c_double),
        ("residual_rel", C.c_double),
        ("incremental_potential", C.c_double),
        ("num_colors", C.c_uint32),
        ("num_particles", C.c_uint32),
        ("num_tets", C.c_uint32),
        ("num_ground", C.c_uint32),
        ("steps_taken", C.c_uint64),
        ("vertex_updates", C.c_uint64),
        ("last_step_ms", C.c_double),
        ("iterations_total", C.c_uint64),
        ("last_step_iterations", C.c_uint32),
        ("pad_", C.c_uint32),
        ("warm_start_predictions", C.c_uint64),
        ("warm_start_fallbacks", C.c_uint64),
        ("steps_at_budget", C.c_uint64),
    ]


class Contact(C.Structure):
    """avbd_contact: one dynamic triangle-rigid / triangle-rod contact."""
    _fields_ = [
        ("slot", C.c_uint32),
        ("idx", C.c_uint32 * 5),
        ("n", C.c_double * 3),
        ("t", C.c_double * 3),
        ("b", C.c_double * 3),
        ("barys", C.c_double * 3),
        ("s", C.c_double),
        ("cp_local", C.c_double * 3),
        ("k_start", C.c_double),
        ("k_n", C.c_double),
        ("k_t", C.c_double),
        ("k_b", C.c_double),
        ("lambda_n", C.c_double),
        ("lambda_t", C.c_double),
        ("lambda_b", C.c_double),
        ("C_n_prev", C.c_double),
        ("use_static", C.c_uint8),
        ("unclamped", C.c_uint8),
        ("pad_", C.c_uint8 * 6),
    ]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_U8 = C.POINTER(C.c_uint8)
_F = C.POINTER(C.c_float)

# name -> (restype, argtypes); every symbol declared in include/avbd.h
SIGNATURES = {
    "avbd_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "avbd_destroy": (None, [_P]),
    "avbd_last_error": (C.c_char_p, [_P]),
    "avbd_add_particles": (C.c_int, [_P, C.c_uint32, _D, _D, _D, _D, _D, _U8, _U32]),
    "avbd_add_energies": (C.c_int, [_P, C.c_int, C.c_void_p, C.c_uint32]),
    "avbd_add_oriented_particles": (C.c_int, [_P, C.c_uint32, _D, _D, _D, _D, _D, _D, _U8, _U32]),
    "avbd_add_rod": (C.c_int, [_P, C.c_uint32, C.c_double, _D, C.c_double, C.c_double, C.c_double, C.c_double, _D,
                               _D, _D, C.c_int, C.c_int, C.c_int, _U32]),
    "avbd_add_sphere": (C.c_int, [_P, C.c_double, C.c_double, _D, _D, C.c_int, C.c_int, _U32]),
    "avbd_set_rotation_state": (C.c_int, [_P, C.c_uint32, C.c_uint32, _D, _D]),
    "avbd_add_mass": (C.c_int, [_P, C.c_uint32, C.c_double, _D]),
    "avbd_set_residual_target": (C.c_int, [_P, C.c_double, C.c_uint32, C.c_uint32]),
    "avbd_get_residual_history": (C.c_int, [_P, _D, C.c_uint32, _U32]),
    "avbd_get_scene_residual_history": (C.c_int, [_P, _D, C.c_uint32, _U32]),
    "avbd_get_scene_iterations": (C.c_int, [_P, _U32, C.c_uint32, _U32]),
    "avbd_get_warm_start_decision": (C.c_int, [_P, _D, _F, C.c_uint32, _U32]),
    "avbd_mlp_param_count": (C.c_uint64, []),
    "avbd_mlp_init": (C.c_int, [_P, C.c_uint64]),
    "avbd_mlp_set_params": (C.c_int, [_P, _F, C.c_uint64]),
    "avbd_mlp_get_params": (C.c_int, [_P, _F, C.c_uint64]),
    "avbd_mlp_set_normalization": (C.c_int, [_P, _F]),
    "avbd_mlp_forward_backward": (C.c_int, [_P, _F, _F, C.c_uint32, _F, C.c_uint64, _F]),
    "avbd_mlp_predict": (C.c_int, [_P, _F, C.c_uint32, _F]),
    "avbd_mlp_train": (C.c_int, [_P, _F, _F, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, _F]),
    "avbd_record_steps": (C.c_int, [_P, C.c_uint32, _F, _F, C.c_uint32, _U32]),
    "avbd_set_warm_start": (C.c_int, [_P, C.c_int]),
    "avbd_mlp_gemm_test": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, _F, _F, _F]),
    "avbd_download_rotations": (C.c_int, [_P, _D, _D]),
    "avbd_add_tet_mesh": (C.c_int, [_P, C.c_uint32, _D, C.c_uint32, _U32, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_double, C.c_double, _U32]),
    "avbd_set_particle_state": (C.c_int, [_P, C.c_uint32, C.c_uint32, _D, _D, _D, _D]),
    "avbd_set_fixed": (C.c_int, [_P, C.c_uint32, C.c_uint32, _U8]),
    "avbd_set_external_accel": (C.c_int, [_P, C.c_uint32, C.c_uint32, _D]),
    "avbd_set_scenes": (C.c_int, [_P, C.c_uint32, _U32]),
    "avbd_build_graph": (C.c_int, [_P]),
    "avbd_step": (C.c_int, [_P, C.c_uint32]),
    "avbd_synchronize": (C.c_int, [_P]),
    "avbd_download_state": (C.c_int, [_P, _D, _D]),
    "avbd_download_inertial": (C.c_int, [_P, _D]),
    "avbd_download_energies": (C.c_int, [_P, C.c_int, C.c_void_p, C.c_uint32]),
    "avbd_download_coloring": (C.c_int, [_P, _U32, _U32]),
    "avbd_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "avbd_num_particles": (C.c_uint32, [_P]),
    "avbd_num_energies": (C.c_uint32, [_P, C.c_int]),
    "avbd_get_launch_config": (C.c_int, [_P, _U32, _U32, _U32]),
    "avbd_get_step_mode": (C.c_int, [_P, _U32, _U32]),
    "avbd_set_timing": (C.c_int, [_P, C.c_int]),
    "avbd_get_solve_timing": (C.c_int, [_P, _D, _U32]),
    "avbd_get_broadphase_timing": (C.c_int, [_P, _D]),
    "avbd_calibrate_copy": (C.c_int, [_P, C.c_uint64, C.c_uint32, _D]),
    "avbd_time_solve_launches": (C.c_int, [_P, C.c_uint32, _D, _U32]),
    "avbd_add_collision_triangles": (C.c_int, [_P, C.c_uint32, _U32]),
    "avbd_detect_collisions": (C.c_int, [_P]),
    "avbd_download_bvh": (C.c_int, [_P, C.POINTER(C.c_uint64), _U32, _U32, _U32, _U32, _D, _U32]),
    "avbd_download_candidates": (C.c_int, [_P, _U32, C.c_uint32, _U32]),
    "avbd_add_collision_segments": (C.c_int, [_P, C.c_uint32, _U32, C.c_double]),
    "avbd_add_collision_sphere": (C.c_int, [_P, C.c_uint32, C.c_double]),
    "avbd_set_dynamic_contacts": (C.c_int, [_P, C.c_double, C.c_uint32]),
    "avbd_set_self_collision": (C.c_int, [_P, C.c_int]),
    "avbd_download_contacts": (C.c_int, [_P, C.c_uint32, C.POINTER(Contact), C.c_uint32, _U32]),
    "avbd_download_detected": (C.c_int, [_P, C.POINTER(C.c_uint64), _U32, _U8, C.c_uint32, _U32]),
    "avbd_download_conflicts": (C.c_int, [_P, _U8]),
    "avbd_checkpoint_size": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "avbd_save_checkpoint": (C.c_int, [_P, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "avbd_load_checkpoint": (C.c_int, [_P, C.c_void_p, C.c_uint64]),
    "avbd_detect_collisions_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_double,
                                                 C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"avbd_amd native library not found at {LIB_PATH}; run __graft_entry__.build() "
            "(or `python -m avbd_amd.build`) first -- there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class AvbdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"avbd error {code}: {msg}")
        self.code = code


def check(rc: int, ctx=None) -> None:
    if rc != AVBD_OK:
        msg = lib.avbd_last_error(ctx)
        raise AvbdError(rc, msg.decode() if msg else "")
