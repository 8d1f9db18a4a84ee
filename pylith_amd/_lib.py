"""ctypes binding of ``libpylith_amd.so`` (the C ABI declared in include/pylith_amd.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpylith_amd.so")

HEX_Q1, TET_P2 = 1, 2
FP64, FP32 = 0, 1
GF_PC_JACOBI, GF_PC_AMG = 0, 1
PC_PBJACOBI, PC_AMG = 0, 1
ALL_KEYS = -1  # label value selecting every key of a part (one per material)
ADD_VALUES, INSERT_VALUES = 0, 1
LHS, RHS, LHS_LUMPED_INV, LHS_WEIGHTED = 0, 1, 2, 3
FN_NONE = 0
F1V_ISOLINEAR_ELASTICITY_3D = 1
JF3VU_ISOLINEAR_ELASTICITY_3D = 2
F1V_GENMAXWELL_3D = 3
JF3VU_GENMAXWELL_3D = 4
VISCOUS_STRAIN_GENMAXWELL_3D = 5
TOTAL_STRAIN_3D = 6
F0V_BODYFORCE = 7
F0V_GRAV = 8
F0V_GRAVBODYFORCE = 9
F1V_ISOLINEAR_ELASTICITY_REFSTATE_3D = 10
KSP_CG, KSP_FGMRES_IR, KSP_IR = 0, 1, 2
ALGO_AUTO, ALGO_QUADRATURE_POINTS, ALGO_FACTORED, ALGO_POINTFN = 0, 1, 2, 3
ACCUMULATE_AUTO, ACCUMULATE_COLORED, ACCUMULATE_LDS_ATOMIC = 0, 1, 2

_c_int, _c_dbl, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dblp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "plx_dm_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _i32p, _i32p,
                               _dblp, _i32p]),
    "plx_dm_destroy": (_c_int, [_vp]),
    "plx_dm_create_ordered": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _i32p, _i32p,
                                       _dblp, _i32p, _i32p]),
    "plx_dm_set_point_order": (_c_int, [_vp, _c_int, _i32p]),
    "plx_dm_set_stream": (_c_int, [_vp, _vp]),
    "plx_dm_last_error": (ctypes.c_char_p, [_vp]),
    "plx_dm_get_sizes": (_c_int, [_vp, _i64p, _i64p]),
    "plx_dm_set_kernel_options": (_c_int, [_vp, _c_int, _c_int, _c_int]),
    "plx_dm_set_constraints": (_c_int, [_vp, _c_int, _i32p]),
    "plx_dm_add_residual": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int]),
    "plx_dm_add_jacobian": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "plx_dm_set_auxiliary": (_c_int, [_vp, _c_int, _c_int, _c_int, _i32p, _c_int, _dblp]),
    "plx_dm_set_constants": (_c_int, [_vp, _c_int, _c_int, _c_int, _dblp]),
    "plx_compute_residual": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _vp, _vp, _vp, _c_int]),
    "plx_compute_residual_noalloc": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _vp, _vp, _vp, _c_int]),
    "plx_compute_jacobian_action": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _vp, _vp,
                                             _vp, _vp, _c_int]),
    "plx_compute_jacobian_pbdiag": (_c_int, [_vp, _c_int, _c_int, _c_dbl, _c_dbl, _vp, _vp]),
    "plx_local_to_global": (_c_int, [_vp, _c_int, _vp, _vp, _c_int]),
    "plx_global_to_local": (_c_int, [_vp, _c_int, _vp, _vp]),
    "plx_insert_boundary_values": (_c_int, [_vp, _c_int, _vp, _vp]),
    "plx_vec_dot": (_c_int, [_vp, _c_int, ctypes.c_int64, _vp, _vp, _dblp]),
    "plx_vec_norm2": (_c_int, [_vp, _c_int, ctypes.c_int64, _vp, _dblp]),
    "plx_solver_create": (_c_int, [ctypes.POINTER(_vp), _vp, _c_int, _c_int]),
    "plx_solver_destroy": (_c_int, [_vp]),
    "plx_solver_setup": (_c_int, [_vp]),
    "plx_solver_set_inner": (_c_int, [_vp, _c_dbl, _c_int, _c_int]),
    "plx_solver_solve": (_c_int, [_vp, _c_int, _vp, _vp, _c_dbl, _c_int, _vp]),
    "plx_dm_add_fault": (_c_int, [_vp, _c_int, _i32p, _i32p, _i32p, _i32p, _c_int, _dblp, _dblp]),
    "plx_dm_set_fault_slip": (_c_int, [_vp, _dblp]),
    "plx_dm_set_fault_slip_device": (_c_int, [_vp, _vp]),
    "plx_dm_get_fault_info": (_c_int, [_vp, _i64p, ctypes.POINTER(_c_int)]),
    "plx_compute_residual_cohesive": (_c_int, [_vp, _c_dbl, _vp, _vp, _c_int]),
    "plx_compute_jacobian_action_cohesive": (_c_int, [_vp, _vp, _vp, _c_int]),
    "plx_update_state_vars": (_c_int, [_vp, _c_int, _c_int, _c_dbl, _vp]),
    "plx_dm_get_auxiliary": (_c_int, [_vp, _c_int, _c_int, _dblp]),
    "plx_gf_create": (_c_int, [ctypes.POINTER(_vp), _vp, _c_int, _c_int]),
    "plx_gf_destroy": (_c_int, [_vp]),
    "plx_gf_set_options": (_c_int, [_vp, _c_int, _c_dbl, _c_int, _c_int]),
    "plx_gf_set_preconditioner": (_c_int, [_vp, _c_int]),
    "plx_solver_set_preconditioner": (_c_int, [_vp, _c_int]),
    "plx_gf_autotune": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "plx_gf_probe_times": (_c_int, [_vp, _dblp]),
    "plx_gf_apply": (_c_int, [_vp, _c_int, _c_int, _vp, _vp]),
    "plx_gf_solve": (_c_int, [_vp, _c_int, _vp, _vp, _c_dbl, _c_int, _vp]),
    "plx_jacobian_pattern": (_c_int, [_vp, _i64p, _i32p, _i32p]),
    "plx_compute_jacobian": (_c_int, [_vp, _c_int, _c_int, _c_dbl, _c_dbl, _vp, _vp, _vp, _c_int]),
    "plx_dm_add_residual_pointfn": (_c_int, [_vp, _c_int, _c_int, _vp]),
    "plx_dm_add_jacobian_pointfn": (_c_int, [_vp, _c_int, _c_int, _vp, _vp]),
    "plx_gf_slots_open": (_c_int, [_vp, _c_int, _c_dbl, _c_int]),
    "plx_gf_slots_load": (_c_int, [_vp, _c_int, _i32p, _vp]),
    "plx_gf_slots_step": (_c_int, [_vp, _i32p]),
    "plx_gf_slot_take": (_c_int, [_vp, _c_int, _vp, _vp]),
    "plx_dm_set_event_logging": (_c_int, [_vp, _c_int]),
    "plx_dm_get_event_stats": (_c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_c_int), _dblp]),
    "plx_dm_reset_event_stats": (_c_int, [_vp]),
}



class SolveStats(ctypes.Structure):
    _fields_ = [("outer_iterations", _c_int), ("inner_iterations", _c_int), ("applies_fp64", _c_int),
                ("converged", _c_int), ("initial_residual", _c_dbl), ("final_residual", _c_dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GfStats(ctypes.Structure):
    _fields_ = [("block", _c_int), ("num_blocks", _c_int), ("outer_iterations", _c_int),
                ("inner_iterations", _c_int), ("converged", _c_int), ("max_relative_residual", _c_dbl)]


class SlotStats(ctypes.Structure):
    _fields_ = [("converged", _c_int), ("outer_iterations", _c_int), ("inner_iterations", _c_int),
                ("relative_residual", _c_dbl)]


SLOT_EMPTY, SLOT_RUNNING, SLOT_CONVERGED, SLOT_FAILED = 0, 1, 2, 3

_lib = None


def load():
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(the MI355X path has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib

# Test hook (include/pylith_amd_debug.h): host emulation of the tile kernels.
SIGNATURES["plx_debug_emulate_apply"] = (_c_int, [_c_int, _c_int, _c_int, _c_int, _i32p, _i32p,
                                                  _dblp, _i32p, _c_int, _dblp, _c_int, _vp, _vp,
                                                  _c_int, _c_int])
SIGNATURES["plx_debug_plan_info"] = (_c_int, [_vp, _c_int, _c_int, ctypes.POINTER(_c_int), _i64p,
                                              ctypes.POINTER(_c_int)])
