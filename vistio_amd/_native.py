"""ctypes binding of the C ABI in include/vistio_hip.h (libvistio_hip.so).

The library is built in-tree (``__graft_entry__.build()`` or
``make -C vistio_amd/csrc``). Loading fails loudly when it is missing: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvistio_hip.so")

VH_OK = 0
STATUS_NAMES = {
    0: "VH_OK", 1: "VH_E_INVALID_CONFIG", 2: "VH_E_NOT_INITIALIZED", 3: "VH_E_NOT_SPD",
    4: "VH_E_UNSUPPORTED", 5: "VH_E_INVALID_MESH", 6: "VH_E_GPU", 7: "VH_E_INTERNAL",
    8: "VH_E_INVALID_ARG", 9: "VH_E_NONFINITE",
}
MODEL_ARAP, MODEL_COROTATIONAL, MODEL_ISOTROPIC_LINEAR, MODEL_ORTHOTROPIC_LINEAR, MODEL_ANISOTROPIC_COROTATIONAL = \
    0, 1, 2, 3, 4
BENDING_NONE, BENDING_DIHEDRAL, BENDING_DISCRETE_SHELLS = 0, 1, 2
PREC_F64, PREC_F32_IR = 0, 1

# Every symbol include/vistio_hip.h declares.
EXPORTS = (
    "vh_create", "vh_destroy", "vh_last_error", "vh_pd_init", "vh_pd_step", "vh_pd_step_resident",
    "vh_state_set_f64", "vh_state_get_f64", "vh_state_set_inv_mass", "vh_pd_set_ground", "vh_pd_residual",
    "vh_pd_solve",
    "vh_pd_info_get", "vh_pd_lumped_masses", "vh_mesh_quad_grid", "vh_mesh_interior_edges", "vh_pd_profile",
    "vh_ipc_init", "vh_pd_step_ipc", "vh_pd_step_ipc_resident", "vh_ipc_detect", "vh_ipc_ccd_step", "vh_bvh_pairs",
    "vh_state_lambda_set_f64", "vh_state_lambda_get_f64", "vh_ipc_crossings", "vh_pd_init_ex",
    "vh_backend_axpy", "vh_backend_dot", "vh_backend_fill", "vh_backend_dot_noalloc", "vh_pd_blocks_get",
    "vh_pd_profile_graph", "vh_narrow_vertex_triangle",
)


class PdParams(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("bending", C.c_int32),
        ("k_membrane", C.c_double), ("k_bending", C.c_double), ("density", C.c_double),
        ("dt_matrix", C.c_double), ("gravity", C.c_double * 3), ("damping", C.c_double),
        ("rayleigh", C.c_double), ("max_iterations", C.c_int32), ("tolerance", C.c_double),
        ("has_ground", C.c_int32), ("ground_height", C.c_double), ("precision", C.c_int32),
        ("ir_sweeps", C.c_int32), ("ir_final_sweeps", C.c_int32), ("leaf_size", C.c_int32),
        ("max_band_levels", C.c_int32), ("band_launch_us", C.c_double), ("band_gbps", C.c_double),
        ("task_entries", C.c_int32),
        ("warp_stiffness", C.c_double), ("weft_stiffness", C.c_double),
        ("bend_warp", C.c_double), ("bend_weft", C.c_double),
    ]


class ContactPairC(C.Structure):
    """vh_contact_pair (ContactPair, contact.rs:22-44, as VertexTriangleTest fills it)."""
    _fields_ = [("indices", C.c_uint32 * 4), ("distance", C.c_float), ("normal", C.c_float * 3),
                ("barycentric", C.c_float * 3), ("contact_type", C.c_int32), ("is_self", C.c_int32)]


class MaterialArrays(C.Structure):
    """vh_material_arrays: per-element k_membrane (t), k_bending (h), density (t); NULL = uniform."""
    _fields_ = [("k_membrane", C.POINTER(C.c_double)), ("k_bending", C.POINTER(C.c_double)),
                ("density", C.POINTER(C.c_double))]


class StepResultC(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("final_residual", C.c_double),
                ("converged", C.c_int32), ("wall_time", C.c_double)]


class StateF32(C.Structure):
    _fields_ = [(nm, C.POINTER(C.c_float)) for nm in (
        "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z",
        "prev_x", "prev_y", "prev_z", "pred_x", "pred_y", "pred_z")] + [("inv_mass", C.POINTER(C.c_float))]


class PdInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("t", C.c_uint32), ("h", C.c_uint32), ("supernodes", C.c_uint32),
                ("levels", C.c_uint32), ("nnz_a", C.c_int64), ("nnz_factor", C.c_int64),
                ("max_front", C.c_int32), ("setup_seconds", C.c_double), ("factor_seconds", C.c_double)]


class GraphProfile(C.Structure):
    """vh_graph_profile: one solve sweep and the local step, timed in graph replay."""
    _fields_ = [("reps", C.c_uint32), ("solve_launches", C.c_uint32), ("solve_us", C.c_double),
                ("solve_bytes", C.c_double), ("local_us", C.c_double), ("local_bytes", C.c_double)]


class IpcParams(C.Structure):
    _fields_ = [
        ("has_ground", C.c_int32), ("ground_height", C.c_double),
        ("has_sphere", C.c_int32), ("sphere_center", C.c_double * 3), ("sphere_radius", C.c_double),
        ("self_collision", C.c_int32), ("exclusion_depth", C.c_int32),
        ("d_hat", C.c_double), ("kappa", C.c_double), ("thickness", C.c_double),
        ("al_max_iterations", C.c_int32), ("al_mu_initial", C.c_double), ("al_mu_growth", C.c_double),
        ("al_tolerance", C.c_double),
        ("compliant_contact", C.c_int32), ("compliant_d_hat_scale", C.c_double),
        ("friction_coefficient", C.c_double), ("contact_damping", C.c_double),
        ("adaptive_contact_damping", C.c_int32), ("contact_damping_max", C.c_double),
        ("contact_velocity_threshold", C.c_double), ("max_pairs", C.c_uint32),
        ("detect_keeps_d_hat", C.c_int32), ("ccd_edge_edge", C.c_int32),
    ]


class IpcStats(C.Structure):
    _fields_ = [("al_iterations", C.c_int32), ("al_converged", C.c_int32), ("mu", C.c_double),
                ("max_violation", C.c_double), ("active", C.c_int64), ("self_contacts", C.c_int64),
                ("pairs", C.c_int64), ("last_alpha", C.c_double)]


VH_MAX_BANDS = 16


class Profile(C.Structure):
    _fields_ = [("bands", C.c_uint32), ("iterations", C.c_uint32), ("sweeps", C.c_uint32),
                ("launches", C.c_uint32), ("local_us", C.c_double),
                ("local_bytes", C.c_double), ("fwd_us", C.c_double * VH_MAX_BANDS),
                ("fwd_bytes", C.c_double * VH_MAX_BANDS), ("bwd_us", C.c_double * VH_MAX_BANDS),
                ("bwd_bytes", C.c_double * VH_MAX_BANDS), ("ir_us", C.c_double), ("ir_bytes", C.c_double),
                ("iteration_us", C.c_double)]


_lib = None


def load():
    """Load libvistio_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "or `make -C vistio_amd/csrc` (the HIP path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp, fp, up = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    sig = {
        "vh_create": ([C.c_int32, C.POINTER(vp)], C.c_int),
        "vh_destroy": ([vp], C.c_int),
        "vh_last_error": ([vp], C.c_char_p),
        "vh_pd_init": ([vp, C.c_uint32, fp, fp, fp, C.c_uint32, up, C.c_uint32, up, C.POINTER(C.c_uint8),
                        C.POINTER(PdParams), fp], C.c_int),
        "vh_pd_init_ex": ([vp, C.c_uint32, fp, fp, fp, C.c_uint32, up, C.c_uint32, up, C.POINTER(C.c_uint8),
                           C.POINTER(PdParams), C.POINTER(MaterialArrays), fp], C.c_int),
        "vh_backend_axpy": ([vp, fp, C.c_uint64, fp, C.c_uint64, C.c_float, fp, C.c_uint64], C.c_int),
        "vh_backend_dot": ([vp, fp, C.c_uint64, fp, C.c_uint64, dp], C.c_int),
        "vh_backend_dot_noalloc": ([vp, fp, C.c_uint64, fp, C.c_uint64, dp], C.c_int),
        "vh_backend_fill": ([vp, fp, C.c_uint64, C.c_float], C.c_int),
        "vh_pd_step": ([vp, C.c_double, C.POINTER(StateF32), C.POINTER(StepResultC)], C.c_int),
        "vh_pd_step_resident": ([vp, C.c_double, C.c_uint32, C.POINTER(StepResultC)], C.c_int),
        "vh_state_set_f64": ([vp, C.c_int32, dp, dp, dp], C.c_int),
        "vh_state_get_f64": ([vp, C.c_int32, dp, dp, dp], C.c_int),
        "vh_state_set_inv_mass": ([vp, fp], C.c_int),
        "vh_pd_set_ground": ([vp, C.c_int32, C.c_double], C.c_int),
        "vh_pd_residual": ([vp, C.c_double, dp], C.c_int),
        "vh_pd_solve": ([vp, dp, dp, C.c_int32], C.c_int),
        "vh_pd_info_get": ([vp, C.POINTER(PdInfo)], C.c_int),
        "vh_pd_blocks_get": ([vp, C.c_int32, dp, C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
        "vh_pd_profile_graph": ([vp, C.c_double, C.c_uint32, C.POINTER(GraphProfile)], C.c_int),
        "vh_pd_lumped_masses": ([vp, fp], C.c_int),
        "vh_pd_profile": ([vp, C.c_double, C.c_uint32, C.POINTER(Profile)], C.c_int),
        "vh_mesh_quad_grid": ([C.c_uint32, C.c_uint32, C.c_float, C.c_float, fp, fp, fp, up], C.c_int),
        "vh_mesh_interior_edges": ([C.c_uint32, C.c_uint32, up, up, C.c_uint32, up], C.c_int),
        "vh_ipc_init": ([vp, C.POINTER(IpcParams)], C.c_int),
        "vh_pd_step_ipc": ([vp, C.c_double, C.POINTER(StateF32), fp, fp, fp, C.POINTER(StepResultC)], C.c_int),
        "vh_pd_step_ipc_resident": ([vp, C.c_double, C.c_uint32, C.POINTER(StepResultC), C.POINTER(IpcStats)],
                                    C.c_int),
        "vh_ipc_detect": ([vp, dp, C.c_double, C.c_double, dp, dp, C.POINTER(C.c_uint8), dp, C.POINTER(IpcStats)],
                          C.c_int),
        "vh_ipc_ccd_step": ([vp, dp, dp, dp], C.c_int),
        "vh_bvh_pairs": ([vp, dp, C.c_double, up, up, C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
        "vh_ipc_crossings": ([vp, dp, C.POINTER(C.c_uint64)], C.c_int),
        "vh_narrow_vertex_triangle": ([vp, C.c_uint32, fp, fp, fp, C.c_uint32, up, C.c_uint64, up, up, C.c_float,
                                       C.POINTER(ContactPairC), C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
        "vh_state_lambda_set_f64": ([vp, dp], C.c_int),
        "vh_state_lambda_get_f64": ([vp, dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
