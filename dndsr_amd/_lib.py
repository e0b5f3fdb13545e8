"""ctypes binding of libdndsr_amd.so (the C-ABI declared in include/dndsr_amd.h).

Loading fails loudly: there is no CPU fallback for any operator of this package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdndsr_amd.so")

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
vp = C.c_void_p


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("nNode", C.c_int64), ("coords", dp), ("nCell", C.c_int64),
                ("cellType", i32p), ("cell2nodePtr", i64p), ("cell2node", i64p),
                ("cell2facePtr", i64p), ("cell2face", i64p), ("nFace", C.c_int64),
                ("faceType", i32p), ("face2nodePtr", i64p), ("face2node", i64p),
                ("face2cell", i64p), ("faceShift", dp), ("faceZone", i32p), ("nBCZone", C.c_int32)]


class VRSettings(C.Structure):
    _fields_ = [("maxOrder", C.c_int), ("intOrder", C.c_int), ("dirWeightScheme", C.c_int),
                ("geomWeightScheme", C.c_int), ("geomWeightBias", C.c_double),
                ("geomWeightPower", C.c_double), ("scaleMultiplier", C.c_double),
                ("jacobiRelax", C.c_double), ("nIterCellSmoothScale", C.c_int), ("bcWeight", C.c_double),
                ("zoneDirWeight", dp), ("SORInstead", C.c_int), ("subs2ndOrder", C.c_int),
                ("subs2ndOrderGGScheme", C.c_int)]


class EulerSettings(C.Structure):
    _fields_ = [("gamma", C.c_double), ("rsType", C.c_int), ("rsMeanValueEig", C.c_int),
                ("rsFixScale", C.c_double), ("rsIncFScale", C.c_double), ("viscous", C.c_int),
                ("muModel", C.c_int), ("Rgas", C.c_double), ("muGas", C.c_double), ("prGas", C.c_double),
                ("TRef", C.c_double), ("CSutherland", C.c_double), ("useRoeJacobian", C.c_int)]


class ImplicitSettings(C.Structure):
    _fields_ = [("dt", C.c_double), ("CFL", C.c_double), ("nRec", C.c_int), ("sgsIter", C.c_int),
                ("gsUpdate", C.c_int), ("maxIter", C.c_int), ("minIter", C.c_int), ("tol", C.c_double),
                ("fp32Sweeps", C.c_int), ("incrementRelax", C.c_double), ("refRho", C.c_double),
                ("refP", C.c_double), ("gmresCode", C.c_int), ("nGmresSpace", C.c_int), ("nGmresIter", C.c_int)]


# Every symbol of include/dndsr_amd.h with its argument types.
SIGNATURES = {
    "dndsr_create": [C.POINTER(vp), C.c_int],
    "dndsr_destroy": [vp],
    "dndsr_last_error": [],
    "dndsr_sync": [vp],
    "dndsr_set_stream": [vp, vp],
    "dndsr_vr_construct": [vp, C.POINTER(MeshDesc), C.POINTER(VRSettings)],
    "dndsr_mesh_info": [vp, i64p, i64p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                        C.POINTER(C.c_int), C.POINTER(C.c_int), dp],
    "dndsr_get_cell_metrics": [vp, vp, vp, vp, vp, vp, vp],
    "dndsr_get_face_metrics": [vp, vp, vp, vp, vp, vp, vp, vp],
    "dndsr_get_cell_quad": [vp, vp, vp],
    "dndsr_get_vr_coeffs": [vp, vp, vp],
    "dndsr_set_vr_coeffs": [vp, vp, vp],
    "dndsr_get_colors": [vp, vp],
    "dndsr_h2d_cellfield": [vp, vp, vp, C.c_int],
    "dndsr_d2h_cellfield": [vp, vp, vp, C.c_int],
    "dndsr_vr_iter": [vp, C.c_int, vp, vp, vp, C.c_int, C.c_double],
    "dndsr_vr_iter_diff": [vp, C.c_int, vp, vp, vp, vp],
    "dndsr_set_bc": [vp, C.c_int, i32p, dp, C.c_double],
    "dndsr_set_bc_gas": [vp, C.c_int, i32p, dp, C.c_double, C.c_double],
    "dndsr_set_vr_dirichlet": [vp, C.c_int, vp],
    "dndsr_eval_dt": [vp, C.POINTER(EulerSettings), vp, C.c_double, C.c_double, C.c_int, vp, vp, vp, vp, vp],
    "dndsr_eval_rhs": [vp, C.POINTER(EulerSettings), vp, vp, vp, vp],
    "dndsr_eval_rhs_noalloc": [vp, C.POINTER(EulerSettings), vp, vp, vp, vp],
    "dndsr_model_rhs": [vp, C.c_int, C.c_double, C.c_double, C.c_double, vp, vp, vp],
    "dndsr_lusgs_init": [vp, vp, C.c_double, C.c_double, vp, vp],
    "dndsr_update_sgs": [vp, C.POINTER(EulerSettings), C.c_double, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int,
                         C.c_int, vp],
    "dndsr_update_sgs_transpose": [vp, C.POINTER(EulerSettings), C.c_double, vp, vp, vp, vp, vp, vp, C.c_int,
                                   C.c_int, C.c_int, vp],
    "dndsr_update_sgs_f32": [vp, C.POINTER(EulerSettings), C.c_float, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int,
                             C.c_int, vp],
    "dndsr_eval_rhs_roe_lambda": [vp, C.POINTER(EulerSettings), vp, vp, vp, vp],
    "dndsr_lusgs_init_block": [vp, C.POINTER(EulerSettings), vp, C.c_double, C.c_double, vp, vp, vp, vp, vp],
    "dndsr_update_sgs_block": [vp, C.POINTER(EulerSettings), C.c_double, vp, vp, vp, vp, vp, vp, vp, C.c_int,
                               C.c_int, C.c_int, vp],
    "dndsr_update_sgs_block_transpose": [vp, C.POINTER(EulerSettings), C.c_double, vp, vp, vp, vp, vp, vp, vp,
                                         C.c_int, C.c_int, C.c_int, vp],
    "dndsr_sgs_transpose_symmetric": [vp, C.POINTER(EulerSettings), C.c_double, vp, vp, vp, vp, vp, vp, C.c_int,
                                      C.c_int, vp, vp],
    "dndsr_update_sgs_block_f32": [vp, C.POINTER(EulerSettings), C.c_float, vp, vp, vp, vp, vp, vp, vp, C.c_int,
                                   C.c_int, C.c_int, vp],
    "dndsr_ssprk3": [vp, C.POINTER(EulerSettings), vp, vp, C.c_int, C.c_double, C.c_int, C.c_double],
    "dndsr_implicit_euler_step": [vp, C.POINTER(EulerSettings), C.POINTER(ImplicitSettings), vp, vp,
                                  C.POINTER(C.c_int), dp, dp, dp],
    "dndsr_implicit_bdf_step": [vp, C.POINTER(EulerSettings), C.POINTER(ImplicitSettings), C.c_int,
                                C.POINTER(vp), vp, vp, C.POINTER(C.c_int), dp, dp, dp],
    "dndsr_implicit_pp_stats": [vp, i64p, dp],
    "dndsr_lusgs_matvec": [vp, C.POINTER(EulerSettings), C.c_double, vp, vp, vp, vp, vp],
    "dndsr_gmres_lusgs": [vp, C.POINTER(EulerSettings), C.c_double, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                          C.c_int, C.c_int, dp],
    "dndsr_eval_rhs_vjp": [vp, C.POINTER(EulerSettings), vp, vp, vp, vp, vp, vp],
    "dndsr_vr_iter_vjp": [vp, C.c_int, C.c_int, C.c_double, vp, vp, vp, vp, vp],
    "dndsr_zone_force": [vp, C.POINTER(EulerSettings), vp, vp, vp, C.c_int, dp],
    "dndsr_zone_force_vjp": [vp, C.POINTER(EulerSettings), vp, vp, vp, C.c_int, dp, vp, vp],
    "dndsr_stream_probe": [vp, C.c_int, vp, vp, C.c_int64, C.c_int, dp],
    "dndsr_set_timing": [vp, C.c_int],
    "dndsr_get_timing": [vp, dp, i64p, C.c_int],
}

KERNEL_IDS = ["vr_iter", "rhs_face", "rhs_gather", "dt_face", "dt_cell", "rk_axpby", "sgs_colour", "lusgs_init",
              "dual_residual", "increment", "to_f32", "adjoint", "linsolve", "rhs_viscous"]

_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dndsr_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.dndsr_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        raise RuntimeError("dndsr_amd: " + lib().dndsr_last_error().decode())
