"""ctypes binding of the C ABI in include/rheocell.h (librheocell.so).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librheocell.so")

RC_OK = 0
RC_ERR = {1: "RC_ERR_ARG", 2: "RC_ERR_HIP", 3: "RC_ERR_OVERLAP", 4: "RC_ERR_GROWTH", 5: "RC_ERR_CG",
          6: "RC_ERR_STATE", 7: "RC_ERR_TILE", 8: "RC_ERR_NAN"}
RC_WARM_REFERENCE = 0
RC_WARM_PREVIOUS = 1
RC_WARM_LEARNED = 2
RC_CG_SITES = 0
RC_CG_ROWS = 1

# timing kinds (rheocell.h RC_T_*)
T_KINDS = ["box_gather", "stage_fused_pred", "stage_fused_corr", "stage_cf", "stage_fs_pred", "stage_fs_corr",
           "cg_diag", "cg_gather_x", "cg_gather_p", "cg_row_resid", "cg_box_pap", "cg_row_update", "cg_box_n4",
           "cg_solve", "step", "cg_gather_p_first", "cg_reduce", "cg_learned", "nem_force", "nem_sums",
           "scg_gather", "scg_init", "scg_row_init", "scg_iter", "scg_final", "scg_reduce", "scg_init_w"]
# kernel symbol of each timing kind (as rocprofv3 names it)
T_SYMBOL = {"box_gather": "k_box_gather", "stage_fused_pred": "k_stage<0>", "stage_fused_corr": "k_stage<0>",
            "stage_cf": "k_stage<1>", "stage_fs_pred": "k_stage<2>", "stage_fs_corr": "k_stage<2>",
            "cg_diag": "k_cg_diag", "cg_gather_x": "k_cg_gather<0>", "cg_gather_p_first": "k_cg_gather<1>",
            "cg_gather_p": "k_cg_gather<2>", "cg_row_resid": "k_cg_row<0>", "cg_box_pap": "k_cg_box<1>",
            "cg_row_update": "k_cg_row<2>", "cg_box_n4": "k_cg_box<0>", "cg_reduce": "k_cg_reduce",
            "nem_force": "k_nem_force", "nem_sums": "k_box_field", "scg_gather": "k_cg_gather<5>",
            "scg_init": "k_scg_init", "scg_row_init": "k_scg_row_init", "scg_iter": "k_scg_iter",
            "scg_final": "k_scg_final", "scg_reduce": "k_scg_", "scg_init_w": "k_scg_init_w"}

EXPORTS = ["rc_create", "rc_upload_cells", "rc_init_forces", "rc_apply_changes_after_equilibration", "rc_step",
           "rc_download_cells", "rc_get_rows", "rc_get_site_array", "rc_total_rows", "rc_cg_stats",
           "rc_set_warm_start", "rc_timing", "rc_enable_timing", "rc_last_error", "rc_destroy",
           "rc_net_create", "rc_net_set_params", "rc_net_get_params", "rc_net_forward", "rc_net_grads",
           "rc_net_train_step", "rc_net_last_error", "rc_net_destroy", "rc_attach_net", "rc_snapshot_box",
           "rc_record_solves", "rc_get_recorded", "rc_set_learned_stages", "rc_set_stage_shape",
           "rc_stage_groups", "rc_stage_persistent", "rc_set_graphs", "rc_graph_stats",
           "rc_growth_stats", "rc_box_field", "rc_box_field_noalloc", "rc_set_cg_form"]


class rc_model_params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("lambda_", C.c_double), ("omega", C.c_double), ("mu", C.c_double),
        ("kappa", C.c_double), ("R", C.c_int), ("anchoring", C.c_int),
        ("zetaQ_self", C.c_double), ("zetaQ_inter", C.c_double), ("J_Q", C.c_double),
        ("friction", C.c_double), ("friction_cell", C.c_double), ("CGtolerance", C.c_double),
        ("wall_slip", C.c_double), ("passive_alpha", C.c_double), ("beta", C.c_double),
        ("J0", C.c_double), ("dt", C.c_double), ("cg_max_iter", C.c_int), ("interaction", C.c_int),
    ]


RC_INTERACTION = {"wetpolarmodel": 0, "wetmodel": 1}


class rc_ext_force(C.Structure):
    _fields_ = [("kind", C.c_int), ("kappa_wall", C.c_double), ("lambda_", C.c_double), ("shear_rate", C.c_double)]


RC_MAX_FORCES = 4


class rc_box_desc(C.Structure):
    _fields_ = [("L", C.c_int), ("ncells", C.c_int), ("params", rc_model_params), ("walls", C.POINTER(C.c_double)),
                ("nforces", C.c_int), ("forces", rc_ext_force * RC_MAX_FORCES)]


class rc_cell(C.Structure):
    _fields_ = [
        ("LsubX", C.c_int), ("LsubY", C.c_int), ("offset", C.c_int * 2), ("corner", C.c_int),
        ("corner_old", C.c_int), ("x_sub_left", C.c_int), ("y_sub_bottom", C.c_int),
        ("CoM", C.c_double * 2), ("CoM_old", C.c_double * 2), ("unrap", C.c_double * 2),
        ("Q00", C.c_double), ("Q01", C.c_double), ("thetaQ", C.c_double), ("thetaQ_old", C.c_double),
        ("area", C.c_double), ("area_old", C.c_double), ("sumF", C.c_double), ("sumF_old", C.c_double),
        ("perimeter", C.c_double), ("S00", C.c_double), ("S01", C.c_double),
        ("total_force_repulsion", C.c_double * 2), ("nemQ", C.c_double * 2), ("nemQ_old", C.c_double * 2),
    ]


class rc_cg_stats_t(C.Structure):
    _fields_ = [("solves", C.c_longlong), ("iters_total", C.c_longlong), ("last_iters", C.c_int * 2),
                ("last_relres", C.c_double * 2), ("last_cold", C.c_int), ("last_solve_ms", C.c_double),
                ("total_solve_ms", C.c_double), ("sweeps_total", C.c_longlong),
                ("learned_used", C.c_longlong), ("learned_tried", C.c_longlong), ("worst_relres", C.c_double)]


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load librheocell.so and declare the C ABI.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("librheocell.so not found at %s: build it with `python -c 'import "
                           "__graft_entry__ as g; g.build()'` (there is no CPU fallback)" % path)
    L = C.CDLL(path)
    vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    L.rc_create.argtypes = [C.POINTER(rc_box_desc), ip, C.POINTER(vp)]
    L.rc_upload_cells.argtypes = [vp, C.POINTER(rc_cell), dp]
    L.rc_init_forces.argtypes = [vp]
    L.rc_apply_changes_after_equilibration.argtypes = [vp]
    L.rc_step.argtypes = [vp, ip, dp]
    L.rc_download_cells.argtypes = [vp, C.POINTER(rc_cell), dp]
    L.rc_get_rows.argtypes = [vp, ip, dp]
    L.rc_get_site_array.argtypes = [vp, ip, dp]
    L.rc_total_rows.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.rc_cg_stats.argtypes = [vp, C.POINTER(rc_cg_stats_t)]
    L.rc_set_warm_start.argtypes = [vp, ip]
    L.rc_set_cg_form.argtypes = [vp, ip]
    L.rc_timing.argtypes = [vp, ip, dp, C.POINTER(C.c_longlong)]
    L.rc_enable_timing.argtypes = [vp, ip]
    L.rc_last_error.argtypes = [vp]
    L.rc_last_error.restype = C.c_char_p
    L.rc_destroy.argtypes = [vp]
    L.rc_box_field.argtypes = [vp, ip, ip, ip, dp, dp, dp, dp]
    # outputs of the allocation-free variant may be device pointers: plain addresses
    L.rc_box_field_noalloc.argtypes = [vp, ip, ip, ip, vp, vp, vp, vp]
    L.rc_destroy.restype = None
    fp = C.POINTER(C.c_float)
    L.rc_net_create.argtypes = [ip, ip, ip, C.POINTER(vp)]
    L.rc_net_set_params.argtypes = [vp, fp, fp]
    L.rc_net_get_params.argtypes = [vp, fp, fp]
    L.rc_net_forward.argtypes = [vp, fp, fp]
    L.rc_net_grads.argtypes = [vp, fp, fp, fp, fp, dp]
    L.rc_net_train_step.argtypes = [vp, fp, fp, C.c_float, dp]
    L.rc_net_last_error.argtypes = [vp]
    L.rc_net_last_error.restype = C.c_char_p
    L.rc_net_destroy.argtypes = [vp]
    L.rc_net_destroy.restype = None
    L.rc_attach_net.argtypes = [vp, vp, fp]
    L.rc_snapshot_box.argtypes = [vp, fp]
    L.rc_record_solves.argtypes = [vp, ip]
    L.rc_get_recorded.argtypes = [vp, fp, C.POINTER(C.c_int)]
    L.rc_set_learned_stages.argtypes = [vp, ip]
    L.rc_set_stage_shape.argtypes = [vp, ip, ip]
    L.rc_set_graphs.argtypes = [vp, ip]
    L.rc_growth_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
    L.rc_graph_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.rc_stage_groups.argtypes = [vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]
    L.rc_stage_persistent.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
    for name in EXPORTS:
        getattr(L, name)
    _lib = L
    return L
