"""ctypes binding of libfemx_amd.so (the C-ABI in include/femx_amd.h).

The library is built in-tree by femx_amd.build.  Loading fails loudly when
it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfemx_amd.so")

_i32 = C.c_int32
_f64 = C.c_double
_vp = C.c_void_p
_pi = C.POINTER(C.c_int32)

_SIGS = {
    "fx_version": ([], C.c_int),
    "fx_last_error": ([], C.c_char_p),
    "fx_mesh_p1": ([_i32, _vp, _vp, _vp], C.c_int),
    "fx_mesh_p2": ([_i32, _vp, _vp, _vp, _vp], C.c_int),
    "fx_th_dofs": ([_i32, _vp, _vp, _vp], C.c_int),
    "fx_th_mesh": ([_i32, _i32, _vp, C.POINTER(_i32), _vp, _vp, _vp], C.c_int),
    "fx_map_create": ([_i32, _i32, _i32, _vp, _vp, _vp, _vp, C.POINTER(_vp)], C.c_int),
    "fx_map_info_get": ([_vp, _vp], C.c_int),
    "fx_map_export_pattern": ([_vp, _vp, _vp], C.c_int),
    "fx_map_export_jac": ([_vp, _vp, _vp], C.c_int),
    "fx_map_export_rowinc3": ([_vp, _vp, _vp], C.c_int),
    "fx_map_export_rowinc": ([_vp, _vp, _vp, _vp], C.c_int),
    "fx_map_export_rowinc15": ([_vp, _vp, _vp, _vp], C.c_int),
    "fx_map_th_blocks": ([_vp, _i32, _i32, _vp, _vp, _vp, _vp], C.c_int),
    "fx_map_th_lanes": ([_vp, _i32, _vp, _vp, _vp], C.c_int),
    "fx_map_th_rounds": ([_i32, _i32] + [_vp] * 9, C.c_int),
    "fx_map_gls_plan": ([_vp, _i32] + [_vp] * 9, C.c_int),
    "fx_map_destroy": ([_vp], None),
    "fx_csr_transpose_map": ([_i32, _i32, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "fx_bcmap_create": ([_i32, _vp, _vp, _i32, _vp, C.POINTER(_vp)], C.c_int),
    "fx_bcmap_info_get": ([_vp, _vp], C.c_int),
    "fx_bcmap_export": ([_vp] * 7, C.c_int),
    "fx_bcmap_export_elim": ([_vp] * 5, C.c_int),
    "fx_bcmap_destroy": ([_vp], None),
    "fx_bc_replace_rows": ([_i32, _vp, _vp, _vp, _vp, _f64, _vp], C.c_int),
    "fx_bc_replace_res": ([_i32, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "fx_bc_zero": ([_i32, _vp, _vp, _vp], C.c_int),
    "fx_bc_prepare_forward": ([_i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
                              C.c_int),
    "fx_csr_matvec": ([_i32, _i32, _vp, _vp, _vp, _vp, _vp, _f64, _f64, _vp], C.c_int),
    "fx_csr_matvec16": ([_i32, _i32, _vp, _vp, _vp, _vp, _vp, _f64, _f64, _vp], C.c_int),
    "fx_csr_col_off16": ([_i32, _vp, _vp, _vp], C.c_int),
    "fx_p1_stencil_plan": ([_i32, _vp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp], C.c_int),
    "fx_p1_assemble_stencil": ([_i32] * 4 + [_vp] * 3 + [_vp, C.c_int] + [_vp] * 9 + [_vp],
                               C.c_int),
    "fx_csr_stencil_plan": ([_i32, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp], C.c_int),
    "fx_csr_matvec_multi_st": ([_i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _i32,
                                _f64, _f64, _vp], C.c_int),
    "fx_csr_matvec_t": ([_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _f64, _f64, _vp], C.c_int),
    "fx_vec_axpby": ([_i32, _f64, _vp, _f64, _vp, _vp], C.c_int),
    "fx_dot_work_len": ([], _i32),
    "fx_vec_dot": ([_i32, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "fx_p1_assemble": ([_i32, _i32, _i32, _i32] + [_vp] * 4 + [C.c_int] + [_vp] * 7 + [_vp],
                       C.c_int),
    "fx_p1_vjp": ([_i32, _i32] + [_vp] * 11 + [_vp], C.c_int),
    "fx_p1_vjp_patch": ([_i32, _i32] + [_vp] * 9 + [_vp], C.c_int),
    "fx_p1_asm_destroy": ([_vp], None),
    "fx_p1_asm_create": ([_i32, _i32] + [_vp] * 5 + [_i32, _i32, C.POINTER(_vp)], C.c_int),
    "fx_p1_asm_info_get": ([_vp, _vp], C.c_int),
    "fx_p1_asm_export": ([_vp] * 12, C.c_int),
    "fx_p1_assemble_patch": ([_i32, _i32, _i32] + [_vp] * 4 + [C.c_int] + [_vp] * 8 + [_vp],
                             C.c_int),
    "fx_th_assemble": ([_i32] * 4 + [_vp] * 5 + [_f64] + [_vp] * 10 + [_vp], C.c_int),
    "fx_th_assemble_rounds": ([_i32] * 4 + [_vp] * 5 + [_f64] + [_vp] * 10 + [_vp], C.c_int),
    "fx_csr_matvec_multi": ([_i32, _i32, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _i32, _f64,
                             _f64, _vp], C.c_int),
    "fx_vec_div": ([_i32, _vp, _vp, _vp, _vp], C.c_int),
    "fx_vec_fill": ([_i32, _f64, _vp, _vp], C.c_int),
    "fx_vec_gather": ([_i32, _vp, _vp, _i32, _vp, _vp], C.c_int),
    "fx_vec_scatter": ([_i32, _vp, _vp, _vp, _i32, _vp], C.c_int),
    "fx_gls_hist_vjp": ([_i32] * 3 + [_vp] * 5 + [_f64] * 3 + [_i32] * 3 + [_vp] * 7 + [_vp],
                        C.c_int),
    "fx_gls_hist_vjp_patch": ([_i32] * 3 + [_vp] * 3 + [_f64] * 3 + [_i32] * 2 + [_vp] * 6 + [_vp],
                              C.c_int),
    "fx_gls_assemble": ([_i32] * 5 + [_vp] * 5 + [_f64] * 3 + [_i32] * 2 + [_vp] * 9 + [_vp],
                        C.c_int),
    "fx_th_vjp": ([_i32] * 2 + [_vp] * 5 + [_f64] + [_vp] * 8 + [_vp], C.c_int),
    "fx_th_vjp_noalloc": ([_i32] * 2 + [_vp] * 5 + [_f64] + [_vp] * 7 + [_vp], C.c_int),
    "fx_patches_create": ([_i32, _i32, _i32, _vp, _vp, _i32, _vp, _vp, _i32, C.POINTER(_vp)],
                          C.c_int),
    "fx_patches_info_get": ([_vp, _vp], C.c_int),
    "fx_patches_export": ([_vp] * 11, C.c_int),
    "fx_patches_destroy": ([_vp], None),
    "fx_th_vjp_patch": ([_i32] * 2 + [_vp] * 3 + [_f64] + [_vp] * 4 + [_vp] * 2 + [_vp], C.c_int),
    "fx_batch_create": ([_i32, _vp, _i32, _vp, _i32, _vp, _i32, C.POINTER(_vp)], C.c_int),
    "fx_batch_work": ([_vp, _vp, _vp], C.c_int),
    "fx_batch_run": ([_vp, C.c_int, _vp], C.c_int),
    "fx_batch_capture": ([_vp, C.c_int], C.c_int),
    "fx_batch_launch": ([_vp, _vp], C.c_int),
    "fx_batch_destroy": ([_vp], None),
    "fx_batch_graph_create": ([_i32, _vp, C.c_int, C.POINTER(_vp)], C.c_int),
    "fx_batch_graph_launch": ([_vp, _vp], C.c_int),
    "fx_batch_graph_destroy": ([_vp], None),
}






class PatchInfo(C.Structure):
    _fields_ = [(n, _i32) for n in ("num_patches", "num_ldofs", "num_seam", "num_parts",
                                    "max_patch_ldofs", "ldof_stride")]


class PatchPlan(C.Structure):
    _fields_ = ([(n, _i32) for n in ("num_patches", "num_seam", "num_parts", "max_patch_ldofs")]
                + [(n, _vp) for n in ("patch_off", "patch_elem", "elem_vert", "elem_ldof", "ldof_off",
                                      "ldof_dof", "ldof_part", "ldof_end", "seam_off", "seam_dof",
                                      "parts")])


class P1Instance(C.Structure):
    _fields_ = ([(n, _i32) for n in ("num_nodes", "num_elems", "nnz", "max_row_len")]
                + [(n, _vp) for n in ("x", "y", "conn", "m", "f", "u", "inc_off", "inc", "row_ptr",
                                      "vals", "res", "lam", "bc_mask", "jt", "g")]
                + [("vjp", PatchPlan)])


class THInstance(C.Structure):
    _fields_ = ([(n, _i32) for n in ("ndof", "num_elems", "num_blocks", "max_block_nnz")]
                + [(n, _vp) for n in ("xv", "yv", "vconn3", "dofs15", "state")]
                + [("nu", _f64)]
                + [(n, _vp) for n in ("nu_e", "inc_off", "inc_elem", "inc_meta", "blk_row",
                                      "blk_flags", "row_ptr", "vals", "res", "lam", "bc_mask",
                                      "jt", "g", "blk_lanes")]
                + [("vjp", PatchPlan)]
                + [(n, _vp) for n in ("rt_off", "rt_elem", "rt_meta")])


class GLSInstance(C.Structure):
    _fields_ = ([(n, _i32) for n in ("num_nodes", "num_states", "num_elems", "max_tile")]
                + [(n, _vp) for n in ("x", "y", "conn3", "dofs9", "node_rows")]
                + [(n, _f64) for n in ("rho", "mu", "dt")]
                + [(n, _i32) for n in ("step", "num_hist")]
                + [(n, _vp) for n in ("hist", "nxt", "inc_off", "inc_elem", "inc_meta", "blk_runs",
                                      "node_tile", "vals", "res", "adj", "out_lag0", "out_lag1")]
                + [("vjp", PatchPlan)])


class P1AsmInfo(C.Structure):
    _fields_ = [(n, _i32) for n in ("num_patches", "num_entries", "num_ldofs", "num_runs",
                                    "num_left", "max_patch_ldofs", "max_tile", "max_runs")]


class P1AsmPlan(C.Structure):
    _fields_ = ([(n, _i32) for n in ("num_patches", "num_left", "max_patch_ldofs", "max_tile",
                                     "max_runs")]
                + [(n, _vp) for n in ("patch_off", "patch_elem", "elem_ldof", "elem_pos",
                                      "ldof_off", "ldof_dof", "ldof_tile", "patch_tile",
                                      "run_off", "runs", "rows_left")])


class MapInfo(C.Structure):
    _fields_ = [(n, _i32) for n in ("num_elems", "num_res", "num_states", "nnz", "jac_len",
                                    "max_res", "max_state", "max_jac", "max_row_len", "num_inc")]


class BcMapInfo(C.Structure):
    _fields_ = [(n, _i32) for n in ("rows", "num_bcs", "num_col_entries", "num_elim_rows",
                                    "num_elim_entries")]


_LIB = None


def lib():
    """The loaded library.  Raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"femx_amd: {LIB_PATH} is missing; build it with `python -m femx_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGS)


def check(rc):
    """Raise RuntimeError with the library's message for a non-zero status,
    as femx's require() throws std::runtime_error (femx/common/Checks.hpp:11-17)."""
    if rc != 0:
        msg = lib().fx_last_error().decode()
        raise RuntimeError(msg or f"femx_amd error {rc}")
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))
