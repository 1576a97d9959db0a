"""ctypes binding of the C-ABI in include/mom_gpu.h (lib/libmomgpu.so).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing, and every call raises MomError on a
non-zero status with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libmomgpu.so")

MOM_PREC_FP64 = 64
MOM_PREC_MIXED = 32

# every symbol declared in include/mom_gpu.h
EXPORTS = [
    "mom_last_error", "mom_ctx_create", "mom_ctx_destroy", "mom_mesh_create", "mom_mesh_destroy",
    "mom_mesh_near_pairs", "mom_compute_operator", "mom_sysmat_pmchwt", "mom_srcvec_pw",
    "mom_solve_multi_linsys", "mom_system_create", "mom_system_destroy", "mom_system_fill",
    "mom_system_get_matrix", "mom_system_set_matrix", "mom_system_factor", "mom_system_solve",
    "mom_system_solve_adjoint", "mom_cs_abs_gradient", "mom_pmchwt_deps_apply",
    "mom_system_solve_pw", "mom_system_profile", "mom_system_stats", "mom_cs_prtsrf", "mom_far_fields",
    "mom_rcs_ext", "mom_nlsurf_coef", "mom_sysmat_pmchwt_nls", "mom_system_fill_nls", "mom_problem_build", "mom_problem_load_gmsh", "mom_problem_destroy", "mom_problem_sizes", "mom_problem_domain",
    "mom_problem_topology", "mom_problem_mesh_create", "mom_problem_set_lines", "mom_problem_classify_edges",
    "mom_problem_edge_bnd", "mom_symm_group", "mom_symm_edge_bc", "mom_system_create_order",
    "mom_system_fill_symm", "mom_srcvec_pw_symm", "mom_cs_prtsrf_symm", "mom_rcs_ext_symm",
    "mom_pmchwt_vertex_adjoint", "mom_cs_abs_vertex_explicit", "mom_ctx_fill_tuning", "mom_far_fields_noalloc",
    "mom_ctx_fill_profile", "mom_mesh_classify_like", "mom_far_fields_adjoint", "mom_nlsurf_coef_adjoint",
    "mom_domain_block_apply", "mom_system_set_pivoting", "mom_system_set_precision",
    "mom_system_set_cooperative", "mom_symm_group_c",
]
# test hooks declared in include/mom_gpu_testing.h
TEST_EXPORTS = ["mom_test_mfma_layout", "mom_test_gemm_sub", "mom_test_panel_mode", "mom_test_lookahead",
                "mom_test_unit_lower_inverse", "mom_test_u_solve", "mom_test_vertex_fd", "mom_test_batch_fallback",
                "mom_test_tournament", "mom_test_factor_graph", "mom_test_fill_lossless",
                "mom_test_fill_spw"]


class MomError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("status %d: %s" % (status, msg))
        self.status = status


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_d = C.c_double


def lib():
    """Load libmomgpu.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmomgpu.so is not built: %s (run __graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.mom_last_error.restype = C.c_char_p
    L.mom_ctx_create.argtypes = [_i32, C.POINTER(_vp)]
    L.mom_ctx_destroy.argtypes = [_vp]
    L.mom_mesh_create.argtypes = [_vp, _i64, _dp, _i64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _i64, _ip, _ip,
                                  _dp, _ip, _d, C.POINTER(_vp)]
    L.mom_mesh_destroy.argtypes = [_vp]
    L.mom_mesh_near_pairs.restype = _i64
    L.mom_mesh_near_pairs.argtypes = [_vp]
    L.mom_compute_operator.argtypes = [_vp, _vp, C.c_char, _d, _d, _d, _i32, _dp, _dp, _dp, _i64]
    L.mom_sysmat_pmchwt.argtypes = [_vp, _i32, C.POINTER(_vp), _dp, _i64, _d, _i32, _dp, _dp, _dp, _i64]
    L.mom_srcvec_pw.argtypes = [_vp, _vp, _i64, _d, _d, _d, _i64, _dp, _i32, _dp, _dp, _dp, _i64]
    L.mom_solve_multi_linsys.argtypes = [_vp, _i64, _dp, _i64, _i64, _dp, _i64, C.c_void_p]
    L.mom_system_create.argtypes = [_vp, _i64, _i32, C.POINTER(_vp)]
    L.mom_system_destroy.argtypes = [_vp]
    L.mom_system_fill.argtypes = [_vp, _i32, C.POINTER(_vp), _dp, _d, _i32, _dp, _dp]
    L.mom_system_get_matrix.argtypes = [_vp, _dp, _i64]
    L.mom_system_set_matrix.argtypes = [_vp, _dp, _i64]
    L.mom_system_factor.argtypes = [_vp]
    L.mom_system_solve.argtypes = [_vp, _i64, _dp, _i64, _d, _i32, C.POINTER(_i32), C.POINTER(_d)]
    L.mom_cs_abs_gradient.argtypes = [_vp, _vp, _i64, _d, _d, _d, _dp, _i32, _dp, _dp, _dp, C.POINTER(_d)]
    L.mom_pmchwt_deps_apply.argtypes = [_vp, _vp, _i64, _d, _d, _d, _i32, _dp, _dp, _dp, _dp]
    L.mom_system_solve_adjoint.argtypes = [_vp, _i64, _dp, _i64, _d, _i32, C.POINTER(_i32), C.POINTER(_d)]
    L.mom_system_solve_pw.argtypes = [_vp, _vp, _d, _d, _d, _i64, _dp, _i32, _dp, _dp, _d, _i32, _dp, _i64,
                                      C.POINTER(_i32), C.POINTER(_d)]
    L.mom_system_stats.argtypes = [_vp, _dp]
    L.mom_system_profile.argtypes = [_vp, _i32]
    L.mom_test_mfma_layout.argtypes = [_vp, _i32, _dp]
    L.mom_test_gemm_sub.argtypes = [_vp, _i32, _i64, _i64, _i64, _dp, _i64, _dp, _i64, _dp, _i64]
    L.mom_test_panel_mode.argtypes = [_i32]
    L.mom_test_lookahead.argtypes = [_i32]
    L.mom_test_u_solve.argtypes = [_i32]
    L.mom_test_vertex_fd.argtypes = [_i32]
    L.mom_test_batch_fallback.argtypes = [_i32]
    L.mom_test_tournament.argtypes = [_i32]
    L.mom_test_factor_graph.argtypes = [_i32, C.POINTER(_i64)]
    L.mom_test_fill_lossless.argtypes = [_i32]
    L.mom_test_fill_spw.argtypes = [C.c_void_p, _i32]
    L.mom_ctx_fill_tuning.argtypes = [_vp, C.POINTER(_i32), _dp]
    L.mom_ctx_fill_profile.argtypes = [_vp, _i32, _dp]
    L.mom_mesh_classify_like.argtypes = [_vp, _vp]
    L.mom_far_fields_adjoint.argtypes = [_vp, _vp, _i64, _d, _d, _d, _d, _i64, _dp, _dp, _dp, _i32, _dp, _dp, _dp]
    L.mom_nlsurf_coef_adjoint.argtypes = [_vp, _vp, _i64, _d, _d, _d, _dp, _dp, _dp, _i32, _dp, _dp, _dp, _dp,
                                          _dp]
    L.mom_system_set_pivoting.argtypes = [_vp, _i32]
    L.mom_system_set_precision.argtypes = [_vp, _i32]
    L.mom_system_set_cooperative.argtypes = [_vp, _i32]
    L.mom_symm_group_c.argtypes = [_i32, _ip, C.POINTER(_i32), _dp, _dp, _dp, _ip]
    L.mom_domain_block_apply.argtypes = [_vp, _vp, _i64, _d, _d, _d, _i32, _i32, _dp, _dp, _dp, _dp]
    L.mom_test_unit_lower_inverse.argtypes = [_vp, _vp, _i64, _vp]
    L.mom_cs_prtsrf.argtypes = [_vp, _vp, _i64, _d, _d, _d, _dp, _d, _d, _d, _i32, _dp, _dp, C.POINTER(_d),
                                C.POINTER(_d)]
    L.mom_far_fields.argtypes = [_vp, _vp, _i64, _d, _d, _d, _dp, _d, _i64, _dp, _dp, _i32, _dp, _dp, _dp]
    L.mom_far_fields_noalloc.argtypes = [_vp, _vp, _i64, _d, _d, _d, _vp, _d, _i64, _vp, _vp, _i32, _dp, _dp, _vp]
    L.mom_rcs_ext.argtypes = [_vp, _vp, _i64, _d, _d, _d, _dp, _i32, _i32, _i32, C.c_void_p, _i32, _dp, _dp, _dp]
    L.mom_nlsurf_coef.argtypes = [_vp, _vp, _i64, _d, _d, _d, _dp, _dp, _dp, _i32, _dp, _dp, _dp, _dp]
    L.mom_sysmat_pmchwt_nls.argtypes = [_vp, _i32, C.POINTER(_vp), _dp, _i64, _d, _i32, _dp, _dp, _i64, _dp, _i64,
                                        _dp, _i64, _dp, _i64]
    L.mom_system_fill_nls.argtypes = [_vp, _i32, C.POINTER(_vp), _dp, _d, _i32, _dp, _dp, _i64, _dp, _i64, _dp,
                                      _i64]
    L.mom_problem_build.argtypes = [_i64, _dp, _i64, _ip, _ip, _d, _i32, _ip, _ip, C.POINTER(_vp)]
    L.mom_problem_load_gmsh.argtypes = [C.c_char_p, _d, _i32, _ip, _ip, C.POINTER(_vp)]
    L.mom_problem_destroy.argtypes = [_vp]
    L.mom_problem_sizes.argtypes = [_vp, _i32] + [C.POINTER(_i64)] * 4
    L.mom_problem_domain.argtypes = [_vp, _i32, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _ip,
                                     C.POINTER(_d)]
    L.mom_problem_topology.argtypes = [_vp, _i32, _ip, _ip, _ip, _ip]
    L.mom_problem_mesh_create.argtypes = [_vp, _vp, _i32, C.POINTER(_vp)]
    L.mom_problem_set_lines.argtypes = [_vp, _i64, _ip, _ip]
    L.mom_problem_classify_edges.argtypes = [_vp, _i32, _i32, C.POINTER(_i64)]
    L.mom_problem_edge_bnd.argtypes = [_vp, _i32, _ip]
    L.mom_symm_group.argtypes = [_i32, _ip, C.POINTER(_i32), _dp, _dp, _dp, _ip]
    L.mom_symm_edge_bc.argtypes = [_i64, _ip, _i32, _ip, _i32, _ip, C.POINTER(_i64)]
    L.mom_system_create_order.argtypes = [_vp, _i64, _i32, C.POINTER(_vp)]
    L.mom_system_fill_symm.argtypes = [_i32, C.POINTER(_vp), _i32, C.POINTER(_vp), _dp, _i64, _ip, _i32, _ip, _d,
                                       _i32, _dp, _dp]
    L.mom_srcvec_pw_symm.argtypes = [_vp, _vp, _i64, _d, _d, _d, _i32, _ip, _i64, _dp, _i32, _dp, _dp, _dp, _i64]
    L.mom_cs_prtsrf_symm.argtypes = [_vp, _vp, _i64, _d, _d, _d, _i32, _ip, _dp, _d, _d, _d, _i32, _dp, _dp,
                                     C.POINTER(_d), C.POINTER(_d)]
    L.mom_rcs_ext_symm.argtypes = [_vp, _vp, _i64, _d, _d, _d, _i32, _ip, _dp, _i32, _i32, _i32, C.c_void_p, _i32,
                                   _dp, _dp, _dp]
    L.mom_pmchwt_vertex_adjoint.argtypes = [_vp, _vp, _i64, _d, _d, _d, _i32, _dp, _dp, _dp, _dp, _dp]
    L.mom_cs_abs_vertex_explicit.argtypes = [_vp, _vp, _i64, _d, _d, _d, _dp, _dp, _d, _d, _d, _i32, _dp, _dp, _dp]
    _lib = L
    return L


def check(status):
    if status != 0:
        raise MomError(status, lib().mom_last_error().decode(errors="replace"))
    return status
