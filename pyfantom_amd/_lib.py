"""ctypes binding of libfantom_hip.so (the gfx950 kernels).

The library is loaded from this package directory.  There is no fallback:
if the library is missing or no HIP device is visible, every operation that
needs it raises.  Status codes map to the reference's exception types
(invalid argument -> ValueError, runtime failure -> RuntimeError).
"""
from __future__ import annotations

import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfantom_hip.so")

KU_NODE = 0
KU_TILE = 1
FIN_STORE, FIN_ALPHA, FIN_BETA, FIN_RZ0 = 0, 1, 2, 3
S_RZ, S_PQ, S_ALPHA, S_RR, S_BETA = 0, 1, 2, 3, 4
N_SCALARS = 16

_c_int, _c_long, _c_double, _vp = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p

# name -> (restype, argtypes); the single source for binding and for the
# export check in the test suite
SIGNATURES = {
    "fantom_abi_version": (_c_int, []),
    "fantom_last_error": (ctypes.c_char_p, []),
    "fantom_stream_sync": (_c_int, [_vp]),
    "fantom_ku_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_ku_f32": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_ku_jacobi_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _c_double, _vp,
                                      _vp, _vp]),
    "fantom_ku_jacobi_prolong_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_double, _vp, _vp, _vp,
                                              _vp]),
    "fantom_ku_jacobi_prolong_f32": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_double, _vp, _vp, _vp,
                                              _vp]),
    "fantom_ku_jacobi0_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _c_double, _vp, _vp]),
    "fantom_ku_residual_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_ku_dot_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_ku_partials": (_c_long, [_c_int, _c_int, _c_int, _c_int]),
    "fantom_hex8_blocks": (_c_int, [_vp, _vp, _vp]),
    "fantom_diag_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp]),
    "fantom_dc_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fantom_fea_locals_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_restrict_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fantom_prolong_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp]),
    "fantom_coarsen_E_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "fantom_coarsen_mask": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "fantom_dense_assemble_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "fantom_csr_row_nnz": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "fantom_mesh_elements": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp]),
    "fantom_mesh_nodes_f64": (_c_int, [_c_int, _c_int, _c_int, _c_double, _c_double, _c_double, _vp, _vp]),
    "fantom_mesh_nodes_f32": (_c_int, [_c_int, _c_int, _c_int, _c_double, _c_double, _c_double, _vp, _vp]),
    "fantom_csr_assemble_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_spd_inverse_f64": (_c_int, [_c_int, _vp, _vp, _vp]),
    "fantom_gemv_f64": (_c_int, [_c_int, _vp, _vp, _vp, _vp]),
    "fantom_vec_partials": (_c_long, [_c_long]),
    "fantom_dot_partials_f64": (_c_int, [_c_long, _vp, _vp, _vp, _vp]),
    "fantom_dot2_partials_f64": (_c_int, [_c_long, _vp, _vp, _vp, _vp, _vp]),
    "fantom_finalize_f64": (_c_int, [_vp, _c_long, _vp, _c_int, _c_int, _vp]),
    "fantom_pcg_xr_f64": (_c_int, [_c_long, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_pcg_p_f64": (_c_int, [_c_long, _vp, _vp, _vp, _c_int, _vp]),
    "fantom_scale_f64": (_c_int, [_c_long, _c_double, _vp, _vp, _vp, _vp]),
    "fantom_axpby_f64": (_c_int, [_c_long, _c_double, _vp, _c_double, _vp, _vp]),
    "fantom_filter_normalizer_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fantom_filter_fwd_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "fantom_filter_adj_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_oc_partials": (_c_long, [_c_long]),
    "fantom_simp_f64": (_c_int, [_c_long, _vp, _c_int, _c_double, _c_double, _c_double, _c_int, _c_double,
                                 _c_double, _vp, _vp, _vp]),
    "fantom_mul_f64": (_c_int, [_c_long, _vp, _vp, _vp, _vp]),
    "fantom_mg_workspace_size": (_c_long, [_c_int, _c_int, _c_int, _c_int, _c_int]),
    "fantom_mg_create": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _c_int, _c_int, _vp, _vp,
                                  _c_long, ctypes.POINTER(_vp)]),
    "fantom_mg_destroy": (_c_int, [_vp]),
    "fantom_mg_set_omega_schedule": (_c_int, [_vp, _c_int, _c_double, _c_double]),
    "fantom_mg_set_smoother": (_c_int, [_vp, _c_int, _c_double, _c_int, _c_int]),
    "fantom_mg_smoother_rho": (_c_int, [_vp, _vp, _c_int]),
    "fantom_chebyshev4_weights": (_c_int, [_c_int, _vp]),
    "fantom_mg_setup": (_c_int, [_vp, _vp, _vp, _vp]),
    "fantom_mg_precondition": (_c_int, [_vp, _vp, _vp, _vp]),
    "fantom_mg_solve": (_c_int, [_vp, _vp, _vp, _c_double, _c_int, _c_int, _c_int, ctypes.POINTER(_c_int),
                                 ctypes.POINTER(_c_double), _vp]),
    "fantom_mx_workspace_size": (_c_long, [_c_int, _c_int, _c_int, _c_int, _c_int]),
    "fantom_mx_create": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _c_int, _vp, _vp, _c_long,
                                  ctypes.POINTER(_vp)]),
    "fantom_mx_destroy": (_c_int, [_vp]),
    "fantom_mx_set_omega_schedule": (_c_int, [_vp, _c_int, _c_double, _c_double]),
    "fantom_mx_set_smoother": (_c_int, [_vp, _c_int, _c_double, _c_int, _c_int]),
    "fantom_mx_set_operator": (_c_int, [_vp, _c_int]),
    "fantom_mx_set_flexible": (_c_int, [_vp, _c_int]),
    "fantom_sweep_bf16_table_size": (_c_long, [_c_int]),
    "fantom_sweep_bf16_tables": (_c_int, [_c_int, _vp, _vp]),
    "fantom_sweep_bf16": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _c_double, _vp, _vp]),
    "fantom_mx_precondition": (_c_int, [_vp, _vp, _vp, _vp]),
    "fantom_mx_smoother_rho": (_c_int, [_vp, _vp, _c_int]),
    "fantom_mx_setup": (_c_int, [_vp, _vp, _vp, _vp]),
    "fantom_mx_solve": (_c_int, [_vp, _vp, _vp, _c_double, _c_double, _c_int, _c_int, _c_int, _c_int,
                                 ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(_c_double), _vp]),
    "fantom_mx_solve_pcg": (_c_int, [_vp, _vp, _vp, _c_double, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "fantom_conv3d_fwd_bf16": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                        _c_int, _vp]),
    "fantom_conv3d_wgrad_parts": (_c_long, [_c_int, _c_int, _c_int, _c_int, _c_int, ctypes.POINTER(_c_long)]),
    "fantom_conv3d_wgrad_bf16": (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "fantom_bias_grad_bf16": (_c_int, [_c_long, _c_int, _vp, _vp, _vp, _vp]),
    "fantom_flip_transpose_bf16": (_c_int, [_c_int, _c_int, _vp, _vp, _vp]),
    "fantom_relu_mask_bf16": (_c_int, [_c_long, _vp, _vp, _vp, _vp]),
    "fantom_add_bf16": (_c_int, [_c_long, _vp, _vp, _vp]),
    "fantom_vol_restrict_bf16": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, ctypes.c_float, _vp]),
    "fantom_vol_prolong_bf16": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, ctypes.c_float, _c_int, _vp]),
    "fantom_adam_f32": (_c_int, [_c_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, _c_int, _vp]),
    "fantom_f32_to_bf16": (_c_int, [_c_long, _vp, _vp, _vp]),
    "fantom_net_input": (_c_int, [_c_long, _vp, _vp, _vp, _c_int, _vp, _vp]),
    "fantom_net_output": (_c_int, [_c_long, _vp, _vp, _c_int, _vp, _c_int, _vp]),
    "fantom_mse_bf16": (_c_int, [_c_long, _vp, _vp, _vp, _c_int, _vp, _vp, _vp]),
    "fantom_node_density": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "fantom_galerkin_tables": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "fantom_gal1_ku_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_gal1_classes": (_c_int, [_vp, _vp, _vp]),
    "fantom_gal1_ku_cls_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fantom_gal_build_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "fantom_elem_ku_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "fantom_elem_from_fine_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fantom_galerkin_tables2": (_c_int, [_vp, _vp]),
    "fantom_gal2_build_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fantom_stencil_build_f64": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    "fantom_stencil_ku_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _c_double, _vp,
                                       _vp]),
    "fantom_stencil_partials": (_c_long, [_c_int, _c_int, _c_int]),
    "fantom_stencil_size": (_c_long, [_c_int, _c_int, _c_int]),
    "fantom_ku_general_partials": (_c_long, [_c_int, _c_int]),
    "fantom_ku_general_workspace": (_c_long, [_c_int, _c_int, _vp]),
    "fantom_ku_general_f64": (_c_int, [_c_int, _c_int, _c_int, _c_int] + [_vp] * 12),
    "fantom_diag_general_f64": (_c_int, [_c_int, _c_int, _c_int] + [_vp] * 8),
    "fantom_dc_general_f64": (_c_int, [_c_int, _c_int] + [_vp] * 4),
    "fantom_oc_update_f64": (_c_int, [_c_long, _vp, _vp, _vp, _c_double, _c_double, _vp, _vp, _vp, _vp, _vp]),
    "fantom_oc_update_noalloc_f64": (_c_int, [_c_long, _vp, _vp, _vp, _c_double, _c_double, _vp, _vp, _vp]),
    "fantom_mma_workspace_size": (_c_long, [_c_long, _c_int]),
    "fantom_mma_update_f64": (_c_int, [_c_long, _c_int, _c_int, _vp, _vp, _vp, _c_double, _c_double, _vp, _vp, _vp,
                                       _vp, _vp, _c_double, _c_double, _vp, _vp, _vp, _c_double, _c_int, _vp, _vp,
                                       _c_long, _vp, _vp]),
}

_lib = None


class ElemGroup(ctypes.Structure):
    """fantom_elem_group: one element size of an unstructured mesh (device arrays)."""
    _fields_ = [("n_el", ctypes.c_int), ("npe", ctypes.c_int), ("el_ids", ctypes.c_void_p),
                ("nodes", ctypes.c_void_p), ("K", ctypes.c_void_p)]


class FantomLibraryMissing(RuntimeError):
    pass


def load():
    """Load and bind the library (no device access)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FantomLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m pyfantom_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("pyfantom_amd needs a HIP device (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def check(status: int):
    if status == 0:
        return
    msg = load().fantom_last_error().decode(errors="replace")
    if status == 1:
        raise ValueError(msg)
    if status == 3:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def ptr(t) -> int:
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def call(name: str, *args):
    """Call an int-status entry point, appending the current stream."""
    check(getattr(load(), name)(*args, stream_ptr()))
