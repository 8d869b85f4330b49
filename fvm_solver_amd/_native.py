"""ctypes binding of libfvm_hip.so (the C ABI declared in include/fvm_hip.h).

The library is the only compute path of this package: if it cannot be loaded, or no
GPU is visible, every entry point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvm_hip.so")
ABI_VERSION = 18

FVM_ERR_ROW_OVERFLOW = 2

c_float_p = ctypes.POINTER(ctypes.c_float)
c_void_p = ctypes.c_void_p
i32, i64, f32, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double


class FvmMeshDev(ctypes.Structure):
    _fields_ = [("dim", i32), ("n_groups", i32), ("n_cells", i64), ("n_facets", i64), ("n_bc", i64),
                ("cell_facet", c_void_p), ("cell_nbr", c_void_p), ("cell_info", c_void_p),
                ("cell_disp", c_void_p), ("cell_gval", c_void_p), ("cell_inv_vol", c_void_p),
                ("cell_eps", c_void_p), ("fac_geo", c_void_p), ("fac_cell", c_void_p),
                ("bc_cell", c_void_p), ("bc_group", c_void_p), ("bc_flags", c_void_p), ("bc_val", c_void_p),
                ("bc_l", c_void_p), ("bc_ga", c_void_p), ("bc_gb", c_void_p), ("bc_char", c_void_p),
                ("n_slots", i32), ("n_meshes", i32), ("blk_slot", c_void_p), ("slot_blk", c_void_p),
                ("slot_g0", c_void_p), ("slot_ng", c_void_p), ("group_s0", c_void_p), ("group_s1", c_void_p),
                ("group_count", c_void_p), ("cell_mesh", c_void_p), ("mesh_phys", c_void_p),
                ("persist_blocks", i32), ("fz_threads", i32),
                ("fz_tile", c_void_p), ("fz_halo", c_void_p), ("fz_hrec", c_void_p), ("fz_cw", c_void_p),
                ("fz_hw", c_void_p), ("fz_fdesc", c_void_p), ("fz_ffac", c_void_p), ("fz_fdesc_x", c_void_p),
                ("fz_geo", c_void_p), ("fz_geo_x", c_void_p), ("fz_ncl", i32), ("fz_ngl", i32), ("fz_nfl", i32),
                ("fz_nh1", i32)]


class FvmPhys(ctypes.Structure):
    _fields_ = [("inv_cv", f32), ("cv", f32), ("R", f32), ("gR", f32), ("mu0", f32), ("T0", f32), ("T0pS", f32),
                ("S", f32), ("mub0", f32), ("T0sq", f32), ("kappa", f32), ("v_factor", f32), ("lim_p", i32)]


class FvmWork(ctypes.Structure):
    _fields_ = [("ubc", c_void_p), ("gp", c_void_p), ("bc_part", c_void_p), ("red", c_void_p),
                ("fflux", c_void_p), ("gp_prev", c_void_p),
                ("bc_part_next", c_void_p)]


class FvmStage(ctypes.Structure):
    _fields_ = [("prim", c_void_p), ("U0", c_void_p), ("acc", c_void_p), ("acc2", c_void_p), ("kbuf", c_void_p),
                ("prim_next", c_void_p), ("U_out", c_void_p), ("dt", c_void_p),
                ("stage", i32), ("last", i32), ("store_k", i32), ("has_prev", i32), ("bc_ready", i32),
                ("b", f32), ("b2", f32), ("n_inc", i32), ("inc_idx", i32 * 10), ("inc_coef", f32 * 10),
                ("mode", i32), ("n_hist", i32), ("f_m1", c_void_p), ("f_m2", c_void_p), ("f_out", c_void_p),
                ("tile_list", c_void_p), ("n_tile_list", i32)]


_PROTOS = {
    "fvm_last_error": (ctypes.c_char_p, []),
    "fvm_abi_version": (ctypes.c_int, []),
    "fvm_ell_spmm_f32": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, i32, i64, i64, i64, i32, c_void_p, i64, i64,
                                        i32, c_void_p, c_void_p, i64, i64, c_void_p]),
    "fvm_ell_spmm_f32_shape": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, i32, i64, i64, i64, i32, c_void_p, i64,
                                              i64, i32, c_void_p, c_void_p, i64, i64, i32, i32, c_void_p]),
    "fvm_ell_fast_path": (ctypes.c_int, [i64, i64, i32, i64, i64, i32, i64, i64, i32, i64, i64, c_void_p, c_void_p]),
    "fvm_csr_max_row_nnz": (ctypes.c_int, [c_void_p, i64, c_void_p, c_void_p]),
    "fvm_ell_pack_f32": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, i64, i32, c_void_p, c_void_p, i64, c_void_p,
                                        c_void_p]),
    "fvm_rhs_f32": (ctypes.c_int, [ctypes.POINTER(FvmMeshDev), ctypes.POINTER(FvmPhys), ctypes.POINTER(FvmWork),
                                   c_void_p, c_void_p, i32, c_void_p]),
    "fvm_rk_stage_f32": (ctypes.c_int, [ctypes.POINTER(FvmMeshDev), ctypes.POINTER(FvmPhys),
                                        ctypes.POINTER(FvmWork), ctypes.POINTER(FvmStage), c_void_p]),
    "fvm_rk_stage_part_f32": (ctypes.c_int, [ctypes.POINTER(FvmMeshDev), ctypes.POINTER(FvmPhys),
                                             ctypes.POINTER(FvmWork), ctypes.POINTER(FvmStage), i32, c_void_p]),
    "fvm_grad_f32": (ctypes.c_int, [ctypes.POINTER(FvmMeshDev), ctypes.POINTER(FvmPhys), ctypes.POINTER(FvmWork),
                                    c_void_p, i32, c_void_p]),
    "fvm_max_lds_per_workgroup": (ctypes.c_int, [c_void_p]),
    "fvm_update_stepsize": (ctypes.c_int, [ctypes.POINTER(FvmWork), i32, i32, f32, f32, f32, f32, f32, f32,
                                           c_void_p, c_void_p]),
    "fvm_state_to_prim_f32": (ctypes.c_int, [c_void_p, c_void_p, i64, i32, f32, c_void_p]),
    "fvm_conservation_f64": (ctypes.c_int, [c_void_p, c_void_p, i64, i32, c_void_p, c_void_p, c_void_p]),
    "fvm_conservation_f64_noalloc": (ctypes.c_int, [c_void_p, c_void_p, i64, i32, c_void_p, c_void_p]),
    "fvm_cfl_dt": (ctypes.c_int, [c_void_p, c_void_p, i64, i32, f64, f64, c_void_p, c_void_p, c_void_p]),
    "fvm_geom_facets_f64": (ctypes.c_int, [c_void_p, i32, c_void_p, i64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fvm_geom_cells_f64": (ctypes.c_int, [c_void_p, i32, c_void_p, i64, c_void_p, c_void_p, c_void_p]),
    "fvm_geom_disp_signs_f64": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, i32, c_void_p, i64, c_void_p, c_void_p,
                                               c_void_p]),
    "fvm_geom_lsq_f64": (ctypes.c_int, [c_void_p, c_void_p, i32, c_void_p, c_void_p, c_void_p, i64, c_void_p,
                                        c_void_p]),
    "fvm_gemm_bf16": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, i64, i64, i64, i64, i64, i32, c_void_p,
                                     c_void_p, i64, c_void_p, i32, c_void_p, c_void_p, c_void_p]),
    "fvm_mp_aggregate": (ctypes.c_int, [c_void_p, i32, c_void_p, i64, i32, c_void_p, c_void_p]),
    "fvm_mp_aggregate_bwd": (ctypes.c_int, [c_void_p, i32, c_void_p, i64, i32, c_void_p, c_void_p, c_void_p]),
    "fvm_mp_layer": (ctypes.c_int, [c_void_p, i32, c_void_p, i64, i32, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    "fvm_transpose_bf16": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, c_void_p]),
    "fvm_relu_residual": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, i64, c_void_p]),
    "fvm_relu_bwd": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, i64, c_void_p]),
    "fvm_colsum_f32": (ctypes.c_int, [c_void_p, i64, i32, c_void_p, c_void_p, c_void_p]),
    "fvm_mse_loss": (ctypes.c_int, [c_void_p, i64, c_void_p, i64, i32, c_void_p, c_void_p, c_void_p]),
    "fvm_adam_f32": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, i64, f32, f32, f32, f32, i32,
                                    c_void_p]),
    "fvm_cast_f32_bf16": (ctypes.c_int, [c_void_p, c_void_p, i64, c_void_p]),
    "fvm_mp_encode": (ctypes.c_int, [c_void_p, i64, i32, c_void_p, c_void_p, i32, c_void_p, c_void_p]),
    "fvm_mp_layer_inf": (ctypes.c_int, [c_void_p, i32, c_void_p, i64, i32, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p, i32, c_void_p, c_void_p]),
    "fvm_mp_window_features": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p, i32, c_void_p, i64, i32, c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_void_p, i32, c_void_p]),
    "fvm_mp_window_apply": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, f32, f32, c_void_p, c_void_p, i64,
                                           c_void_p, c_void_p, i64, i32, c_void_p, c_void_p]),
    "fvm_mp_window_predict": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, f32, f32, f32, f32, c_void_p, c_void_p,
                                             i64, c_void_p, c_void_p, i64, i32, c_void_p, c_void_p, c_void_p]),
    "fvm_mesh_rms2": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, i32, i32, f32, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "fvm_mesh_rms": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, i32, i32, f32, c_void_p, c_void_p, c_void_p]),
    "fvm_mp_defect3": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      i32, i32, c_void_p, c_void_p, c_void_p]),
    "fvm_mp_defect": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, i32, i32, c_void_p,
                                     c_void_p, c_void_p]),
}

_lib = None


def exported_symbols():
    return list(_PROTOS)


def load_library(path: str = LIB_PATH):
    """Load libfvm_hip.so and bind its prototypes (no GPU needed to load)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m fvm_solver_amd.build` "
                           "(this package has no fallback compute path)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.fvm_abi_version() != ABI_VERSION:
        raise RuntimeError("libfvm_hip.so ABI version mismatch: rebuild the library")
    _lib = lib
    return lib


def lib():
    return _lib if _lib is not None else load_library()


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().fvm_last_error().decode()
        if rc == FVM_ERR_ROW_OVERFLOW:
            raise ValueError(msg or "Some rows have more nonzeros than K")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def require_device(t: torch.Tensor, name: str = "tensor"):
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} must be a GPU tensor (device 'cuda'); this library has no CPU path")


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def max_lds_per_workgroup(device=None) -> int:
    """LDS bytes one workgroup may use on `device` (the tile tables of the fused stage are
    sized to it)."""
    out = ctypes.c_int32(0)
    with torch.cuda.device(torch.device(device) if device is not None else torch.cuda.current_device()):
        check(lib().fvm_max_lds_per_workgroup(ctypes.byref(out)), "fvm_max_lds_per_workgroup")
    return int(out.value)
