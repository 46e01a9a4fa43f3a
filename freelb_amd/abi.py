"""ctypes binding of libflb_hip.so, the C ABI declared in include/flb_abi.h.

The library is built in-tree by ``freelb_amd.build`` (``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FLB_LIB_PATH selects another build of the library (default: the in-tree one).
LIB_PATH = os.environ.get("FLB_LIB_PATH") or os.path.join(HERE, "libflb_hip.so")

# ---- enums (include/flb_abi.h) ---------------------------------------------
(FLB_OK, FLB_ERR_ARG, FLB_ERR_STATE, FLB_ERR_HIP, FLB_ERR_UNSUPPORTED, FLB_ERR_LAYOUT, FLB_ERR_NOMEM,
 FLB_ERR_NONFINITE) = range(8)
D2Q9, D3Q19, D2Q5, D3Q7 = 9, 19, 5, 7
STORE_F32, STORE_F16S, STORE_BF16S, STORE_F64, STORE_F16C = range(5)
STORAGE = {"f32": STORE_F32, "f16s": STORE_F16S, "bf16s": STORE_BF16S, "f64": STORE_F64, "f16c": STORE_F16C}
(DYN_NONE, DYN_BGK, DYN_BGK_FIELD, DYN_BOUNCEBACK, DYN_BOUNCEBACK_MW, DYN_MRT, DYN_SMAG_BGK,
 DYN_SMAG_MRT, DYN_BGK_FORCE, DYN_BGK_WRITE, DYN_BGK_SCALAR_FORCE, DYN_SCALAR_BGK,
 DYN_SCALAR_BGK_FIELD) = range(13)
HWBC_NORMAL, HWBC_MOVING, HWBC_ANTI = 0, 1, 2

# Every symbol include/flb_abi.h declares (checked by the CPU test suite).
EXPORTS = (
    "flb_last_error", "flb_abi_version", "flb_block_layout", "flb_block_create", "flb_block_destroy", "flb_get_info",
    "flb_synchronize", "flb_set_tasks", "flb_set_halfway_bc", "flb_set_mrt_rates", "flb_set_omega",
    "flb_set_force", "flb_share_velocity", "flb_add_buoyancy", "flb_clear_buoyancy",
    "flb_upload_flags", "flb_set_field_rho", "flb_set_field_u", "flb_upload_pops_f32",
    "flb_upload_pops_f64", "flb_download_pops_f32", "flb_download_pops_f64", "flb_init_equilibrium",
    "flb_apply_cell_dynamics", "flb_stream", "flb_set_level", "flb_set_watchdog", "flb_apply_halfway_bc", "flb_step", "flb_moments",
    "flb_reduce_mass_momentum", "flb_batch_create", "flb_batch_destroy", "flb_batch_info",
    "flb_batch_set_launch", "flb_batch_step", "flb_batch_capture", "flb_batch_replay",
    "flb_batch_synchronize", "flb_batch_autotune", "flb_batch_set_halo", "flb_set_ghost_overlap",
    "flb_halo_create", "flb_halo_destroy", "flb_halo_info", "flb_halo_exchange",
    "flb_refine_create", "flb_refine_destroy", "flb_refine_info", "flb_refine_set_levels", "flb_refine_average",
    "flb_refine_interpolate", "flb_refine_communicate", "flb_refine_run",
    "flb_download_fields", "flb_enable_tolerance", "flb_tolerance", "flb_store_moments", "flb_write_vtm", "flb_adjoint_create", "flb_adjoint_destroy",
    "flb_adjoint_set_solid", "flb_adjoint_set_fraction", "flb_adjoint_set_fraction_at",
    "flb_adjoint_init_equilibrium", "flb_adjoint_forward", "flb_adjoint_backward", "flb_adjoint_state",
    "flb_adjoint_timing", "flb_adjoint_batch_create", "flb_adjoint_batch_destroy", "flb_adjoint_batch_forward",
    "flb_adjoint_batch_backward", "flb_adjoint_batch_timing", "flb_bandwidth_probe",
    "flb_bandwidth_probe_ex", "flb_reduce_mass_momentum_noalloc",
)


class BlockDesc(C.Structure):
    _fields_ = [("q", C.c_int), ("storage", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("overlap", C.c_int), ("periodic", C.c_int * 3),
                ("omega", C.c_double), ("stream", C.c_void_p)]


class Task(C.Structure):
    _fields_ = [("mask", C.c_uint8), ("dynamics", C.c_uint8), ("pad", C.c_uint8 * 2),
                ("smagorinsky", C.c_float)]


class HWBC(C.Structure):
    _fields_ = [("cell_mask", C.c_uint8), ("void_mask", C.c_uint8), ("kind", C.c_uint8),
                ("pad", C.c_uint8)]


class BlockInfo(C.Structure):
    _fields_ = [("q", C.c_int), ("storage", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("overlap", C.c_int), ("periodic", C.c_int * 3),
                ("n_cells", C.c_int64), ("n_interior", C.c_int64), ("stride", C.c_int64),
                ("step", C.c_int64), ("parity", C.c_int), ("d_pops", C.c_void_p),
                ("d_flags", C.c_void_p)]


class Layout(C.Structure):
    _fields_ = [("pitch", C.c_int), ("xs", C.c_int), ("plane", C.c_int64), ("stride", C.c_int64),
                ("front_pad", C.c_int64), ("dpop", C.c_int64), ("ypop", C.c_int64), ("zpop", C.c_int64),
                ("pop_elems", C.c_int64), ("pop_bytes", C.c_int64)]


class LaunchCfg(C.Structure):
    _fields_ = [("waves_per_group", C.c_int), ("grid", C.c_int)]


class AdjointDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("periodic", C.c_int * 3),
                ("storage", C.c_int), ("omega", C.c_double), ("force", C.c_double * 3),
                ("brinkman_alpha", C.c_double), ("outlet_x", C.c_int), ("stream", C.c_void_p)]


class FLBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[flb error {code}] {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load libflb_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m freelb_amd.build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    L.flb_last_error.restype = C.c_char_p
    L.flb_last_error.argtypes = []
    L.flb_abi_version.restype = C.c_int
    sig = {
        "flb_block_layout": [C.POINTER(BlockDesc), C.POINTER(Layout)],
        "flb_block_create": [C.POINTER(BlockDesc), C.POINTER(P)],
        "flb_block_destroy": [P],
        "flb_get_info": [P, C.POINTER(BlockInfo)],
        "flb_synchronize": [P],
        "flb_set_tasks": [P, C.POINTER(Task), C.c_int],
        "flb_set_halfway_bc": [P, C.POINTER(HWBC), C.c_int],
        "flb_set_mrt_rates": [P, C.POINTER(C.c_double), C.c_int],
        "flb_set_omega": [P, C.c_double],
        "flb_set_force": [P, C.POINTER(C.c_double), C.c_int],
        "flb_share_velocity": [P, P],
        "flb_add_buoyancy": [P, P, C.c_uint8, C.c_double, C.c_double],
        "flb_clear_buoyancy": [P],
        "flb_upload_flags": [P, C.POINTER(C.c_uint8)],
        "flb_set_field_rho": [P, C.POINTER(C.c_double)],
        "flb_set_field_u": [P, C.POINTER(C.c_double)],
        "flb_upload_pops_f32": [P, C.POINTER(C.c_float)],
        "flb_upload_pops_f64": [P, C.POINTER(C.c_double)],
        "flb_download_pops_f32": [P, C.POINTER(C.c_float)],
        "flb_download_pops_f64": [P, C.POINTER(C.c_double)],
        "flb_init_equilibrium": [P, C.POINTER(C.c_double)],
        "flb_apply_cell_dynamics": [P, C.c_int64],
        "flb_stream": [P, C.c_int64],
        "flb_set_level": [P, C.c_int, C.c_int],
        "flb_set_watchdog": [P, C.c_int],
        "flb_apply_halfway_bc": [P],
        "flb_step": [P, C.c_int],
        "flb_moments": [P, C.c_uint8, P],
        "flb_reduce_mass_momentum": [P, C.c_uint8, C.POINTER(C.c_double)],
        "flb_batch_create": [C.POINTER(P), C.c_int, P, C.POINTER(P)],
        "flb_batch_destroy": [P],
        "flb_batch_info": [P, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(LaunchCfg)],
        "flb_batch_set_launch": [P, C.POINTER(LaunchCfg)],
        "flb_batch_step": [P, C.c_int],
        "flb_batch_capture": [P, C.c_int],
        "flb_batch_replay": [P],
        "flb_batch_synchronize": [P],
        "flb_batch_autotune": [P, C.c_int, C.POINTER(LaunchCfg), C.POINTER(C.c_double)],
        "flb_batch_set_halo": [P, P],
        "flb_set_ghost_overlap": [P, C.c_int],
        "flb_halo_create": [C.POINTER(P), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                            C.POINTER(C.c_int), P, C.POINTER(P)],
        "flb_halo_destroy": [P],
        "flb_halo_info": [P, C.POINTER(C.c_int), C.POINTER(C.c_int64)],
        "flb_halo_exchange": [P],
        "flb_refine_create": [P, P, C.POINTER(C.c_int), P, C.POINTER(P)],
        "flb_refine_destroy": [P],
        "flb_refine_info": [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
        "flb_refine_set_levels": [P, C.c_int, C.c_int],
        "flb_refine_average": [P, C.c_int64],
        "flb_refine_interpolate": [P, C.c_int64],
        "flb_refine_communicate": [P, C.c_int64],
        "flb_refine_run": [P, C.c_int64, C.c_int],
        "flb_download_fields": [P, C.POINTER(C.c_double)],
        "flb_enable_tolerance": [P],
        "flb_tolerance": [P, C.c_int, C.c_int, C.POINTER(C.c_double)],
        "flb_store_moments": [P, C.c_uint8],
        "flb_write_vtm": [C.POINTER(P), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.POINTER(C.c_int), C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_char_p,
                          C.c_char_p],
        "flb_adjoint_create": [C.POINTER(AdjointDesc), C.POINTER(P)],
        "flb_adjoint_destroy": [P],
        "flb_adjoint_set_solid": [P, C.POINTER(C.c_uint8)],
        "flb_adjoint_set_fraction": [P, C.POINTER(C.c_double)],
        "flb_adjoint_set_fraction_at": [P, C.c_int64, C.c_double],
        "flb_adjoint_init_equilibrium": [P, C.POINTER(C.c_double)],
        "flb_adjoint_forward": [P, C.c_int, C.c_int, C.POINTER(C.c_double)],
        "flb_adjoint_backward": [P, C.POINTER(C.c_double)],
        "flb_adjoint_state": [P, C.POINTER(C.c_double)],
        "flb_adjoint_timing": [P, C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "flb_adjoint_batch_create": [C.POINTER(P), C.c_int, P, C.POINTER(P)],
        "flb_adjoint_batch_destroy": [P],
        "flb_adjoint_batch_forward": [P, C.c_int, C.c_int, C.POINTER(C.c_double)],
        "flb_adjoint_batch_backward": [P, C.POINTER(C.c_double)],
        "flb_adjoint_batch_timing": [P, C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "flb_bandwidth_probe": [C.c_size_t, C.c_int, C.POINTER(C.c_double)],
        "flb_reduce_mass_momentum_noalloc": [P, C.c_uint8, C.c_void_p],
        "flb_bandwidth_probe_ex": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_int,
                                   C.POINTER(C.c_double)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != FLB_OK:
        raise FLBError(rc, lib().flb_last_error().decode(errors="replace"))


def ptr(a: np.ndarray, T):
    return a.ctypes.data_as(C.POINTER(T))
