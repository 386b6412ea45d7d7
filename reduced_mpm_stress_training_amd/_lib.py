"""ctypes binding of libmpm_hip.so (C-ABI in include/mpm_hip.h).

The library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

from ._build import LIB_PATH

MPM_ABI_VERSION = 15
MAX_COLLIDERS = 16

# plane indices of the particle state buffer
P_X0, P_U, P_V, P_C, P_F, P_TAU, P_MASS, P_VOL, P_YIELD, P_PID = 0, 3, 6, 9, 18, 27, 36, 37, 38, 39
NPLANES = 40
# adjoint planes
ADJ_X, ADJ_V, ADJ_C, ADJ_F, ADJ_TAU = 0, 3, 6, 15, 24
ADJ_NPLANES = 33


class MpmParams(C.Structure):
    _fields_ = [("n_grid", C.c_int32), ("material", C.c_int32),
                ("dx", C.c_float), ("inv_dx", C.c_float),
                ("mu", C.c_float), ("lam", C.c_float), ("alpha", C.c_float),
                ("gravity", C.c_float * 3), ("hardening", C.c_int32), ("xi", C.c_float)]


class MpmCollider(C.Structure):
    _fields_ = [("kind", C.c_int32), ("surface_type", C.c_int32),
                ("point", C.c_float * 3), ("normal", C.c_float * 3),
                ("start_time", C.c_float), ("end_time", C.c_float),
                ("friction", C.c_float), ("velocity", C.c_float * 3),
                ("threshold", C.c_float), ("radius", C.c_float), ("height", C.c_float), ("width", C.c_float),
                ("v_scale", C.c_float), ("x_unit", C.c_float * 3), ("y_unit", C.c_float * 3),
                ("reset", C.c_int32)]


class MpmState(C.Structure):
    _fields_ = [("n", C.c_int64), ("stride", C.c_int64), ("planes", C.c_void_p)]


class MpmBins(C.Structure):
    _fields_ = [("nb", C.c_int32), ("bin_cap", C.c_int32), ("overflow_cap", C.c_int32), ("cur", C.c_int32),
                ("epoch", C.c_int32), ("since_sort", C.c_int32),
                ("stride", C.c_int64), ("planes", C.c_void_p * 2), ("count", C.c_void_p),
                ("wg_threads", C.c_int32), ("sort_period", C.c_int32), ("record_acc", C.c_void_p)]


class MpmScene(C.Structure):
    _fields_ = [("params", MpmParams), ("n_particles", C.c_int64), ("n_colliders", C.c_int32),
                ("bin_cap", C.c_int32)]


class MpmBatch(C.Structure):
    _fields_ = [("n_scenes", C.c_int32), ("n_bins", C.c_int32), ("n_colliders", C.c_int32), ("max_cap", C.c_int32),
                ("n_slots", C.c_int64), ("n_particles", C.c_int64), ("n_nodes", C.c_int64),
                ("workspace", C.c_void_p)]


class CromNet(C.Structure):
    _fields_ = [("in_dim", C.c_int32), ("out_dim", C.c_int32), ("x_col", C.c_int32), ("reserved", C.c_int32),
                ("weight", C.c_void_p * 6), ("bias", C.c_void_p * 6),
                ("mean_x", C.c_void_p), ("std_x", C.c_void_p), ("mean_q", C.c_void_p), ("std_q", C.c_void_p)]


# exported symbol -> (restype, argtypes)
_SIGNATURES = {
    "mpm_abi_version": (C.c_int, []),
    "mpm_last_error": (C.c_char_p, []),
    "mpm_zero_grid": (C.c_int, [C.POINTER(MpmParams), C.c_void_p, C.c_void_p]),
    "mpm_p2g": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_float, C.c_void_p]),
    "mpm_grid_update": (C.c_int, [C.POINTER(MpmParams), C.c_void_p, C.c_void_p, C.POINTER(MpmCollider),
                                  C.c_int, C.c_float, C.c_float, C.c_void_p]),
    "mpm_g2p": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_float, C.c_void_p]),
    "mpm_workspace_bytes": (C.c_int64, [C.POINTER(MpmParams), C.c_int64]),
    "mpm_step": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.POINTER(MpmCollider), C.c_int,
                           C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p]),
    "mpm_noalloc_max_grid": (C.c_int, []),
    "mpm_step_noalloc": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.POINTER(MpmCollider), C.c_int,
                                   C.c_float, C.c_float, C.c_void_p]),
    "mpm_reduced_workspace_bytes": (C.c_int64, [C.POINTER(MpmParams)]),
    "mpm_step_reduced": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_int64, C.c_void_p,
                                   C.POINTER(MpmCollider), C.c_int, C.c_void_p, C.c_int64, C.c_float, C.c_float,
                                   C.c_void_p]),
    "mpm_g2p_sample": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_int64, C.c_void_p,
                                 C.c_float, C.c_void_p]),
    "mpm_mark_cells_grid_v": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_int64, C.c_void_p,
                                        C.c_void_p]),
    "mpm_bins_workspace_bytes": (C.c_int64, [C.POINTER(MpmParams)]),
    "mpm_bins_insert": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.POINTER(MpmBins), C.c_void_p,
                                  C.c_void_p]),
    "mpm_bins_extract": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.POINTER(MpmState), C.c_void_p]),
    "mpm_bins_extract_planes": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.POINTER(MpmState), C.c_uint64,
                                          C.c_void_p]),
    "mpm_bins_p2g": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.c_void_p, C.c_float, C.c_void_p]),
    "mpm_bins_status_offset": (C.c_int64, [C.POINTER(MpmParams)]),
    "mpm_bins_velocity_offset": (C.c_int64, [C.POINTER(MpmParams)]),
    "mpm_bins_clear_grid": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.c_void_p, C.c_void_p]),
    "mpm_bins_settle_grid": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.c_void_p, C.c_void_p]),
    "mpm_bins_substep": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.POINTER(MpmCollider), C.c_int,
                                   C.c_void_p, C.c_float, C.c_float, C.c_void_p]),
    "mpm_bins_grid_update": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.POINTER(MpmCollider), C.c_int,
                                       C.c_void_p, C.c_float, C.c_float, C.c_void_p]),
    "mpm_bins_g2p2g": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.c_void_p, C.c_float, C.c_void_p]),
    "mpm_bins_profile": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.POINTER(MpmCollider), C.c_int,
                                   C.c_void_p, C.c_double, C.c_float, C.c_int, C.c_void_p,
                                   C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mpm_bins_sort": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.c_void_p, C.c_void_p]),
    "mpm_bins_run": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmBins), C.POINTER(MpmCollider), C.c_int,
                               C.c_void_p, C.c_double, C.c_float, C.c_int, C.c_void_p]),
    "mpm_batch_workspace_bytes": (C.c_int64, [C.POINTER(MpmScene), C.c_int]),
    "mpm_batch_setup": (C.c_int, [C.POINTER(MpmScene), C.c_int, C.POINTER(MpmCollider), C.c_void_p,
                                  C.POINTER(MpmBatch)]),
    "mpm_batch_status_offset": (C.c_int64, [C.POINTER(MpmBatch)]),
    "mpm_batch_insert": (C.c_int, [C.POINTER(MpmBatch), C.POINTER(MpmState), C.POINTER(MpmBins), C.c_void_p]),
    "mpm_batch_extract": (C.c_int, [C.POINTER(MpmBatch), C.POINTER(MpmBins), C.POINTER(MpmState), C.c_void_p]),
    "mpm_batch_clear_grid": (C.c_int, [C.POINTER(MpmBatch), C.POINTER(MpmBins), C.c_void_p]),
    "mpm_batch_p2g": (C.c_int, [C.POINTER(MpmBatch), C.POINTER(MpmBins), C.c_float, C.c_void_p]),
    "mpm_batch_run": (C.c_int, [C.POINTER(MpmBatch), C.POINTER(MpmBins), C.c_double, C.c_float, C.c_int,
                                C.c_void_p]),
    "mpm_batch_graph_create": (C.c_int, [C.POINTER(MpmBatch), C.POINTER(MpmBins), C.c_float, C.c_int,
                                         C.POINTER(C.c_void_p)]),
    "mpm_batch_graph_launch": (C.c_int, [C.c_void_p, C.POINTER(MpmBatch), C.POINTER(MpmBins), C.c_double, C.c_int,
                                         C.c_void_p]),
    "mpm_graph_destroy": (C.c_int, [C.c_void_p]),
    "mpm_adjoint_workspace_bytes": (C.c_int64, [C.POINTER(MpmParams)]),
    "mpm_step_grids_bytes": (C.c_int64, [C.POINTER(MpmParams)]),
    "mpm_adjoint_substep": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_void_p,
                                      C.POINTER(MpmCollider), C.c_int, C.c_void_p, C.c_float, C.c_float, C.c_void_p,
                                      C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mpm_adjoint_segment": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_void_p,
                                      C.POINTER(MpmCollider), C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mpm_record_segment": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_void_p,
                                     C.POINTER(MpmCollider), C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                     C.POINTER(MpmState), C.c_void_p, C.c_int64, C.c_void_p]),
    "mpm_adjoint_recorded": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.c_void_p,
                                       C.POINTER(MpmCollider), C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_int,
                                       C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mpm_run_ordered_workspace_bytes": (C.c_int64, [C.POINTER(MpmParams)]),
    "mpm_run_ordered": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.POINTER(MpmCollider),
                                  C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_void_p]),
    "mpm_order_bytes": (C.c_int64, [C.POINTER(MpmParams), C.c_int64]),
    "mpm_order_perm_offset": (C.c_int64, [C.POINTER(MpmParams), C.c_int64]),
    "mpm_sort_by_bin": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.POINTER(MpmState), C.c_void_p,
                                  C.c_void_p]),
    "mpm_step_ordered": (C.c_int, [C.POINTER(MpmParams), C.POINTER(MpmState), C.c_void_p, C.POINTER(MpmCollider),
                                   C.c_int, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p]),
    "mpm_permute": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_int64, C.c_int,
                              C.c_void_p]),
    "crom_conv_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crom_conv_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "crom_dense_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "crom_dense_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crom_latent_grad": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "crom_packed_bytes": (C.c_int64, []),
    "crom_acts_bytes": (C.c_int64, [C.c_int64]),
    "crom_pack": (C.c_int, [C.POINTER(CromNet), C.c_void_p, C.c_void_p]),
    "crom_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crom_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
    "crom_param_count": (C.c_int64, [C.c_int32, C.c_int32]),
    "crom_wgrad_workspace_bytes": (C.c_int64, [C.c_int64]),
    "crom_wgrad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                             C.c_void_p]),
    "crom_adam": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float,
                            C.c_float, C.c_float, C.c_int32, C.c_void_p]),
    "mpm_import": (C.c_int, [C.POINTER(MpmState), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "mpm_export": (C.c_int, [C.POINTER(MpmState), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
}

_lib = None


class MpmError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with "
            "`python -m reduced_mpm_stress_training_amd._build` (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.mpm_abi_version() != MPM_ABI_VERSION:
        raise ImportError(f"{LIB_PATH}: ABI version {lib.mpm_abi_version()} != {MPM_ABI_VERSION}")
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().mpm_last_error().decode(errors="replace")
        raise MpmError(f"{what} failed ({rc}): {msg}")


def exported_symbols():
    return list(_SIGNATURES)
