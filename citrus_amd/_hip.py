"""Loader of the HIP library ``libcitrus_hip.so`` (ctypes) and the C-ABI argument structures.

The library is built in-tree by ``citrus_amd.build`` for gfx950. There is no fallback: when the library is missing or
fails to load, every entry point raises ``RuntimeError`` naming the cause.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

LIBRARY_PATH = Path(__file__).resolve().parent / "libcitrus_hip.so"
ABI_VERSION = 20
# points per block of the emission kernels' partial sums (CITRUS_EMISSION_BLOCK of include/citrus_hip.h)
EMISSION_BLOCK = 32
# unused floats after each group's region of the block-order first-collision source (CITRUS_FC_GROUP_PAD)
FC_GROUP_PAD = 1088
DISABLE_ENV_VAR = "CITRUS_DISABLE_HIP"

_lib = None
_load_error: str | None = None

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_fp = ctypes.POINTER(ctypes.c_float)
c_void = ctypes.c_void_p


class SweepArgs(ctypes.Structure):
    """Mirror of ``CitrusSweepArgs`` (include/citrus_hip.h)."""

    _fields_ = [
        ("G", c_int), ("nsh", c_int), ("D8", c_int), ("d0", c_int), ("dc", c_int), ("V", c_int),
        ("accumulate", c_int), ("write_psi", c_int),
        ("mu_t", c_void), ("T", c_void), ("source", c_void), ("phi", c_void),
        ("Y", c_void), ("Yw", c_void), ("cdir", c_void), ("faces", c_void), ("source_layout", c_int),
        ("T_bf16", c_void), ("T_low", c_void),
    ]


class SweepView(ctypes.Structure):
    """Mirror of ``CitrusSweepView`` (include/citrus_hip.h)."""

    _fields_ = [
        ("mu_t", c_void), ("T", c_void), ("T_bf16", c_void), ("T_low", c_void), ("source", c_void),
        ("phi", c_void), ("cdir", c_void), ("fc_scale", c_void), ("fc_beta", c_void), ("fc_u", c_void),
    ]


class PersistentArgs(ctypes.Structure):
    """Mirror of ``CitrusPersistentArgs`` (include/citrus_hip.h)."""

    _fields_ = [
        ("G", c_int), ("nsh", c_int), ("D8", c_int), ("d0", c_int), ("dc", c_int), ("accumulate", c_int),
        ("source_layout", c_int), ("Y", c_void), ("Yw", c_void), ("views", ctypes.POINTER(SweepView)),
        ("status", c_void), ("spin_limit", c_int), ("fc_omega", c_void), ("waves_per_cu", c_int),
    ]


class FcArgs(ctypes.Structure):
    """Mirror of ``CitrusFcArgs`` (include/citrus_hip.h)."""

    _fields_ = [
        ("M", c_int), ("G", c_int), ("n_eta", c_int), ("n_map", c_int),
        ("table", c_void), ("eta_map", c_void),
        ("n_slots", c_int), ("slot_material", c_void), ("slot_density", c_void), ("phi_uc", c_void),
        ("R", c_int), ("C", c_int), ("S", c_int),
        ("lattice", ctypes.c_double * 6), ("focal", ctypes.c_double * 3), ("parallel", c_int),
        ("omegas", c_void), ("K", c_int),
    ]


_SIGNATURES = {
    "citrus_last_error": (ctypes.c_char_p, []),
    "citrus_abi_version": (c_int, []),
    "citrus_sweep_plan_create": (c_int, [c_int, c_int, c_int, c_int, ctypes.POINTER(c_void)]),
    "citrus_sweep_plan_destroy": (None, [c_void]),
    "citrus_sweep_face_floats": (c_i64, [c_void, c_int, c_int, c_int]),
    "citrus_sweep_run": (c_int, [c_void, ctypes.POINTER(SweepArgs), c_void]),
    "citrus_persistent_plan_create": (c_int, [c_int, ctypes.POINTER(c_int), c_int, ctypes.POINTER(c_void)]),
    "citrus_persistent_plan_destroy": (None, [c_void]),
    "citrus_persistent_workspace_bytes": (c_i64, [c_void, c_int]),
    "citrus_persistent_num_tasks": (c_i64, [c_void]),
    "citrus_persistent_schedule_table": (c_i64, [c_int, ctypes.POINTER(c_int), c_int, c_void, c_void, c_void]),
    "citrus_persistent_sweep_run": (c_int, [c_void, ctypes.POINTER(PersistentArgs), c_void, c_i64, c_void]),
    "citrus_persistent_check": (c_int, [c_void, c_void, ctypes.POINTER(c_int), ctypes.POINTER(c_i64), c_void]),
    "citrus_sh_expand": (c_int, [c_void, c_int, c_int, c_void, c_i64, c_void, c_void]),
    "citrus_multiple_emission": (c_int, [c_void, c_int, c_int, c_void, c_int, c_i64, c_void, c_int, c_void, c_void,
                                         c_void, c_void]),
    "citrus_attenuation_to_densities": (c_int, [c_void, c_i64, c_int, c_void, c_void, c_void, c_void, c_void,
                                                c_void]),
    "citrus_resample_conservative": (c_int, [c_void, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                             c_void, c_void, c_void, c_int, c_void, c_void, c_void, c_int,
                                             c_void, c_void, c_void, c_int, c_void, c_void]),
    "citrus_state_pack_workspace_floats": (c_i64, [c_int, c_int, c_int, c_i64]),
    "citrus_state_pack": (c_int, [c_void, c_int, c_int, c_int, c_i64, c_void, c_void, c_void, c_i64, c_void]),
    "citrus_state_pack_static": (c_int, [c_void, c_int, c_int, c_int, c_i64, c_void, c_void, c_void]),
    "citrus_state_unpack": (c_int, [c_void, c_void, c_int, c_int, c_int, c_i64, c_void, c_void]),
    "citrus_scatter_moments_dense": (c_int, [c_void, c_int, c_void, c_void, c_int, c_int, c_i64, c_int, c_void]),
    "citrus_scatter_moments_factored": (c_int, [c_void, c_int, c_void, c_int, c_int, c_void, c_void, c_int, c_int,
                                                c_i64, c_int, c_void, c_void]),
    "citrus_scatter_moments_stored": (c_int, [c_void, c_int, c_void, c_int, c_int, c_void, c_void, c_void, c_int,
                                              c_int, c_i64, c_int, c_void]),
    "citrus_ray_segments": (c_int, [c_void, c_int, c_int, c_int, c_int, c_void, c_void, c_void, c_int, c_i64, c_void,
                                    c_void]),
    "citrus_ray_trace_indices": (c_int, [c_int, c_int, c_int, c_void, c_void, c_void, c_i64, c_int, c_void, c_void,
                                         c_void, c_void]),
    "citrus_ray_lattice": (c_int, [c_void, c_int, c_int, c_int, c_int, c_void, c_void, c_void, c_void, c_void,
                                   c_void, c_void]),
    "citrus_ray_primary": (c_int, [c_void, c_int, c_int, c_int, c_int, c_void, c_void, c_void, c_int, c_void,
                                   c_void]),
    "citrus_ray_oblique": (c_int, [c_void, c_i64, c_int, c_int, c_int, c_int, c_void, c_void, c_void, c_void, c_int,
                                   c_void, c_i64, c_i64, c_void]),
    "citrus_ray_downstream": (c_int, [c_void, c_int, c_int, c_int, c_int, c_void, c_void, c_void, c_void, c_int,
                                      c_void, c_void]),
    "citrus_ray_crossing_sequence_bytes": (c_i64, [c_int, c_int, c_int]),
    "citrus_ray_crossing_sequences": (c_int, [c_int, c_int, c_int, c_void, c_void, c_int, c_void, c_void]),
    "citrus_ray_downstream_shared": (c_int, [c_void, c_int, c_int, c_int, c_int, c_void, c_void, c_void, c_void,
                                             c_int, c_void, c_void, c_int, c_void, c_void]),
    "citrus_group_transmission": (c_int, [c_void, c_int, c_i64, c_void, c_void, c_void, c_int, c_int, c_void, c_void,
                                          c_void, c_void]),
    "citrus_uncollided_fluence": (c_int, [c_void, c_int, c_void, c_void, c_void, c_void, c_void, c_int, c_int,
                                          c_void, c_void, c_void, c_void, c_void, c_void, c_void, c_void]),
    "citrus_bf16_operand": (c_int, [c_void, c_int, c_int, c_i64, c_int, c_void, c_void, c_void]),
    "citrus_scatter_moments_mixed": (c_int, [c_void, c_int, c_void, c_int, c_int, c_void, c_void, c_void, c_int,
                                             c_int, c_i64, c_int, c_void, c_void]),
    "citrus_nn_conv3d_forward": (c_int, [c_void, c_i64, c_int, c_int, c_int, c_int, c_int, c_void, c_void, c_int,
                                         c_int, c_int, c_void, c_i64, c_void]),
    "citrus_nn_conv3d_backward_data": (c_int, [c_void, c_i64, c_int, c_int, c_int, c_int, c_int, c_void, c_int, c_int,
                                               c_void, c_i64, c_void]),
    "citrus_nn_wgrad_workspace_floats": (c_i64, [c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
    "citrus_nn_conv3d_backward_weight": (c_int, [c_void, c_i64, c_void, c_i64, c_int, c_int, c_int, c_int, c_int,
                                                 c_int, c_int, c_void, c_void, c_void, c_void]),
    "citrus_nn_leaky_backward": (c_int, [c_void, c_i64, c_void, c_i64, c_i64, c_int, c_void]),
    "citrus_nn_avg_pool2": (c_int, [c_void, c_i64, c_int, c_int, c_int, c_int, c_int, c_void, c_i64, c_void]),
    "citrus_nn_avg_pool2_backward": (c_int, [c_void, c_i64, c_int, c_int, c_int, c_int, c_int, c_void, c_i64, c_int,
                                             c_void]),
    "citrus_nn_upsample2": (c_int, [c_void, c_i64, c_int, c_int, c_int, c_int, c_int, c_void, c_i64, c_void]),
    "citrus_nn_upsample2_backward": (c_int, [c_void, c_i64, c_int, c_int, c_int, c_int, c_int, c_void, c_i64,
                                             c_void]),
    "citrus_nn_adam": (c_int, [c_void, c_void, c_void, c_void, c_void, c_i64, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, c_int, c_void]),
    "citrus_nn_to_bf16": (c_int, [c_void, c_void, c_i64, c_void]),
    "citrus_fc_source": (c_int, [ctypes.POINTER(FcArgs), c_void, c_void]),
    "citrus_fc_sums": (c_int, [ctypes.POINTER(FcArgs), c_void, c_void, c_int, c_void, c_void, c_void]),
    "citrus_fc_sums_blocked": (c_int, [ctypes.POINTER(FcArgs), c_void, c_void, c_int, c_void, c_void, c_int, c_void,
                                       c_void, c_void, c_void]),
    "citrus_fc_correct": (c_int, [ctypes.POINTER(FcArgs), c_void, c_void, c_void, c_void]),
    "citrus_fc_primary_directions": (c_int, [ctypes.POINTER(FcArgs), c_void, c_void]),
    "citrus_fc_correct_means": (c_int, [c_void, c_void, c_void, c_void, c_void, c_int, c_int, c_i64, c_int, c_void]),
    "citrus_fc_moments": (c_int, [ctypes.POINTER(FcArgs), c_void, c_void, c_int, c_int, c_void, c_void]),
    "citrus_attenuated_emission": (c_int, [ctypes.POINTER(FcArgs), c_void, c_int, c_void, c_void, c_void, c_int,
                                           c_void, c_void, c_void, c_void, c_void]),
    "citrus_attenuated_emission_collapsed": (c_int, [ctypes.POINTER(FcArgs), c_void, c_int, c_void, c_void, c_void,
                                                     c_void, c_void, c_void, c_void]),
    "citrus_box_matrix": (c_int, [c_void, c_int, c_int, c_void, c_void]),
    "citrus_footprint_filters": (c_int, [c_void, c_int, c_int, ctypes.c_float, c_void, ctypes.c_float, ctypes.c_float,
                                         c_int, c_int, c_void, c_void, c_void, c_void, c_void]),
    "citrus_footprint_blur": (c_int, [c_void, c_int, c_int, ctypes.c_float, c_void, ctypes.c_float, ctypes.c_float,
                                      c_int, c_int, c_void, c_void, c_void, c_void, c_void, c_void]),
}


def exported_symbols() -> list[str]:
    """Names of every C-ABI entry point this wrapper binds."""
    return list(_SIGNATURES)


def library() -> ctypes.CDLL:
    """The loaded library; raises ``RuntimeError`` if it is missing, disabled or of the wrong ABI version."""
    global _lib, _load_error
    if os.environ.get(DISABLE_ENV_VAR, "") not in ("", "0"):
        raise RuntimeError(f"the HIP library is disabled by {DISABLE_ENV_VAR}")
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise RuntimeError(_load_error)
    if not LIBRARY_PATH.exists():
        _load_error = f"{LIBRARY_PATH} is missing: build it with `python -m citrus_amd.build`"
        raise RuntimeError(_load_error)
    try:
        lib = ctypes.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
    except (OSError, AttributeError) as e:
        _load_error = f"cannot load {LIBRARY_PATH}: {e}"
        raise RuntimeError(_load_error) from e
    if lib.citrus_abi_version() != ABI_VERSION:
        _load_error = f"{LIBRARY_PATH} has ABI {lib.citrus_abi_version()}, expected {ABI_VERSION}: rebuild it"
        raise RuntimeError(_load_error)
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    """Raise ``RuntimeError`` with the library's message when a call returned a nonzero status."""
    if rc != 0:
        msg = _lib.citrus_last_error().decode() if _lib is not None else ""
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


def ptr(t: torch.Tensor | None) -> int | None:
    """Device address of a tensor (None stays None)."""
    return None if t is None else t.data_ptr()


def stream_of(device: torch.device) -> int:
    """Raw HIP stream handle of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu_tensor(t: torch.Tensor, name: str, dtype=torch.float32) -> None:
    """Raise ``ValueError`` unless ``t`` is a contiguous tensor of ``dtype`` on a GPU."""
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on a GPU device, got {t.device}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
