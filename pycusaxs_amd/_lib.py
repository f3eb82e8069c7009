"""ctypes binding of the C ABI declared in include/saxs_engine.h.

The library is loaded from the package tree (``pycusaxs_amd/_lib``).  There
is no fallback: if it is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from ._build import LIB_PATH

SAXS_OK, SAXS_EINVAL, SAXS_ENOMEM, SAXS_EHIP, SAXS_EUNSUP, SAXS_ECAPACITY = 0, -1, -2, -3, -4, -5
ABI_VERSION = 6

# every symbol include/saxs_engine.h declares
EXPORTED = (
    "saxs_plan_create", "saxs_plan_set_tables", "saxs_forward", "saxs_forward_frames",
    "saxs_backward", "saxs_forward_plain", "saxs_plan_workspace_bytes", "saxs_plan_kept",
    "saxs_plan_destroy", "saxs_last_error", "saxs_abi_version", "saxs_plan_set_timing",
    "saxs_plan_kernel_times", "saxs_backward_ex", "saxs_plan_set_padding",
    "saxs_plan_set_point_tables", "saxs_forward_static", "saxs_forward_boxes",
)
KERNEL_NAMES = ("clear", "prep", "scan", "scatter", "spread_z", "yfft", "xfft_epi", "reduce",
                "adj_x", "adj_y", "adj_z", "grad_sum", "plain")


class SaxsDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int32 * 3),
        ("nn", C.c_int32 * 3),
        ("order", C.c_int32),
        ("num_elements", C.c_int32),
        ("num_bins", C.c_int32),
        ("border", C.c_int32),
        ("lut_len", C.c_int32),
        ("bin_lut", C.POINTER(C.c_int32)),
        ("ff_lut", C.POINTER(C.c_float)),
        ("mod_x", C.POINTER(C.c_float)),
        ("mod_y", C.POINTER(C.c_float)),
        ("mod_z", C.POINTER(C.c_float)),
        ("max_atoms", C.c_int32),
        ("max_frames", C.c_int32),
        ("tile_x", C.c_int32),
        ("tile_y", C.c_int32),
        ("flags", C.c_int32),
        ("kmax", C.c_int32 * 3),
        ("cell", C.c_float * 9),
    ]


class SaxsFrameBoxes(C.Structure):
    _fields_ = [
        ("cells", C.POINTER(C.c_float)),
        ("qgeom", C.POINTER(C.c_float)),
        ("bin_qgeom", C.POINTER(C.c_float)),
        ("key_type", C.c_int32),
        ("kmax", C.c_int32 * 3),
        ("lut_len", C.c_int32),
        ("dq", C.c_float),
        ("qcut", C.c_float),
        ("form_coef", C.POINTER(C.c_float)),
    ]


FLAG_PER_ELEMENT_SPREAD = 1
FLAG_NO_PRUNED_Z = 2
FLAG_WIDE_RADIX = 4
FLAG_FORWARD_ONLY = 8     # tile fitted to the forward kernels alone (no backward)
FLAG_NO_KEY_CLASSES = 16  # sort keys by home tile only (tiles fetch every source atom)
BWD_REUSE_FORWARD = 1


_lib = None


class EngineError(RuntimeError):
    pass


class CapacityError(ValueError):
    """The plan's workspace cannot hold the q tables of a new box (more kept
    frequencies than it was sized for); the plan is unchanged."""


def load(path: Path = LIB_PATH):
    """Loads the engine library (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not Path(path).exists():
        raise ImportError(
            f"MI355X SAXS engine library not built: {path} is missing "
            "(build it with: python -m pycusaxs_amd._build)")
    # torch first: its HIP runtime (libtorch_hip asks for libamdhip64.so by
    # its own rpath) then also serves this library's libamdhip64.so.7.  The
    # other order loads a second runtime, on which torch's streams and
    # pointers are meaningless.
    import torch  # noqa: F401
    lib = C.CDLL(str(path))
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    lib.saxs_plan_create.argtypes = [C.POINTER(SaxsDesc), C.POINTER(vp)]
    lib.saxs_plan_set_tables.argtypes = [vp, i32, vp, vp]
    lib.saxs_plan_set_padding.argtypes = [vp, vp]
    lib.saxs_plan_set_point_tables.argtypes = [vp, vp, vp, i32, vp, vp]
    lib.saxs_forward.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, vp]
    lib.saxs_forward_frames.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, vp]
    lib.saxs_forward_static.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, vp]
    lib.saxs_forward_boxes.argtypes = [vp, vp, vp, i32, i32, C.POINTER(SaxsFrameBoxes), vp, vp, vp]
    lib.saxs_backward.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, vp]
    lib.saxs_backward_ex.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, i32, vp]
    lib.saxs_forward_plain.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp]
    lib.saxs_plan_workspace_bytes.argtypes = [vp]
    lib.saxs_plan_workspace_bytes.restype = i64
    lib.saxs_plan_kept.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    lib.saxs_plan_destroy.argtypes = [vp]
    lib.saxs_plan_set_timing.argtypes = [vp, i32]
    lib.saxs_plan_kernel_times.argtypes = [vp, vp, vp]
    lib.saxs_last_error.restype = C.c_char_p
    lib.saxs_abi_version.restype = i32
    for name in ("saxs_plan_create", "saxs_plan_set_tables", "saxs_forward", "saxs_forward_frames",
                 "saxs_backward", "saxs_forward_plain", "saxs_plan_kept", "saxs_plan_destroy",
                 "saxs_plan_set_timing", "saxs_plan_kernel_times", "saxs_backward_ex",
                 "saxs_plan_set_padding", "saxs_plan_set_point_tables", "saxs_forward_static",
                 "saxs_forward_boxes"):
        getattr(lib, name).restype = C.c_int
    if lib.saxs_abi_version() != ABI_VERSION:
        raise ImportError(f"engine ABI {lib.saxs_abi_version()} != expected {ABI_VERSION}")
    _lib = lib
    return lib


def check(rc: int, what: str):
    """Maps return codes to Python exceptions like cuda_bindings.cpp:151-167:
    invalid arguments and unsupported shapes -> ValueError, else RuntimeError."""
    if rc == SAXS_OK:
        return
    msg = load().saxs_last_error().decode(errors="replace")
    if rc == SAXS_ECAPACITY:
        raise CapacityError(f"{what}: {msg}")
    if rc in (SAXS_EINVAL, SAXS_EUNSUP):
        raise ValueError(f"{what}: {msg}")
    raise EngineError(f"{what}: {msg}")
