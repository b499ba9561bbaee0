"""ctypes binding of libdiffstir_hip.so (include/diffstir.h).

The library is loaded from this package directory.  There is no fallback: if the library is
missing or fails to load, every entry point raises `LibraryNotBuilt` (build it with
`python -m diffstir_amd.build`).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiffstir_hip.so")

# the exported entry points of include/diffstir.h
EXPORTS = (
    "ds_create", "ds_destroy", "ds_set_stream", "ds_set_option", "ds_last_error", "ds_set_particles", "ds_set_trajectory", "ds_set_params",
    "ds_reset", "ds_forward", "ds_current_step", "ds_get_forces", "ds_backward", "ds_read_liquid",
    "ds_read_stick", "ds_fallback_count", "ds_substep_noalloc", "ds_synchronize", "ds_last_timing", "ds_kernel_timing", "ds_svd3", "ds_svd3_values", "ds_svd3_backward",
)

KERNEL_NAMES = ("k_p2g_tiled", "k_grid_reduce", "k_g2p_tiled", "k_g2p_adj_tiled", "k_vbar_reduce",
                "k_p2g_adj_tiled", "bucket_sort", "k_p2g_adj_stress", "k_g2p_tiled_factor", "k_p2g_adj_xgrad",
                "k_g2p_p2g_tiled")

ERROR_NAMES = {0: "DS_OK", 1: "DS_EINVAL", 2: "DS_EHIP", 3: "DS_ENOMEM", 4: "DS_ESTATE", 5: "DS_ENONFINITE"}
DS_ENONFINITE = 5


class LibraryNotBuilt(RuntimeError):
    pass


class DiffStirError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class SceneDesc(C.Structure):
    _fields_ = [("n_grid", C.c_int32), ("bound", C.c_int32), ("model", C.c_int32), ("k_from_param", C.c_int32),
                ("dt", C.c_double), ("gravity", C.c_double), ("k_bulk", C.c_double), ("wall_r", C.c_double),
                ("wall_cx", C.c_double), ("wall_cz", C.c_double), ("m_liq", C.c_double), ("vol_liq", C.c_double),
                ("m_stick", C.c_double), ("m_cont", C.c_double)]


_lib = None

F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int32)


def load() -> C.CDLL:
    """Load (once) and return the library; raises LibraryNotBuilt if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryNotBuilt(f"{LIB_PATH} is missing: build it with `python -m diffstir_amd.build`")
    # A torch wheel carries its own HIP runtime under ROCm's soname (libamdhip64.so.7), and the
    # first one loaded serves the whole process; loaded after this library, torch cannot bring up
    # its device.  So torch is imported first when it is installed: torch tensors and streams and
    # this library then share one runtime (a stream handle from torch is valid here).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as e:
        raise LibraryNotBuilt(f"cannot load {LIB_PATH}: {e}") from e
    vp = C.c_void_p
    sig = {
        "ds_create": (C.c_int, [C.POINTER(SceneDesc), C.c_int32, I32P, I32P, I32P, C.c_int32, C.c_int32,
                                C.POINTER(vp)]),
        "ds_destroy": (None, [vp]),
        "ds_set_stream": (C.c_int, [vp, vp]),
        "ds_set_option": (C.c_int, [vp, C.c_int32, C.c_int32]),
        "ds_last_error": (C.c_char_p, [vp]),
        "ds_set_particles": (C.c_int, [vp, C.c_int32, F32P, F32P, F32P]),
        "ds_set_trajectory": (C.c_int, [vp, C.c_int32, F32P]),
        "ds_set_params": (C.c_int, [vp, F64P]),
        "ds_reset": (C.c_int, [vp]),
        "ds_forward": (C.c_int, [vp, C.c_int32]),
        "ds_current_step": (C.c_int, [vp, I32P]),
        "ds_get_forces": (C.c_int, [vp, F64P]),
        "ds_backward": (C.c_int, [vp, F64P, F64P]),
        "ds_read_liquid": (C.c_int, [vp, C.c_int32, F32P, F32P, F32P, F32P]),
        "ds_read_stick": (C.c_int, [vp, C.c_int32, F32P, F32P, F32P]),
        "ds_fallback_count": (C.c_int, [vp, C.POINTER(C.c_int64), C.c_int32]),
        "ds_substep_noalloc": (C.c_int, [C.POINTER(SceneDesc), C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, vp,
                                         vp, vp]),
        "ds_synchronize": (C.c_int, [vp]),
        "ds_last_timing": (C.c_int, [vp, F64P, F64P]),
        "ds_kernel_timing": (C.c_int, [vp, F64P, C.POINTER(C.c_int64), C.c_int32]),
        "ds_svd3": (C.c_int, [F64P, C.c_int32, F64P, F64P, F64P]),
        "ds_svd3_values": (C.c_int, [F64P, C.c_int32, F64P]),
        "ds_svd3_backward": (C.c_int, [F64P, F64P, F64P, F64P, F64P, F64P, C.c_int32, F64P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, ctx=None):
    if code != 0:
        lib = load()
        msg = lib.ds_last_error(ctx)
        raise DiffStirError(code, msg.decode() if msg else "")
