"""ctypes binding of libbbwadg_hip.so (the C ABI declared in include/bbwadg.h).

The library is loaded from the package directory; there is no fallback: a
missing or unloadable library raises ``BBWLibraryError``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbbwadg_hip.so")

BBW_OK = 0
BBW_ERR_ARG = -1
BBW_ERR_HIP = -2
BBW_ERR_STATE = -3
BBW_ERR_UNSUPPORTED = -4
BBW_ERR_NOMEM = -5
MODE_PLAIN = 0
MODE_FUSED = 1
MODE_CURVED = 2

# every symbol include/bbwadg.h declares
EXPORTS = [
    "bbw_create", "bbw_destroy", "bbw_last_error", "bbw_set_stream", "bbw_sync",
    "bbw_upload_operators", "bbw_upload_mesh", "bbw_upload_media", "bbw_upload_elastic_media",
    "bbw_set_source", "bbw_set_receivers", "bbw_get_traces", "bbw_set_state", "bbw_get_state",
    "bbw_set_res", "bbw_get_res", "bbw_set_rhs", "bbw_get_rhs", "bbw_volume", "bbw_surface",
    "bbw_update_wadg", "bbw_update", "bbw_stage_fused", "bbw_step", "bbw_energy", "bbw_energy_wadg",
    "bbw_energy_noalloc", "bbw_upload_media_device", "bbw_misfit_device", "bbw_adjoint_gradient_device",
    "bbw_upload_elastic_media_device",
    "bbw_device_ptr", "bbw_set_option", "bbw_upload_adjoint_maps", "bbw_misfit",
    "bbw_adjoint_gradient", "bbw_batch_create", "bbw_batch_destroy", "bbw_batch_last_error",
    "bbw_batch_upload_operators", "bbw_batch_add_mesh", "bbw_batch_add_shot", "bbw_batch_run",
    "bbw_batch_get_state", "bbw_batch_reset", "bbw_batch_last_ms", "bbw_batch_set_step_cap",
    "bbw_upload_bbwadg", "bbw_update_bbwadg", "bbw_upload_slice_lift", "bbw_upload_curved",
    "bbw_curved_rhs",
]
OPT_FUSED_IMPL, OPT_KBLK_V, OPT_KBLK_S, OPT_KBLK_U, OPT_KBLK_F, OPT_ADJ_CACHE_MB, OPT_ADJ_IMPL = 0, 1, 2, 3, 4, 5, 6
OPT_PLAIN_IMPL = 7
OPT_SRC_STEP_TIME = 8
OPT_PLANAR_SPLIT = 9
OPT_REC_FIELD = 10
OPT_ELAS_FUSED = 11
ELAS_LIFT_NODES = 0
ELAS_LIFT_CUBATURE = 1
OPT_XCD_CHUNK = 12
FUSED_VECTOR, FUSED_MFMA = 0, 1
PLAIN_FLAT, PLAIN_TILED, PLAIN_SLICE = 0, 1, 2


class BBWLibraryError(RuntimeError):
    pass


class BBWError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bbw error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("Nfields", ctypes.c_int), ("K", ctypes.c_int),
                ("Nq", ctypes.c_int), ("KblkV", ctypes.c_int), ("KblkS", ctypes.c_int),
                ("KblkU", ctypes.c_int), ("KblkF", ctypes.c_int), ("device", ctypes.c_int)]


class Ops(ctypes.Structure):
    _fields_ = [("Fmask", ctypes.c_void_p), ("D_ids", ctypes.c_void_p), ("Dvals", ctypes.c_void_p),
                ("EEL_ids", ctypes.c_void_p), ("EEL_vals", ctypes.c_void_p), ("EEL_nnz", ctypes.c_int),
                ("L0_ids", ctypes.c_void_p), ("L0_vals", ctypes.c_void_p), ("L0_nnz", ctypes.c_int),
                ("Vq", ctypes.c_void_p), ("Pq", ctypes.c_void_p), ("wq", ctypes.c_void_p),
                ("Vq64", ctypes.c_void_p),
                ("rk4a", ctypes.c_void_p), ("rk4b", ctypes.c_void_p), ("rk4c", ctypes.c_void_p)]


class CurvedOps(ctypes.Structure):
    _fields_ = [("Nfq", ctypes.c_int)] + [(n, ctypes.c_void_p) for n in (
        "Vq", "Vrq", "Vsq", "Vtq", "Pq", "Prq", "Psq", "Ptq", "VfqFace", "Pfq", "Dr", "Ds", "Dt",
        "LIFT", "Fmask", "wq",
        "rk4a", "rk4b", "rk4c")]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BBWLibraryError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    # torch ships its own libamdhip64.so.7; the library must bind to the same HIP runtime as
    # torch (one runtime per process), so torch is loaded first when it is installed
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise BBWLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    P, I, F, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
    sig = {
        "bbw_create": ([ctypes.POINTER(Config), ctypes.POINTER(P)], I),
        "bbw_destroy": ([P], None),
        "bbw_last_error": ([P], ctypes.c_char_p),
        "bbw_set_stream": ([P, P], I),
        "bbw_sync": ([P], I),
        "bbw_upload_operators": ([P, ctypes.POINTER(Ops)], I),
        "bbw_upload_mesh": ([P, P, P, P, P], I),
        "bbw_upload_media": ([P, P, P], I),
        "bbw_upload_elastic_media": ([P, P, P, P, F, F], I),
        "bbw_set_source": ([P, I, P, P, P, D, D], I),
        "bbw_set_receivers": ([P, I, P, P, I], I),
        "bbw_get_traces": ([P, P, I], I),
        "bbw_set_state": ([P, P], I), "bbw_get_state": ([P, P], I),
        "bbw_set_res": ([P, P], I), "bbw_get_res": ([P, P], I),
        "bbw_set_rhs": ([P, P], I), "bbw_get_rhs": ([P, P], I),
        "bbw_volume": ([P], I), "bbw_surface": ([P], I),
        "bbw_update_wadg": ([P, F, F, F, F], I),
        "bbw_update": ([P, F, F, F], I),
        "bbw_stage_fused": ([P, F, F, F, F, I], I),
        "bbw_step": ([P, I, D, D, I, I], I),
        "bbw_energy": ([P, ctypes.POINTER(D)], I),
        "bbw_energy_wadg": ([P, ctypes.POINTER(D)], I),
        "bbw_energy_noalloc": ([P, ctypes.POINTER(D)], I),
        "bbw_device_ptr": ([P, I, ctypes.POINTER(P)], I),
        "bbw_set_option": ([P, I, I], I),
        "bbw_upload_adjoint_maps": ([P, P], I),
        "bbw_upload_bbwadg": ([P, P, P, P, P, P, P, P], I),
        "bbw_upload_slice_lift": ([P, P, P, P, P], I),
        "bbw_upload_curved": ([P, ctypes.POINTER(CurvedOps), P, P, P, P, I, P, I, P], I),
        "bbw_curved_rhs": ([P], I),
        "bbw_update_bbwadg": ([P], I),
        "bbw_misfit": ([P, I, D, D, P, ctypes.POINTER(D)], I),
        "bbw_misfit_device": ([P, I, D, D, P, ctypes.POINTER(D)], I),
        "bbw_adjoint_gradient_device": ([P, I, D, D, P, I, P, ctypes.POINTER(D)], I),
        "bbw_upload_media_device": ([P, P], I),
        "bbw_upload_elastic_media_device": ([P, P, P, P, F, F], I),
        "bbw_adjoint_gradient": ([P, I, D, D, P, I, P, ctypes.POINTER(D)], I),
        "bbw_batch_create": ([I, I, ctypes.POINTER(P)], I),
        "bbw_batch_destroy": ([P], None),
        "bbw_batch_last_error": ([P], ctypes.c_char_p),
        "bbw_batch_upload_operators": ([P, ctypes.POINTER(Ops)], I),
        "bbw_batch_add_mesh": ([P, I, P, P, P, ctypes.POINTER(I)], I),
        "bbw_batch_add_shot": ([P, I, P, P, I, P, P, P, D, D, D, I, ctypes.POINTER(I)], I),
        "bbw_batch_run": ([P, I, I, I], I),
        "bbw_batch_get_state": ([P, I, P], I),
        "bbw_batch_reset": ([P], I),
        "bbw_batch_last_ms": ([P], D),
        "bbw_batch_set_step_cap": ([P, I], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def ptr(a):
    """Address of a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


def check(ctx, rc):
    if rc != BBW_OK:
        msg = load().bbw_last_error(ctx)
        raise BBWError(rc, msg.decode() if msg else "")
    return rc


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
