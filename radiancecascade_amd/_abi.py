"""ctypes binding of include/rc_gi.h (librcgi.so).

The library is the HIP build in this directory; there is no CPU fallback.
Loading fails loudly when the shared object is missing.  When torch is
importable it is imported first, so that the process holds one HIP runtime
(torch's wheel ships its own libamdhip64.so.7, which then also serves
librcgi.so) and torch tensors can be handed to the library as device memory.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librcgi.so")

ABI_VERSION = 2
MAX_LEVELS = 8
MAX_DIM = 16384
MAX_BOXES = 256

OK, ERR_INVALID, ERR_NULL, ERR_NO_DEVICE, ERR_HIP, ERR_OOM, ERR_STATE = range(7)
STATUS_NAMES = {0: "RC_OK", 1: "RC_ERR_INVALID", 2: "RC_ERR_NULL", 3: "RC_ERR_NO_DEVICE",
                4: "RC_ERR_HIP", 5: "RC_ERR_OOM", 6: "RC_ERR_STATE"}

STORAGE = {"f32": 0, "fp32": 0, "f16": 1, "fp16": 1, "bf16": 2}
FLAG_PROFILE = 0x1
FLAG_KEEP_DEBUG = 0x2
FLAG_GRAPH = 0x8

TILE_SHAPES = {0: (8, 8, 4), 1: (4, 4, 16), 2: (2, 2, 64), 3: (1, 1, 256), 4: (8, 4, 8)}

STAGE_SEED, STAGE_JFA, STAGE_JFA_TAIL, STAGE_LEVEL0 = 0, 1, 2, 3
STAGE_MARCH = STAGE_LEVEL0 + MAX_LEVELS
STAGE_NEAR_SDF, STAGE_MIP, STAGE_CONE, STAGE_SURROGATE, STAGE_FALLBACK = range(STAGE_MARCH + 1, STAGE_MARCH + 6)
NUM_STAGES = STAGE_MARCH + 6

MLP_INPUTS, MLP_OUTPUTS, MLP_PARAMS = 64, 64, 16640
MLP_RADIANCE = 48          # targets / radiance outputs per probe (16 directions x rgb)
SURROGATE_BLOB_MAGIC = b"RCSURR01"
TRAIN_MAX_EPOCHS = 64


class Config(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("levels", ctypes.c_uint32), ("storage", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("device", ctypes.c_int32), ("depth_mode", ctypes.c_uint32), ("z_near", ctypes.c_float),
                ("z_far", ctypes.c_float)]

DEPTH_LINEAR = 0
DEPTH_NDC = 1

SDF_EXACT = 0
SDF_JUMP_FLOOD = 1
SDF_MODES = {"exact": SDF_EXACT, "jfa": SDF_JUMP_FLOOD}


class Plane(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 16), ("channels", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("width", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("data", ctypes.c_void_p)]


class GBuffer(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_void_p), ("albedo", ctypes.c_void_p), ("emissive", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("runs", ctypes.c_uint32),
                ("stage_ms", ctypes.c_double * NUM_STAGES),
                ("stage_launches", ctypes.c_uint32 * NUM_STAGES),
                ("level_rays", ctypes.c_uint64 * MAX_LEVELS),
                ("level_bytes_stored", ctypes.c_uint64 * MAX_LEVELS),
                ("surrogate_frames", ctypes.c_uint64), ("fallback_frames", ctypes.c_uint64),
                ("flagged_tiles", ctypes.c_uint64), ("tiles_per_frame", ctypes.c_uint64)]


class SurrogateConfig(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("mode", ctypes.c_uint32), ("threshold", ctypes.c_float),
                ("reserved", ctypes.c_uint32)]


class TrainConfig(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("levels", ctypes.c_uint32), ("scene_seed0", ctypes.c_uint64), ("scenes", ctypes.c_uint32),
                ("n_boxes", ctypes.c_uint32), ("n_emissive", ctypes.c_uint32), ("epochs", ctypes.c_uint32),
                ("batch", ctypes.c_uint32), ("lr", ctypes.c_float), ("lr_decay", ctypes.c_float),
                ("reserved", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


class TrainReport(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32), ("epochs", ctypes.c_uint32), ("samples", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("epoch_loss", ctypes.c_float * TRAIN_MAX_EPOCHS),
                ("gen_ms", ctypes.c_double), ("train_ms", ctypes.c_double), ("out_scale", ctypes.c_float),
                ("reserved", ctypes.c_uint32)]


class View(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("levels", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("depth", ctypes.c_void_p), ("emissive", ctypes.c_void_p),
                ("irradiance", ctypes.c_void_p)]


BATCH_WORKLIST = 0
BATCH_STREAMS = 1
BATCH_MODES = {"worklist": BATCH_WORKLIST, "streams": BATCH_STREAMS}


class BatchConfig(ctypes.Structure):
    _fields_ = [("storage", ctypes.c_uint32), ("mode", ctypes.c_uint32), ("streams", ctypes.c_uint32),
                ("device", ctypes.c_int32)]


class Box(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("albedo", ctypes.c_double * 3), ("emit", ctypes.c_double * 3)]


# name -> (restype, argtypes); every symbol include/rc_gi.h declares.
_P = ctypes.c_void_p
_u32, _i32, _u64 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
SIGNATURES = {
    "rc_create": (_i32, [ctypes.POINTER(Config), ctypes.POINTER(_P)]),
    "rc_resize": (_i32, [_P, _u32, _u32, _u32]),
    "rc_update_host": (_i32, [_P, ctypes.POINTER(GBuffer)]),
    "rc_render_host": (_i32, [_P, _P]),
    "rc_run_device": (_i32, [_P, ctypes.POINTER(GBuffer), _P, _P]),
    "rc_run_device_noalloc": (_i32, [_u32, _u32, _u32, ctypes.POINTER(GBuffer), _P, _P]),
    "rc_noalloc_scratch_bytes": (ctypes.c_uint64, [_u32, _u32, _u32]),
    "rc_noalloc_last_error": (ctypes.c_char_p, []),
    "rc_destroy": (None, [_P]),
    "rc_set_level_tiling": (_i32, [_P, _u32, _u32, _u32]),
    "rc_get_level_tiling": (_i32, [_P, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "rc_autotune": (_i32, [_P, ctypes.POINTER(GBuffer), _P, _u32, _P]),
    "rc_set_jfa_fuse": (_i32, [_P, _u32]),
    "rc_get_jfa_fuse": (_u32, [_P]),
    "rc_set_jfa_rows": (_i32, [_P, _u32]),
    "rc_get_jfa_rows": (_u32, [_P]),
    "rc_set_sdf_mode": (_i32, [_P, _u32]),
    "rc_get_sdf_mode": (_u32, [_P]),
    "rc_set_march": (_i32, [_P, _u32, _u32, _u32]),
    "rc_set_march_order": (_i32, [_P, _u32, _u32]),
    "rc_get_march_order": (_i32, [_P, _u32, ctypes.POINTER(_u32)]),
    "rc_get_march": (_i32, [_P, ctypes.POINTER(_u32), ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "rc_get_config": (_i32, [_P, ctypes.POINTER(Config)]),
    "rc_get_stats": (_i32, [_P, ctypes.POINTER(Stats)]),
    "rc_reset_stats": (_i32, [_P]),
    "rc_last_error": (ctypes.c_char_p, [_P]),
    "rc_status_string": (ctypes.c_char_p, [_i32]),
    "rc_abi_version": (_u32, []),
    "rc_device_count": (_i32, [ctypes.POINTER(ctypes.c_int)]),
    "rc_debug_seeds": (_i32, [_P, _P, _P]),
    "rc_debug_sdf": (_i32, [_P, _P, _P]),
    "rc_debug_level": (_i32, [_P, _u32, _P, _P]),
    "rc_debug_march": (_i32, [_P, _u32, _P, _P, _P]),
    "rc_error_norms": (_i32, [_P, _P, _P, _u32, _u64, _P, _P]),
    "rc_probe_grid": (_i32, [_u32, _u32, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "rc_level_rays": (_u64, [_u32, _u32, _u32]),
    "rc_levels_for_coverage": (_u32, [_u32, _u32]),
    "rc_set_surrogate": (_i32, [_P, ctypes.POINTER(SurrogateConfig)]),
    "rc_get_surrogate": (_i32, [_P, ctypes.POINTER(SurrogateConfig)]),
    "rc_mlp_set": (_i32, [_P, _P, _P, _P, ctypes.c_float]),
    "rc_mlp_get": (_i32, [_P, _P, _P, _P, ctypes.POINTER(ctypes.c_float)]),
    "rc_surrogate_save": (_i32, [_P, _P, _u64, ctypes.POINTER(_u64)]),
    "rc_surrogate_load": (_i32, [_P, _P, _u64]),
    "rc_mlp_forward": (_i32, [_P, _P, _u32, _P, _P]),
    "rc_mlp_gradient": (_i32, [_P, _P, _P, _u32, _P, _P, _P]),
    "rc_mlp_adam": (_i32, [_P, _P, ctypes.c_float, _P]),
    "rc_surrogate_samples": (_i32, [_P, ctypes.POINTER(GBuffer), _P, _P, _P]),
    "rc_surrogate_train": (_i32, [_P, ctypes.POINTER(TrainConfig), ctypes.POINTER(TrainReport)]),
    "rc_host_alloc": (_i32, [_u64, ctypes.POINTER(_P)]),
    "rc_host_free": (None, [_P]),
    "rc_surrogate_train_rows": (_i32, [_P, _P, _P, ctypes.c_uint64, ctypes.POINTER(TrainConfig),
                                       ctypes.POINTER(TrainReport)]),
    "rc_debug_surrogate": (_i32, [_P, _P, _P, _P, _P]),
    "rc_debug_mip": (_i32, [_P, _u32, _P, _P, _P]),
    "rc_batch_create": (_i32, [ctypes.POINTER(View), _u32, ctypes.POINTER(BatchConfig), ctypes.POINTER(_P)]),
    "rc_batch_info": (_i32, [_P, ctypes.POINTER(_u32), ctypes.POINTER(ctypes.c_uint64)]),
    "rc_batch_autotune": (_i32, [_P, _u32, _P]),
    "rc_batch_run": (_i32, [_P, _P]),
    "rc_batch_view_levels": (_i32, [_P, _u32, ctypes.POINTER(_u32)]),
    "rc_batch_last_error": (ctypes.c_char_p, [_P]),
    "rc_batch_destroy": (None, [_P]),
    "rc_composite": (_i32, [_P, _P, _P, _P, _P, _P, _P]),
    "rc_dump_write": (_i32, [ctypes.c_char_p, ctypes.POINTER(Plane), _u32]),
    "rc_dump_info": (_i32, [ctypes.c_char_p, ctypes.POINTER(Plane), _u32, ctypes.POINTER(_u32)]),
    "rc_dump_read": (_i32, [ctypes.c_char_p, ctypes.c_char_p, _P, _u64]),
    "rc_dump_last_error": (ctypes.c_char_p, []),
    "rc_sha256": (None, [_P, _u64, _P]),
    "rc_srgb_thresholds": (None, [_P]),
    "rc_scene_generate": (_i32, [_u64, _u32, _u32, _P]),
    "rc_gbuffer_raster": (_i32, [_P, _u32, _u32, _u32, _P, _P, _P, _P]),
}

_lib = None


class RcError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


def load() -> ctypes.CDLL:
    """Load librcgi.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m radiancecascade_amd.build` "
                          "(the GI pass has no CPU fallback)")
    try:
        import torch  # noqa: F401  (one HIP runtime per process; see module docstring)
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.rc_abi_version() != ABI_VERSION:
        raise ImportError("librcgi.so ABI version mismatch")
    _lib = lib
    return lib


def check(status: int, ctx=None) -> None:
    if status != OK:
        lib = load()
        msg = lib.rc_last_error(ctx)
        raise RcError(status, msg.decode() if msg else "")
