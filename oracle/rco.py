"""ctypes wrapper around the CPU oracle (oracle/librco.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker.  The product
package (radiancecascade_amd) never imports it.

The functions restate the radiance-cascade maths of the published method
(the reference repository has no radiance-cascade code; its only pass is the
forward raster at src/renderer.rs:559-613).  Each wrapper names the C function
whose header comment states the rule it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librco.so")
_lib = None

SEED_NONE = 0xFFFFFFFF
RAY_EXIT_SCREEN = -1
RAY_REACHED_T1 = -2
MAX_BOXES = 256


class Box(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("albedo", ctypes.c_double * 3), ("emit", ctypes.c_double * 3)]


def build() -> str:
    """Compile oracle/librco.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, u64 = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
        sig = {
            "rco_splitmix64_next": (u64, [ctypes.POINTER(u64)]),
            "rco_u01": (ctypes.c_double, [ctypes.POINTER(u64)]),
            "rco_scene_generate": (i32, [u64, i32, i32, P]),
            "rco_raster": (None, [P, i32, i32, i32, P, P, P]),
            "rco_probe_spacing": (i32, [i32]),
            "rco_num_dirs": (i32, [i32]),
            "rco_probe_grid": (None, [i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
            "rco_interval_t0": (i64, [i32]),
            "rco_interval_t1": (i64, [i32]),
            "rco_level_rays": (i64, [i32, i32, i32]),
            "rco_levels_for_coverage": (i32, [i32, i32]),
            "rco_ray_index": (i64, [i32, i32, i32, i32, i32, i32]),
            "rco_dir_table_f32": (None, [i32, P]),
            "rco_jfa_num_passes": (i32, [i32, i32]),
            "rco_seed_init": (None, [i32, i32, P, P, P]),
            "rco_jfa_pass": (None, [i32, i32, i32, P, P]),
            "rco_jfa": (None, [i32, i32, P, P]),
            "rco_sdf_value_d2": (ctypes.c_double, [u64]),
            "rco_sdf_f32": (None, [i32, i32, P, P]),
            "rco_sdf_f64": (None, [i32, i32, P, P]),
            "rco_nearest_seed_bruteforce": (None, [i32, i32, P, P]),
            "rco_edt_d2": (None, [i32, i32, P, P]),
            "rco_sdf_from_d2_f32": (None, [i32, i32, P, P]),
            "rco_sdf_from_d2_f64": (None, [i32, i32, P, P]),
            "rco_bruteforce_irradiance": (None, [i32, i32, P, P, i32, ctypes.c_double, P, i32, P]),
            "rco_march_level_f64": (None, [i32, i32, i32, P, P, P, P]),
            "rco_march_level_f32": (None, [i32, i32, i32, P, P, P, P]),
            "rco_merge_level_f64": (None, [i32, i32, i32, P, P, P]),
            "rco_merge_level_f32": (None, [i32, i32, i32, P, P, P]),
            "rco_gather_f64": (None, [i32, i32, P, P]),
            "rco_gather_f32": (None, [i32, i32, P, P]),
            "rco_run_f64": (i32, [i32, i32, i32, P, P, P, P, P, i32]),
            "rco_run_f32": (i32, [i32, i32, i32, P, P, P, P, P, i32]),
            "rco_set_threads": (None, [i32]),
            "rco_get_threads": (i32, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- PRNG / scene / G-buffer ------------------------------------------------
def splitmix64_stream(seed: int, n: int) -> list[int]:
    st = ctypes.c_uint64(seed)
    return [lib().rco_splitmix64_next(ctypes.byref(st)) for _ in range(n)]


def scene(seed: int, n_boxes: int, n_emissive: int) -> np.ndarray:
    """Box list as a structured (n, 12) float64 array: lo3, hi3, albedo3, emit3."""
    boxes = (Box * max(n_boxes, 1))()
    r = lib().rco_scene_generate(seed, n_boxes, n_emissive, ctypes.cast(boxes, ctypes.c_void_p))
    if r < 0:
        raise ValueError("invalid scene parameters")
    arr = np.ctypeslib.as_array(ctypes.cast(boxes, ctypes.POINTER(ctypes.c_double)), shape=(max(n_boxes, 1) * 12,))
    return arr.reshape(-1, 12)[:n_boxes].copy()


def raster(boxes: np.ndarray, W: int, H: int):
    boxes = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 12)
    depth = np.empty((H, W), np.float32)
    albedo = np.empty((3, H, W), np.float32)
    emissive = np.empty((3, H, W), np.float32)
    buf = boxes if boxes.size else np.zeros((1, 12))
    lib().rco_raster(_p(buf), boxes.shape[0], W, H, _p(depth), _p(albedo), _p(emissive))
    return depth, albedo, emissive


def gbuffer(seed: int, W: int, H: int, n_boxes: int, n_emissive: int):
    return raster(scene(seed, n_boxes, n_emissive), W, H)


# ---- geometry ---------------------------------------------------------------
def probe_grid(W: int, H: int, level: int) -> tuple[int, int]:
    px, py = ctypes.c_int(), ctypes.c_int()
    lib().rco_probe_grid(W, H, level, ctypes.byref(px), ctypes.byref(py))
    return px.value, py.value


def num_dirs(level: int) -> int:
    return lib().rco_num_dirs(level)


def spacing(level: int) -> int:
    return lib().rco_probe_spacing(level)


def interval(level: int) -> tuple[int, int]:
    return lib().rco_interval_t0(level), lib().rco_interval_t1(level)


def level_rays(W: int, H: int, level: int) -> int:
    return lib().rco_level_rays(W, H, level)


def levels_for_coverage(W: int, H: int) -> int:
    return lib().rco_levels_for_coverage(W, H)


def ray_index(W, H, level, px, py, k) -> int:
    return lib().rco_ray_index(W, H, level, px, py, k)


def dir_table_f32(level: int) -> np.ndarray:
    d = np.empty((num_dirs(level), 2), np.float32)
    lib().rco_dir_table_f32(level, _p(d))
    return d


# ---- depth convention ---------------------------------------------------------
def ndc_from_linear(depth: np.ndarray, z_near: float, z_far: float) -> np.ndarray:
    """The reference renderer's Depth32Float value for linear view depth D:
    NDC depth of glam's perspective_rh (src/camera.rs:77-79; clip z = r (n - D)
    with r = f / (n - f), clip w = D), d = f (D - n) / ((f - n) D), evaluated in
    FP64 and rounded to fp32 as the depth target stores it."""
    D = np.asarray(depth, np.float64)
    n, f = float(z_near), float(z_far)
    return (f * (D - n) / ((f - n) * D)).astype(np.float32)


def linear_from_ndc(d: np.ndarray, z_near: float, z_far: float) -> np.ndarray:
    """Linear view depth as the pass reads an NDC depth plane (RC_DEPTH_NDC):
    D = (n f) / (f - d (f - n)) in fp32, with n f and f - n each rounded to
    fp32 first and every operation rounded once."""
    d = np.asarray(d, np.float32)
    n, f = np.float32(z_near), np.float32(z_far)
    nf = np.float32(n * f)
    fmn = np.float32(f - n)
    return (nf / (f - d * fmn)).astype(np.float32)


# ---- occupancy / jump flood -------------------------------------------------
def jfa_num_passes(W: int, H: int) -> int:
    return lib().rco_jfa_num_passes(W, H)


def seed_init(depth: np.ndarray, emissive: np.ndarray) -> np.ndarray:
    H, W = depth.shape
    seed = np.empty((H, W), np.uint32)
    lib().rco_seed_init(W, H, _p(np.ascontiguousarray(depth, np.float32)),
                        _p(np.ascontiguousarray(emissive, np.float32)), _p(seed))
    return seed


def jfa_pass(seed: np.ndarray, step: int) -> np.ndarray:
    H, W = seed.shape
    out = np.empty_like(seed)
    lib().rco_jfa_pass(W, H, step, _p(np.ascontiguousarray(seed, np.uint32)), _p(out))
    return out


def jfa(seed_init_field: np.ndarray) -> np.ndarray:
    s = np.array(seed_init_field, dtype=np.uint32, copy=True, order="C")
    tmp = np.empty_like(s)
    H, W = s.shape
    lib().rco_jfa(W, H, _p(s), _p(tmp))
    return s


def nearest_seed_bruteforce(seed_init_field: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(seed_init_field, np.uint32)
    H, W = s.shape
    out = np.empty_like(s)
    lib().rco_nearest_seed_bruteforce(W, H, _p(s), _p(out))
    return out


SDF_FAR = 2160.0


def sdf_value_d2(d2: int) -> float:
    """Distance value the march reads for squared distance d2 (conservative code)."""
    return lib().rco_sdf_value_d2(d2)


def sdf_f32(seeds: np.ndarray) -> np.ndarray:
    H, W = seeds.shape
    out = np.empty((H, W), np.float32)
    lib().rco_sdf_f32(W, H, _p(np.ascontiguousarray(seeds, np.uint32)), _p(out))
    return out


def sdf_f64(seeds: np.ndarray) -> np.ndarray:
    H, W = seeds.shape
    out = np.empty((H, W), np.float64)
    lib().rco_sdf_f64(W, H, _p(np.ascontiguousarray(seeds, np.uint32)), _p(out))
    return out


def edt_d2_exact(seed_init_field: np.ndarray) -> np.ndarray:
    """int64 [H][W]: exact squared distance to the nearest occupied pixel
    (-1 where no pixel is occupied), from scipy's exact Euclidean distance
    transform (a library; its nearest-feature indices give d^2 in integers)."""
    from scipy import ndimage
    occ = np.asarray(seed_init_field) != SEED_NONE
    H, W = occ.shape
    if not occ.any():
        return np.full((H, W), -1, np.int64)
    _, (iy, ix) = ndimage.distance_transform_edt(~occ, return_indices=True)
    yy, xx = np.mgrid[0:H, 0:W]
    return ((iy - yy).astype(np.int64) ** 2 + (ix - xx).astype(np.int64) ** 2)


SDF_NONE_EXACT = 1e30


def edt_d2(seed_init_field: np.ndarray) -> np.ndarray:
    """int64 [H][W]: exact squared distance to the nearest occupied pixel, -1
    where no pixel is occupied (rco_edt_d2: Felzenszwalb & Huttenlocher's
    separable transform in exact integer arithmetic)."""
    s = np.ascontiguousarray(seed_init_field, np.uint32)
    H, W = s.shape
    out = np.empty((H, W), np.int64)
    lib().rco_edt_d2(W, H, _p(s), _p(out))
    return out


def sdf_from_d2(d2: np.ndarray, dtype=np.float32) -> np.ndarray:
    """The distance values the march reads, from exact squared distances."""
    d2 = np.ascontiguousarray(d2, np.int64)
    H, W = d2.shape
    out = np.empty((H, W), dtype)
    fn = lib().rco_sdf_from_d2_f32 if np.dtype(dtype) == np.float32 else lib().rco_sdf_from_d2_f64
    fn(W, H, _p(d2), _p(out))
    return out


def sdf_exact_f64(seed_init_field: np.ndarray) -> np.ndarray:
    """float64 [H][W]: the exact field sqrt(d^2) of the nearest occupied pixel
    (1e30 with none) -- the field the one-byte code approximates from below."""
    d2 = edt_d2_exact(seed_init_field)
    return np.where(d2 >= 0, np.sqrt(np.maximum(d2, 0).astype(np.float64)), SDF_NONE_EXACT)


def bruteforce_irradiance(seed_init_field: np.ndarray, emissive: np.ndarray, L: int, pixels=None) -> np.ndarray:
    """Ground truth for the whole path (rco_bruteforce_irradiance): per pixel
    centre, the mean over N_{L-1} directions of the emission of the first
    occupied pixel each ray crosses within t1_{L-1} px.  Returns [3][H][W], or
    [3][n] for a list of pixel indices y*W+x."""
    s = np.ascontiguousarray(seed_init_field, np.uint32)
    H, W = s.shape
    n_dirs, t_max = num_dirs(L - 1), float(interval(L - 1)[1])
    if pixels is None:
        out = np.empty((H * W, 3), np.float64)
        lib().rco_bruteforce_irradiance(W, H, _p(s), _p(np.ascontiguousarray(emissive, np.float32)), n_dirs, t_max,
                                        None, H * W, _p(out))
        return out.T.reshape(3, H, W).copy()
    px = np.ascontiguousarray(pixels, np.int32)
    out = np.empty((px.size, 3), np.float64)
    lib().rco_bruteforce_irradiance(W, H, _p(s), _p(np.ascontiguousarray(emissive, np.float32)), n_dirs, t_max,
                                    _p(px), px.size, _p(out))
    return out.T.copy()


# ---- march / merge / gather -------------------------------------------------
def _flavour(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def march_level(sdf: np.ndarray, emissive: np.ndarray, level: int):
    """(rgba [Py,Px,N,4], hit [Py,Px,N] int32) in sdf's precision."""
    H, W = sdf.shape
    PX, PY = probe_grid(W, H, level)
    N = num_dirs(level)
    rgba = np.empty((PY, PX, N, 4), sdf.dtype)
    hit = np.empty((PY, PX, N), np.int32)
    fn = getattr(lib(), "rco_march_level_" + _flavour(sdf.dtype))
    fn(W, H, level, _p(np.ascontiguousarray(sdf)), _p(np.ascontiguousarray(emissive, np.float32)), _p(rgba), _p(hit))
    return rgba, hit


def merge_level(W: int, H: int, level: int, march_rgba: np.ndarray, upper: np.ndarray | None):
    PX, PY = probe_grid(W, H, level)
    NO = num_dirs(level) // 4
    out = np.empty((PY, PX, NO, 4), march_rgba.dtype)
    fn = getattr(lib(), "rco_merge_level_" + _flavour(march_rgba.dtype))
    up = None if upper is None else np.ascontiguousarray(upper, march_rgba.dtype)
    fn(W, H, level, _p(np.ascontiguousarray(march_rgba)), None if up is None else _p(up), _p(out))
    return out


def gather(W: int, H: int, A0: np.ndarray) -> np.ndarray:
    irr = np.empty((3, H, W), A0.dtype)
    fn = getattr(lib(), "rco_gather_" + _flavour(A0.dtype))
    fn(W, H, _p(np.ascontiguousarray(A0)), _p(irr))
    return irr


SDF_MODES = {"exact": 0, "jfa": 1}


def run(depth: np.ndarray, emissive: np.ndarray, L: int, dtype=np.float64, want_levels=False, want_seeds=False,
        sdf: str = "exact"):
    """Whole path.  Returns irr [3,H,W] (and levels list A_0..A_{L-1}, seeds).
    sdf: "exact" (the distance field of the exact transform) or "jfa" (of the
    jump flood; the only mode with a seed field, want_seeds)."""
    H, W = depth.shape
    dtype = np.dtype(dtype)
    irr = np.empty((3, H, W), dtype)
    levels = None
    lv_ptrs = None
    if want_levels:
        levels = []
        for i in range(L):
            PX, PY = probe_grid(W, H, i)
            levels.append(np.empty((PY, PX, num_dirs(i) // 4, 4), dtype))
        lv_ptrs = (ctypes.c_void_p * L)(*[a.ctypes.data for a in levels])
    seeds = np.empty((H, W), np.uint32) if want_seeds else None
    fn = getattr(lib(), "rco_run_" + _flavour(dtype))
    if want_seeds and sdf != "jfa":
        raise ValueError("want_seeds needs sdf='jfa'")
    r = fn(W, H, L, _p(np.ascontiguousarray(depth, np.float32)), _p(np.ascontiguousarray(emissive, np.float32)),
           _p(irr), lv_ptrs, None if seeds is None else _p(seeds), SDF_MODES[sdf])
    if r != 0:
        raise ValueError("invalid run parameters")
    out = [irr]
    if want_levels:
        out.append(levels)
    if want_seeds:
        out.append(seeds)
    return out[0] if len(out) == 1 else tuple(out)


def run_on_field(sdf: np.ndarray, emissive: np.ndarray, L: int) -> np.ndarray:
    """March, merge and gather on a given distance field (in sdf's precision):
    the composition of the stage functions, top level first.  With the field
    the jump flood produces this is rco_run(sdf="jfa"); with sdf_from_d2 of the
    exact transform it is rco_run(sdf="exact"); with sdf_exact_f64 it is the
    real-valued exact-field flavour."""
    H, W = sdf.shape
    upper = None
    for i in range(L - 1, -1, -1):
        m, _ = march_level(sdf, emissive, i)
        upper = merge_level(W, H, i, m, upper)
    return gather(W, H, upper)


def set_threads(n: int) -> None:
    lib().rco_set_threads(n)


def get_threads() -> int:
    return lib().rco_get_threads()
