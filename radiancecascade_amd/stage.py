"""Host-side mirror of the reference's pass interface for the GI pass.

The reference renderer drives every pass through the Rust trait
``RenderStage<T>`` (src/app.rs:3-7): ``update(&mut self, state, queue)`` runs
first each frame and is the only method that receives the queue, so uploads
happen there (src/renderer.rs:620-631); ``render(&self, state, view,
encoder)`` records the pass (src/renderer.rs:559-613); ``resize(&mut self,
device, config)`` recreates size-dependent resources (src/renderer.rs:615-618).
:class:`RadianceCascadeStage` keeps those three names and that order over the
C ABI of librcgi.so, plus ``run_device`` for callers whose G-buffer is already
in device memory.  Where the reference ``unwrap``s and panics
(src/renderer.rs:176, src/window/app.rs:43,57,64), this raises
:class:`RcError`.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import RcError, check


def _dev_ptr(x, nbytes: int | None = None, name: str = "buffer") -> int:
    """Device address of a torch tensor on the GPU, or an int address."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    try:
        import torch
    except ImportError as e:  # pragma: no cover
        raise TypeError(f"{name}: pass a device address (int)") from e
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor or int address")
    if not x.is_cuda:
        raise ValueError(f"{name}: tensor must be on the GPU")
    if not x.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if nbytes is not None and x.numel() * x.element_size() < nbytes:
        raise ValueError(f"{name}: tensor holds {x.numel() * x.element_size()} bytes, needs {nbytes}")
    return x.data_ptr()


def _stream_handle(stream) -> int:
    if stream is None:
        try:
            import torch
            return torch.cuda.current_stream().cuda_stream
        except Exception:
            return 0
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


@dataclass
class StageStats:
    runs: int
    stage_ms: dict
    stage_launches: dict
    level_rays: list
    level_bytes_stored: list
    surrogate_frames: int = 0
    fallback_frames: int = 0
    flagged_tiles: int = 0
    tiles_per_frame: int = 0


class RadianceCascadeStage:
    """The GI pass: jump flood -> per-level march/merge -> gather."""

    def __init__(self, width: int, height: int, levels: int = 0, storage: str = "f32",
                 device: int = 0, profile: bool = False, debug: bool = False, graph: bool = False,
                 ndc_depth: tuple[float, float] | None = None):
        """ndc_depth=(z_near, z_far): the depth plane holds the reference
        renderer's Depth32Float NDC depth of perspective_rh (src/camera.rs:77-79)
        instead of linear view depth; the pass converts it (RC_DEPTH_NDC)."""
        lib = _abi.load()
        if storage not in _abi.STORAGE:
            raise ValueError(f"storage must be one of {sorted(_abi.STORAGE)}")
        near, far = ndc_depth if ndc_depth is not None else (0.0, 0.0)
        cfg = _abi.Config(version=_abi.ABI_VERSION, width=width, height=height, levels=levels,
                          storage=_abi.STORAGE[storage],
                          flags=(_abi.FLAG_PROFILE if profile else 0) | (_abi.FLAG_KEEP_DEBUG if debug else 0)
                          | (_abi.FLAG_GRAPH if graph else 0),
                          device=device, depth_mode=_abi.DEPTH_NDC if ndc_depth is not None else _abi.DEPTH_LINEAR,
                          z_near=near, z_far=far)
        ctx = ctypes.c_void_p()
        check(lib.rc_create(ctypes.byref(cfg), ctypes.byref(ctx)))
        self._lib = lib
        self._ctx = ctx
        self._sync_config()

    # -- RenderStage::resize (src/app.rs:5) ----------------------------------
    def resize(self, width: int, height: int, levels: int = 0) -> None:
        check(self._lib.rc_resize(self._ctx, width, height, levels), self._ctx)
        self._sync_config()

    # -- RenderStage::update (src/app.rs:6): host G-buffer upload ------------
    def update(self, depth: np.ndarray, emissive: np.ndarray, albedo: np.ndarray | None = None) -> None:
        H, W = self.height, self.width
        depth = np.ascontiguousarray(depth, np.float32)
        emissive = np.ascontiguousarray(emissive, np.float32)
        if depth.shape != (H, W) or emissive.shape != (3, H, W):
            raise ValueError(f"G-buffer must be depth [{H},{W}] and emissive [3,{H},{W}]")
        g = _abi.GBuffer(depth=depth.ctypes.data, albedo=None, emissive=emissive.ctypes.data)
        check(self._lib.rc_update_host(self._ctx, ctypes.byref(g)), self._ctx)

    # -- RenderStage::render (src/app.rs:4): run the pass --------------------
    def render(self, out: np.ndarray | None = None) -> np.ndarray:
        """Irradiance [3][H][W] into `out` (e.g. a host_buffer) or a new array."""
        shape = (3, self.height, self.width)
        if out is None:
            out = np.empty(shape, np.float32)
        elif out.shape != shape or out.dtype != np.float32 or not out.flags["C_CONTIGUOUS"]:
            raise ValueError(f"out must be a contiguous float32 array of shape {shape}")
        check(self._lib.rc_render_host(self._ctx, out.ctypes.data), self._ctx)
        return out

    # -- device-resident path, ordered on `stream` ---------------------------
    def run_device(self, depth, emissive, irradiance, stream=None) -> None:
        n = self.width * self.height * 4
        g = _abi.GBuffer(depth=_dev_ptr(depth, n, "depth"), albedo=None,
                         emissive=_dev_ptr(emissive, 3 * n, "emissive"))
        check(self._lib.rc_run_device(self._ctx, ctypes.byref(g), _dev_ptr(irradiance, 3 * n, "irradiance"),
                                      _stream_handle(stream)), self._ctx)

    # -- level-kernel tiling ---------------------------------------------------
    def set_tiling(self, level: int, shape: int, rays_per_lane: int) -> None:
        check(self._lib.rc_set_level_tiling(self._ctx, level, shape, rays_per_lane), self._ctx)

    def set_march(self, split: int, speculation: int = 1, rays_per_lane: int = 1) -> None:
        """split: 0 in place, 1 one launch for every level >= 1, 2 the same by
        a persistent grid (rc_set_march)."""
        check(self._lib.rc_set_march(self._ctx, int(split), speculation, rays_per_lane), self._ctx)

    def march(self) -> tuple[int, int, int]:
        sp, k, m = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(self._lib.rc_get_march(self._ctx, ctypes.byref(sp), ctypes.byref(k), ctypes.byref(m)), self._ctx)
        return sp.value, k.value, m.value

    def set_march_order(self, level: int, order: int) -> None:
        """Order of level `level`'s rays over the split march's lanes: 0
        canonical, 1..5 wave tiles of probes x directions (rc_set_march_order)."""
        check(self._lib.rc_set_march_order(self._ctx, level, order), self._ctx)

    def march_order(self) -> dict:
        """{level: order} for levels >= 1."""
        out = {}
        for i in range(1, self.levels):
            v = ctypes.c_uint32()
            check(self._lib.rc_get_march_order(self._ctx, i, ctypes.byref(v)), self._ctx)
            out[i] = int(v.value)
        return out

    def set_sdf_mode(self, mode: str) -> None:
        """Distance transform: "exact" (the default; occupancy, band-end and
        row kernels) or "jfa" (the jump flood, whose seed field debug_seeds
        reads)."""
        check(self._lib.rc_set_sdf_mode(self._ctx, _abi.SDF_MODES[mode]), self._ctx)

    def sdf_mode(self) -> str:
        v = self._lib.rc_get_sdf_mode(self._ctx)
        return {k: n for n, k in _abi.SDF_MODES.items()}[v]

    def set_jfa_fuse(self, fused_step: int) -> None:
        check(self._lib.rc_set_jfa_fuse(self._ctx, fused_step), self._ctx)

    def jfa_fuse(self) -> int:
        return self._lib.rc_get_jfa_fuse(self._ctx)

    def set_jfa_rows(self, rows: int) -> None:
        """Image rows per wave of a global jump-flood pass (1, 2 or 4)."""
        check(self._lib.rc_set_jfa_rows(self._ctx, rows), self._ctx)

    def jfa_rows(self) -> int:
        return self._lib.rc_get_jfa_rows(self._ctx)

    def tiling(self) -> dict:
        """{level: ((TX, TY, KD), rays_per_lane)} for levels >= 1."""
        out = {}
        for i in range(1, self.levels):
            sh, r = ctypes.c_uint32(), ctypes.c_uint32()
            check(self._lib.rc_get_level_tiling(self._ctx, i, ctypes.byref(sh), ctypes.byref(r)), self._ctx)
            out[i] = (_abi.TILE_SHAPES[sh.value], r.value)
        return out

    def autotune(self, depth, emissive, irradiance, iters: int = 5, stream=None) -> dict:
        n = self.width * self.height * 4
        g = _abi.GBuffer(depth=_dev_ptr(depth, n, "depth"), albedo=None,
                         emissive=_dev_ptr(emissive, 3 * n, "emissive"))
        check(self._lib.rc_autotune(self._ctx, ctypes.byref(g), _dev_ptr(irradiance, 3 * n, "irradiance"),
                                    iters, _stream_handle(stream)), self._ctx)
        return self.tiling()

    # -- introspection -------------------------------------------------------
    def _sync_config(self) -> None:
        cfg = _abi.Config()
        check(self._lib.rc_get_config(self._ctx, ctypes.byref(cfg)), self._ctx)
        self.width, self.height, self.levels = cfg.width, cfg.height, cfg.levels
        self.storage = {0: "f32", 1: "f16", 2: "bf16"}[cfg.storage]

    def stats(self) -> StageStats:
        s = _abi.Stats()
        check(self._lib.rc_get_stats(self._ctx, ctypes.byref(s)), self._ctx)
        names = {_abi.STAGE_SEED: "seed", _abi.STAGE_JFA: "jfa", _abi.STAGE_JFA_TAIL: "jfa_tail",
                 _abi.STAGE_MARCH: "march", _abi.STAGE_NEAR_SDF: "near_sdf", _abi.STAGE_MIP: "mip",
                 _abi.STAGE_CONE: "cone", _abi.STAGE_SURROGATE: "surrogate", _abi.STAGE_FALLBACK: "fallback"}
        for i in range(_abi.MAX_LEVELS):
            names[_abi.STAGE_LEVEL0 + i] = f"level{i}"
        ms = {names[i]: s.stage_ms[i] for i in range(_abi.NUM_STAGES) if s.stage_launches[i]}
        la = {names[i]: s.stage_launches[i] for i in range(_abi.NUM_STAGES) if s.stage_launches[i]}
        return StageStats(runs=s.runs, stage_ms=ms, stage_launches=la,
                          level_rays=list(s.level_rays)[: self.levels],
                          level_bytes_stored=list(s.level_bytes_stored)[: self.levels],
                          surrogate_frames=s.surrogate_frames, fallback_frames=s.fallback_frames,
                          flagged_tiles=s.flagged_tiles, tiles_per_frame=s.tiles_per_frame)

    def reset_stats(self) -> None:
        check(self._lib.rc_reset_stats(self._ctx), self._ctx)

    def level_shape(self, level: int) -> tuple[int, int, int, int]:
        px, py = probe_grid(self.width, self.height, level)
        return (py, px, max((4 << (2 * level)) // 4, 1), 4)

    def debug_seeds(self, out, stream=None) -> None:
        check(self._lib.rc_debug_seeds(self._ctx, _dev_ptr(out, self.width * self.height * 4, "out"),
                                       _stream_handle(stream)), self._ctx)

    def debug_sdf(self, out, stream=None) -> None:
        check(self._lib.rc_debug_sdf(self._ctx, _dev_ptr(out, self.width * self.height * 4, "out"),
                                     _stream_handle(stream)), self._ctx)

    def debug_level(self, level: int, out, stream=None) -> None:
        n = int(np.prod(self.level_shape(level))) * 4
        check(self._lib.rc_debug_level(self._ctx, level, _dev_ptr(out, n, "out"), _stream_handle(stream)),
              self._ctx)

    def debug_march(self, level: int, hit, index=None, stream=None) -> None:
        """The ray codes the last frame's kernels stored for `level` (int32
        device tensor, canonical order); `index` (int64, optional, level >= 1)
        receives the split march's decode of every slot (rc_debug_march)."""
        rays = level_rays(self.width, self.height, level)
        check(self._lib.rc_debug_march(self._ctx, level, _dev_ptr(hit, rays * 4, "hit"),
                                       None if index is None else _dev_ptr(index, rays * 8, "index"),
                                       _stream_handle(stream)), self._ctx)

    def composite(self, albedo, emissive, irradiance, out, direct=None, stream=None) -> None:
        """The presented frame: out [H][W] uint32 (int32 tensor) of bytes B, G, R,
        A=255 from linear (albedo * irradiance + emissive) + direct, 8-bit sRGB
        (rc_composite).  All planar [3][H][W] fp32 device tensors; direct optional."""
        n = self.width * self.height
        check(self._lib.rc_composite(self._ctx, _dev_ptr(albedo, 3 * n * 4, "albedo"),
                                     _dev_ptr(emissive, 3 * n * 4, "emissive"),
                                     _dev_ptr(irradiance, 3 * n * 4, "irradiance"),
                                     None if direct is None else _dev_ptr(direct, 3 * n * 4, "direct"),
                                     _dev_ptr(out, n * 4, "out"), _stream_handle(stream)), self._ctx)

    def error_norms(self, x, ref, stream=None) -> dict:
        """rel-L2 and max-abs of device fp32 `x` against device `ref` (fp32 or
        fp64 tensor of the same length), reduced in FP64 on the GPU."""
        import torch
        n = x.numel()
        if ref.numel() != n:
            raise ValueError("x and ref differ in length")
        out = torch.empty(3, dtype=torch.float64, device=x.device)
        check(self._lib.rc_error_norms(self._ctx, _dev_ptr(x, n * 4, "x"),
                                       _dev_ptr(ref, n * ref.element_size(), "ref"),
                                       int(ref.dtype == torch.float64), n, _dev_ptr(out, 24, "out"),
                                       _stream_handle(stream)), self._ctx)
        sd, sr, mx = out.cpu().tolist()
        return {"rel_l2": (sd / sr) ** 0.5 if sr > 0 else sd ** 0.5, "max_abs": mx, "sum_sq_diff": sd, "sum_sq_ref": sr}

    # ---- learned far field ----------------------------------------------------
    def set_surrogate(self, enabled: bool, threshold: float = float("inf")) -> None:
        """Replace levels >= 2 by the cone features + MLP (enabled) or run the
        plain cascades.  Tiles whose mean predicted squared error (normalised
        units) exceeds `threshold` fall back to the plain cascades."""
        c = _abi.SurrogateConfig(_abi.ABI_VERSION, int(bool(enabled)), float(threshold), 0)
        check(self._lib.rc_set_surrogate(self._ctx, ctypes.byref(c)), self._ctx)

    def surrogate(self) -> tuple[bool, float]:
        c = _abi.SurrogateConfig()
        check(self._lib.rc_get_surrogate(self._ctx, ctypes.byref(c)), self._ctx)
        return bool(c.mode), float(c.threshold)

    def mlp_set(self, params: np.ndarray, in_mean: np.ndarray, in_istd: np.ndarray, out_scale: float) -> None:
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        m = np.ascontiguousarray(in_mean, np.float32).reshape(-1)
        d = np.ascontiguousarray(in_istd, np.float32).reshape(-1)
        if p.size != _abi.MLP_PARAMS or m.size != _abi.MLP_INPUTS or d.size != _abi.MLP_INPUTS:
            raise ValueError("parameter arrays have the wrong size")
        check(self._lib.rc_mlp_set(self._ctx, p.ctypes.data, m.ctypes.data, d.ctypes.data, float(out_scale)), self._ctx)

    def mlp_get(self) -> dict:
        p = np.empty(_abi.MLP_PARAMS, np.float32)
        m = np.empty(_abi.MLP_INPUTS, np.float32)
        d = np.empty(_abi.MLP_INPUTS, np.float32)
        o = ctypes.c_float()
        check(self._lib.rc_mlp_get(self._ctx, p.ctypes.data, m.ctypes.data, d.ctypes.data, ctypes.byref(o)), self._ctx)
        return {"params": p, "in_mean": m, "in_istd": d, "out_scale": float(o.value)}

    def weights_blob(self) -> bytes:
        n = ctypes.c_uint64()
        check(self._lib.rc_surrogate_save(self._ctx, None, 0, ctypes.byref(n)), self._ctx)
        buf = ctypes.create_string_buffer(n.value)
        check(self._lib.rc_surrogate_save(self._ctx, buf, n.value, ctypes.byref(n)), self._ctx)
        return buf.raw[: n.value]

    def load_weights_blob(self, blob: bytes) -> None:
        buf = ctypes.create_string_buffer(blob, len(blob))
        check(self._lib.rc_surrogate_load(self._ctx, buf, len(blob)), self._ctx)

    def save_weights(self, path: str, meta: dict | None = None) -> str:
        """Write the weights blob and a `<path>.json` sidecar with its SHA-256."""
        import hashlib
        import json
        blob = self.weights_blob()
        with open(path, "wb") as f:
            f.write(blob)
        side = {"file": os.path.basename(path), "bytes": len(blob), "sha256": hashlib.sha256(blob).hexdigest(),
                "format": _abi.SURROGATE_BLOB_MAGIC.decode()}
        if meta:
            side.update(meta)
        with open(path + ".json", "w") as f:
            json.dump(side, f, indent=1)
        return side["sha256"]

    def load_weights(self, path: str | None = None, verify: bool = True) -> None:
        """Load a weights blob (default: the packaged network), checking the
        SHA-256 recorded in its `.json` sidecar."""
        import hashlib
        import json
        path = path or packaged_weights()
        with open(path, "rb") as f:
            blob = f.read()
        if verify:
            with open(path + ".json") as f:
                side = json.load(f)
            if hashlib.sha256(blob).hexdigest() != side["sha256"] or len(blob) != side["bytes"]:
                raise RcError(_abi.ERR_INVALID, f"{path}: SHA-256 does not match its sidecar")
        self.load_weights_blob(blob)

    def mlp_forward(self, feats, out, stream=None) -> None:
        """out [n][64] = raw network outputs for features [n][MLP_INPUTS] (device fp32)."""
        n = feats.shape[0]
        check(self._lib.rc_mlp_forward(self._ctx, _dev_ptr(feats, n * _abi.MLP_INPUTS * 4, "feats"), n,
                                       _dev_ptr(out, n * 64 * 4, "out"), _stream_handle(stream)), self._ctx)

    def mlp_gradient(self, feats, targets, grad, loss, stream=None) -> None:
        n = feats.shape[0]
        check(self._lib.rc_mlp_gradient(self._ctx, _dev_ptr(feats, n * _abi.MLP_INPUTS * 4, "feats"),
                                        _dev_ptr(targets, n * _abi.MLP_RADIANCE * 4, "targets"), n,
                                        _dev_ptr(grad, _abi.MLP_PARAMS * 4, "grad"), _dev_ptr(loss, 4, "loss"),
                                        _stream_handle(stream)), self._ctx)

    def mlp_adam(self, grad, lr: float, stream=None) -> None:
        check(self._lib.rc_mlp_adam(self._ctx, _dev_ptr(grad, _abi.MLP_PARAMS * 4, "grad"), float(lr),
                                    _stream_handle(stream)), self._ctx)

    def surrogate_samples(self, depth, emissive, feats, targets, stream=None) -> None:
        """Training rows of one device G-buffer: feats [Px_2*Py_2][MLP_INPUTS],
        targets [Px_2*Py_2][48]."""
        px, py = probe_grid(self.width, self.height, 2)
        n = px * py
        g = _abi.GBuffer(_dev_ptr(depth, self.width * self.height * 4, "depth"), None,
                         _dev_ptr(emissive, 3 * self.width * self.height * 4, "emissive"))
        check(self._lib.rc_surrogate_samples(self._ctx, ctypes.byref(g), _dev_ptr(feats, n * _abi.MLP_INPUTS * 4, "feats"),
                                             _dev_ptr(targets, n * _abi.MLP_RADIANCE * 4, "targets"),
                                             _stream_handle(stream)),
              self._ctx)

    def train_surrogate(self, width: int = 512, height: int = 512, levels: int = 6, scene_seed0: int = 10000,
                        scenes: int = 10000, n_boxes: int = 32, n_emissive: int = 4, epochs: int = 8,
                        batch: int = 16384, lr: float = 2e-3, lr_decay: float = 0.7, seed: int = 1) -> dict:
        """Train this context's network on the GPU (rc_surrogate_train)."""
        cfg = _abi.TrainConfig(_abi.ABI_VERSION, width, height, levels, scene_seed0, scenes, n_boxes, n_emissive,
                               epochs, batch, lr, lr_decay, 0, seed)
        rep = _abi.TrainReport()
        check(self._lib.rc_surrogate_train(self._ctx, ctypes.byref(cfg), ctypes.byref(rep)), self._ctx)
        return {"epochs": rep.epochs, "samples": rep.samples, "steps": rep.steps,
                "epoch_loss": list(rep.epoch_loss)[: rep.epochs], "gen_ms": rep.gen_ms, "train_ms": rep.train_ms,
                "out_scale": rep.out_scale}

    def train_surrogate_rows(self, feats, targets, epochs: int = 8, batch: int = 16384, lr: float = 2e-3,
                             lr_decay: float = 0.7, seed: int = 1, frames: int = 0) -> dict:
        """Train this context's network on rows the caller made
        (rc_surrogate_train_rows): feats [n][MLP_INPUTS], targets [n][48] as
        surrogate_samples writes them; targets is scaled in place."""
        n = feats.shape[0]
        if targets.shape[0] != n:
            raise ValueError("feats and targets need the same number of rows")
        cfg = _abi.TrainConfig(_abi.ABI_VERSION, 0, 0, 0, 0, frames, 0, 0, epochs, batch, lr, lr_decay, 0, seed)
        rep = _abi.TrainReport()
        check(self._lib.rc_surrogate_train_rows(self._ctx, _dev_ptr(feats, n * _abi.MLP_INPUTS * 4, "feats"),
                                                _dev_ptr(targets, n * _abi.MLP_RADIANCE * 4, "targets"), n,
                                                ctypes.byref(cfg), ctypes.byref(rep)), self._ctx)
        return {"epochs": rep.epochs, "samples": rep.samples, "steps": rep.steps,
                "epoch_loss": list(rep.epoch_loss)[: rep.epochs], "gen_ms": rep.gen_ms, "train_ms": rep.train_ms,
                "out_scale": rep.out_scale}

    def surrogate_shapes(self) -> dict:
        px, py = probe_grid(self.width, self.height, 1)
        qx, qy = probe_grid(self.width, self.height, 2)
        return {"level1": (py, px), "level2": (qy, qx), "tiles": ((qy + 7) // 8, (qx + 7) // 8)}

    def debug_surrogate(self, cone=None, level1=None, tile_err=None, stream=None) -> None:
        sh = self.surrogate_shapes()
        (qy, qx), (py, px), (ty, tx) = sh["level2"], sh["level1"], sh["tiles"]
        check(self._lib.rc_debug_surrogate(
            self._ctx, None if cone is None else _dev_ptr(cone, qx * qy * 64 * 4, "cone"),
            None if level1 is None else _dev_ptr(level1, px * py * 16 * 4, "level1"),
            None if tile_err is None else _dev_ptr(tile_err, tx * ty * 4, "tile_err"), _stream_handle(stream)),
            self._ctx)

    def debug_mip(self, level: int, emission=None, coverage=None, stream=None) -> None:
        """Mip level `level` of the last learned frame: emission [h][w][4] (mean
        E*[emitter] rgb, emitter area), coverage [h][w][4] (emitter column and
        row fractions, then the dark occluders')."""
        w = -(-self.width // (1 << level))
        h = -(-self.height // (1 << level))
        check(self._lib.rc_debug_mip(self._ctx, level,
                                     None if emission is None else _dev_ptr(emission, w * h * 16, "emission"),
                                     None if coverage is None else _dev_ptr(coverage, w * h * 16, "coverage"),
                                     _stream_handle(stream)), self._ctx)

    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._lib.rc_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RadianceCascadeBatch:
    """Many views of uneven sizes as one graph launch (rc_batch_*).

    views: sequence of (depth [H,W], emissive [3,H,W], irradiance [3,H,W])
    device tensors; levels: per-view cascade counts (0 = coverage rule).
    mode "worklist": one launch per stage over all views (shared workspace);
    "streams": one context per view, frames captured on `streams` streams."""

    def __init__(self, views, levels=None, storage: str = "f32", mode: str = "worklist", streams: int = 8,
                 device: int = 0):
        lib = _abi.load()
        n = len(views)
        arr = (_abi.View * n)()
        if n == 0:
            raise ValueError("a batch needs at least one view")
        for i, (d, e, o) in enumerate(views):
            if d is None or e is None or o is None:
                raise ValueError(f"view {i}: depth, emissive and irradiance buffers are required")
            H, W = d.shape[-2], d.shape[-1]
            arr[i] = _abi.View(W, H, 0 if levels is None else int(levels[i]), 0, _dev_ptr(d, W * H * 4, "depth"),
                               _dev_ptr(e, 3 * W * H * 4, "emissive"), _dev_ptr(o, 3 * W * H * 4, "irradiance"))
        h = ctypes.c_void_p()
        cfg = _abi.BatchConfig(_abi.STORAGE[storage], _abi.BATCH_MODES[mode], streams, device)
        st = lib.rc_batch_create(arr, n, ctypes.byref(cfg), ctypes.byref(h))
        if st != _abi.OK:
            raise RcError(st, (lib.rc_batch_last_error(None) or b"").decode())
        self._lib, self._h, self._views = lib, h, list(views)
        self.sizes = [(int(a.width), int(a.height)) for a in arr]

    def _check(self, st):
        if st != _abi.OK:
            raise RcError(st, (self._lib.rc_batch_last_error(self._h) or b"").decode())

    def autotune(self, iters: int = 2, stream=None) -> None:
        self._check(self._lib.rc_batch_autotune(self._h, iters, _stream_handle(stream)))

    def run(self, stream=None) -> None:
        self._check(self._lib.rc_batch_run(self._h, _stream_handle(stream)))

    def levels(self, i: int) -> int:
        v = ctypes.c_uint32()
        self._check(self._lib.rc_batch_view_levels(self._h, i, ctypes.byref(v)))
        return int(v.value)

    def info(self) -> dict:
        """Kernel nodes of one batch launch (streams mode: after the first
        run) and the bytes of the shared workspace (work-list mode)."""
        k, b = ctypes.c_uint32(), ctypes.c_uint64()
        self._check(self._lib.rc_batch_info(self._h, ctypes.byref(k), ctypes.byref(b)))
        return {"kernel_nodes": int(k.value), "workspace_bytes": int(b.value)}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.rc_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_SIZES = (256, 512, 768, 1024, 1536, 2048)
_M64 = (1 << 64) - 1


def batch_views_spec(seed0: int = 1000, n: int = 64, n_emissive: int = 4):
    """The multi-view workload: view j uses scene seed seed0 + j; its width,
    height (each from BATCH_SIZES) and box count (16..64) are the first three
    draws u of SplitMix64(seed ^ 0x5EED000000000000), index floor(6u),
    floor(6u), 16 + floor(49u).  Returns [(seed, W, H, n_boxes, n_emissive)]."""
    out = []
    for j in range(n):
        seed = seed0 + j
        state = (seed ^ 0x5EED000000000000) & _M64
        u = []
        for _ in range(3):
            state = (state + 0x9E3779B97F4A7C15) & _M64
            z = state
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            z ^= z >> 31
            u.append((z >> 11) * 2.0 ** -53)
        out.append((seed, BATCH_SIZES[int(6 * u[0])], BATCH_SIZES[int(6 * u[1])], 16 + int(49 * u[2]), n_emissive))
    return out


def packaged_weights() -> str:
    """Path of the network trained on the GPU and shipped with the package."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "surrogate.bin")


# ---- geometry and scenes ------------------------------------------------------
# ---- dump files (rc_dump_*) ----------------------------------------------------
def _dump_check(status: int) -> None:
    if status != 0:
        raise RcError(status, (_abi.load().rc_dump_last_error() or b"").decode())


def host_buffer(shape, dtype=np.float32) -> np.ndarray:
    """A numpy array in page-locked host memory (rc_host_alloc), freed with the
    array: G-buffer planes uploaded from it and irradiance rendered into it
    move by DMA."""
    lib = _abi.load()
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    ptr = ctypes.c_void_p()
    check(lib.rc_host_alloc(max(n, 1), ctypes.byref(ptr)))
    raw = (ctypes.c_ubyte * max(n, 1)).from_address(ptr.value)
    fin = weakref.finalize(raw, lib.rc_host_free, ptr.value)
    fin.atexit = False  # at exit the process's pages go with it; the runtime may be gone
    arr = np.frombuffer(raw, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
    return arr


def dump_write(path: str, planes: dict) -> None:
    """Write named arrays ([C][H][W] or [H][W], converted to fp32) to an
    RCDUMP01 file with a SHA-256 of its plane records and payload (rc_dump_write)."""
    lib = _abi.load()
    keep, recs = [], (_abi.Plane * max(1, len(planes)))()
    for i, (name, a) in enumerate(planes.items()):
        nb = name.encode()
        if len(nb) > 16:
            raise ValueError(f"plane name {name!r} longer than 16 bytes")
        a = np.ascontiguousarray(a, np.float32)
        if a.ndim == 2:
            a = a[None]
        if a.ndim != 3:
            raise ValueError(f"plane {name!r}: expected [C][H][W] or [H][W]")
        keep.append(a)
        recs[i].name, (recs[i].channels, recs[i].height, recs[i].width) = nb, a.shape
        recs[i].data = a.ctypes.data
    _dump_check(lib.rc_dump_write(os.fsencode(path), recs, len(planes)))


def dump_info(path: str) -> dict:
    """{name: (channels, height, width)} of an RCDUMP01 file (checksum verified)."""
    lib = _abi.load()
    n = ctypes.c_uint32()
    _dump_check(lib.rc_dump_info(os.fsencode(path), None, 0, ctypes.byref(n)))
    recs = (_abi.Plane * max(1, n.value))()
    _dump_check(lib.rc_dump_info(os.fsencode(path), recs, n.value, ctypes.byref(n)))
    return {recs[i].name.decode(): (recs[i].channels, recs[i].height, recs[i].width) for i in range(n.value)}


def dump_read(path: str, names=None) -> dict:
    """{name: float32 [C][H][W]} from an RCDUMP01 file (rc_dump_read)."""
    lib = _abi.load()
    out = {}
    for name, shape in dump_info(path).items():
        if names is not None and name not in names:
            continue
        a = np.empty(shape, np.float32)
        _dump_check(lib.rc_dump_read(os.fsencode(path), name.encode(), a.ctypes.data, a.size))
        out[name] = a
    return out


def sha256(data: bytes) -> str:
    """Hex SHA-256 computed by the library (rc_sha256)."""
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    out = ctypes.create_string_buffer(32)
    _abi.load().rc_sha256(buf, len(data), out)
    return out.raw.hex()


def srgb_thresholds() -> np.ndarray:
    """The composite's 256 sRGB encoding thresholds (rc_srgb_thresholds)."""
    t = np.empty(256, np.float32)
    _abi.load().rc_srgb_thresholds(t.ctypes.data)
    return t


def probe_grid(width: int, height: int, level: int) -> tuple[int, int]:
    lib = _abi.load()
    px, py = ctypes.c_uint32(), ctypes.c_uint32()
    check(lib.rc_probe_grid(width, height, level, ctypes.byref(px), ctypes.byref(py)))
    return px.value, py.value


def level_rays(width: int, height: int, level: int) -> int:
    return _abi.load().rc_level_rays(width, height, level)


def levels_for_coverage(width: int, height: int) -> int:
    return _abi.load().rc_levels_for_coverage(width, height)


def device_count() -> int:
    n = ctypes.c_int()
    st = _abi.load().rc_device_count(ctypes.byref(n))
    return n.value if st == _abi.OK else 0


def scene_generate(seed: int, n_boxes: int, n_emissive: int) -> np.ndarray:
    """Box list as an (n, 12) float64 array: lo xyz, hi xyz, albedo rgb, emit rgb."""
    boxes = (_abi.Box * max(n_boxes, 1))()
    check(_abi.load().rc_scene_generate(seed, n_boxes, n_emissive, ctypes.cast(boxes, ctypes.c_void_p)))
    arr = np.ctypeslib.as_array(ctypes.cast(boxes, ctypes.POINTER(ctypes.c_double)), shape=(max(n_boxes, 1) * 12,))
    return arr.reshape(-1, 12)[:n_boxes].copy()


def gbuffer_raster(boxes: np.ndarray, width: int, height: int, depth, albedo, emissive, stream=None) -> None:
    """Rasterise a box scene into device G-buffer tensors (planar fp32)."""
    b = np.ascontiguousarray(boxes, np.float64).reshape(-1, 12)
    buf = b if b.size else np.zeros((1, 12))
    n = width * height * 4
    check(_abi.load().rc_gbuffer_raster(buf.ctypes.data, b.shape[0], width, height,
                                        _dev_ptr(depth, n, "depth"), _dev_ptr(albedo, 3 * n, "albedo"),
                                        _dev_ptr(emissive, 3 * n, "emissive"), _stream_handle(stream)))
