#!/usr/bin/env python3
"""GI frames/s of the radiance-cascade pass on one MI355X.

Workload (default): 1920x1080 G-buffer rasterised on the GPU from the seeded
box scene (seed 0, 32 boxes, 4 emissive), 6 cascade levels, fp32 storage; a
step is one full GI frame (occupancy + jump flood + SDF + 6 levels of
march/merge + gather) with the G-buffer already resident in HBM, its kernels
launched one by one on one stream (the library's default; a replayed
hipGraph leaves ~4 us between frames on this GPU).

Prints ONE JSON line.  `value` is the plain path (the form that meets the
2% rel-L2 bound: its error against the FP64 oracle is ~1e-4, pinned by the
parity tests).  `surrogate` reports the learned far field (levels >= 2
replaced by the cone-traced features + MLP, packaged weights) on the same
G-buffer: its frames/s, the ratio to the plain path, and its rel-L2 judged by
the plain path on the GPU (FP64 reduction).  `batch` reports the
multi-view workload (64 seeded views of uneven sizes, one graph launch) in
aggregate megapixels/s.  `roofline` is the dominant
kernel's algorithmic bytes per launch over its mean launch time, measured
here with HIP events recorded on the stream the kernels run on;
`cpu_baseline` times the FP32 CPU restatement of the same maths (the oracle's
FP32 flavour, OpenMP) on the same G-buffer on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def texel_bytes_contract(storage: str) -> int:
    return 16 if storage == "f32" else 8


def kernel_alg_bytes(W, H, L, storage, rays):
    """Algorithmic bytes per launch of each kernel, from the compulsory-traffic
    model B_alg = W H (28 + 8 + 12) + 2 b sum_{i>=1} R_i split by the kernel
    that moves each term: seed kernel reads the 28 B/px G-buffer; the JFA
    tail writes the 4 B/px SDF; level i >= 1 writes merged level i (b R_i)
    and reads merged level i+1 (b R_{i+1}); level 0 reads the SDF (4 B/px),
    merged level 1 (b R_1) and writes 12 B/px irradiance."""
    b = texel_bytes_contract(storage)
    n = W * H
    # a global jump-flood pass is charged the informative 8 B/px (seed field
    # read once, written once); the contract itself has no JFA term.
    out = {"seed": 28 * n, "jfa": 8 * n, "jfa_tail": 4 * n}
    for i in range(1, L):
        out[f"level{i}"] = b * rays[i] + (b * rays[i + 1] if i + 1 < L else 0)
    out["level0"] = 4 * n + 12 * n + (b * rays[1] if L > 1 else 0)
    # split march of all levels >= 1: reads the SDF once (4 B/px) and produces
    # one b-byte ray-interval result per ray (the march row of the model);
    # the merge kernels are then charged the same contract terms as above.
    out["march"] = 4 * n + b * sum(rays[1:L])
    return out


def b_alg(W, H, L, storage, rays):
    return W * H * 48 + 2 * texel_bytes_contract(storage) * sum(rays[1:L])


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "frame_1080p_hbm_traffic.md")
KERNEL_NAMES = {"march": "k_march_all", "jfa": "k_jfa_pass<false>", "seed": "k_seed_init", "level0": "k_level0<"}


def profiled_traffic(stage, W, H, L, storage):
    """HBM bytes per launch of the stage's kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel factor there),
    for the workload that summary was taken on; None otherwise."""
    if (W, H, L, storage) != (1920, 1080, 6, "f32") or stage not in KERNEL_NAMES or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].strip("`").startswith("rc::" + KERNEL_NAMES[stage]):
            try:
                return (float(cells[3]) + float(cells[4])) * 1e6
            except ValueError:
                return None
    return None


def cpu_baseline(depth, emissive, L, seconds):
    """FP32 CPU restatement (OpenMP) on this host's cores: frames/s."""
    from oracle import rco
    threads = max(1, min(16, os.cpu_count() or 1))
    rco.set_threads(threads)
    rco.run(depth, emissive, L, np.float32)  # warm-up (page-in, thread pool)
    frames, t0 = 0, time.perf_counter()
    while True:
        rco.run(depth, emissive, L, np.float32)
        frames += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": frames / el, "unit": "GI frames/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full frames of the same workload in {el:.1f} s "
                      f"(oracle FP32 flavour, OpenMP, {threads} threads)"}


# MLP FLOP per level-1 probe (48 -> 64 -> 64 -> 64 -> 64, multiply-adds x 2)
MLP_FLOP_PER_PROBE = 2 * (64 * 64 + 3 * 64 * 64)
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense bf16


def surrogate_leg(rc, torch, W, H, L, args, depth, emis, irr_plain, stream, fps_plain):
    """The learned far field on the same G-buffer: frames/s (direct launches,
    as the plain path), rel-L2 against the plain path's irradiance (GPU FP64
    reduction), per-stage times and the MFMA roofline of the network kernel."""
    sur = rc.RadianceCascadeStage(W, H, levels=L, storage=args.storage)
    sur.load_weights(args.weights)
    sur.set_surrogate(True)
    out = torch.empty_like(irr_plain)
    if not args.no_autotune:
        sur.autotune(depth, emis, out, iters=3, stream=stream)   # as the plain path (its level 1 too)
    for _ in range(args.warmup):
        sur.run_device(depth, emis, out, stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        sur.run_device(depth, emis, out, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    err = sur.error_norms(out, irr_plain, stream)
    prof = rc.RadianceCascadeStage(W, H, levels=L, storage=args.storage, profile=True)
    prof.load_weights(args.weights)
    prof.set_surrogate(True)
    for _ in range(5):
        prof.run_device(depth, emis, out, stream)
    prof.reset_stats()
    for _ in range(50):
        prof.run_device(depth, emis, out, stream)
    torch.cuda.synchronize()
    st = prof.stats()
    per_frame = {k: st.stage_ms[k] / st.runs for k in st.stage_ms}
    px, py = rc.probe_grid(W, H, 2)   # the network runs once per level-2 probe
    flop = MLP_FLOP_PER_PROBE * px * py
    mlp_ms = per_frame.get("surrogate", float("nan"))
    tflops = flop / (mlp_ms * 1e-3) / 1e12
    sur.close()
    prof.close()
    return {"fps": 1000.0 / ms, "ms_per_step": ms, "ratio_vs_plain": (1000.0 / ms) / fps_plain,
            "rel_l2_vs_plain": err["rel_l2"], "max_abs_vs_plain": err["max_abs"],
            "meets_2pct": bool(err["rel_l2"] <= 0.02),
            "fallback": "off (threshold +inf)",
            "stage_ms_per_frame": {k: round(v, 5) for k, v in per_frame.items()},
            "mlp_roofline": {"bound": "mfma", "kernel": "k_surr_mlp (features + MLP + level-2 texels)",
                             "flop_per_launch": flop, "achieved": tflops, "peak": BF16_MFMA_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": tflops / BF16_MFMA_PEAK_TFLOPS}}


def host_staging_leg(rc, W, H, L, storage, depth, emis, frames=20):
    """rc_update_host + rc_render_host per frame (host G-buffer in, host
    irradiance out: the PCIe-inclusive rate of the staging boundary)."""
    def run(d, e, out):
        st = rc.RadianceCascadeStage(W, H, levels=L, storage=storage)
        st.update(d, e)
        st.render(out)
        t0 = time.perf_counter()
        for _ in range(frames):
            st.update(d, e)
            st.render(out)
        el = time.perf_counter() - t0
        st.close()
        return el

    d = depth.cpu().numpy()
    e = emis.cpu().numpy()
    el = run(d, e, np.empty((3, H, W), np.float32))
    # the same frames with the planes in page-locked memory (rc_host_alloc)
    dp, ep, op = rc.host_buffer((H, W)), rc.host_buffer((3, H, W)), rc.host_buffer((3, H, W))
    dp[...] = d
    ep[...] = e
    elp = run(dp, ep, op)
    return {"fps": frames / el, "ms_per_frame": 1e3 * el / frames,
            "pinned_fps": frames / elp, "pinned_ms_per_frame": 1e3 * elp / frames,
            "bytes_per_frame": 4 * W * H * (1 + 3 + 3),
            "note": "host G-buffer upload + frame + irradiance download; pageable, then page-locked host buffers"}


def batch_leg(rc, torch, args, stream):
    """The multi-view workload: 64 seeded views (seeds 1000..1063) of uneven
    sizes (rc.batch_views_spec), plain cascades with the coverage rule's level
    count, run as one graph launch (8 captured streams), tiling autotuned per
    distinct view size.  Aggregate megapixels of irradiance per second."""
    dev = torch.device("cuda:0")
    spec = rc.batch_views_spec(1000, 64)
    views = []
    for seed, W, H, nb, ne in spec:
        d = torch.empty((H, W), device=dev)
        e = torch.empty((3, H, W), device=dev)
        rc.gbuffer_raster(rc.scene_generate(seed, nb, ne), W, H, d, None, e)
        views.append((d, e, torch.empty((3, H, W), device=dev)))
    b = rc.RadianceCascadeBatch(views, storage=args.storage, streams=8)
    b.autotune(iters=1, stream=stream)
    for _ in range(3):
        b.run(stream)
    torch.cuda.synchronize()
    n = 10
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(n):
        b.run(stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / n
    mp = sum(W * H for _, W, H, _, _ in spec) / 1e6
    levels = [b.levels(i) for i in range(len(spec))]
    b.close()
    return {"views": len(spec), "megapixels": mp, "ms_per_batch": ms, "mpix_per_s": mp / (ms * 1e-3),
            "views_per_s": len(spec) / (ms * 1e-3), "levels": {str(L): levels.count(L) for L in sorted(set(levels))},
            "streams": 8, "launches": "one graph launch per batch"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--boxes", type=int, default=32)
    ap.add_argument("--emissive", type=int, default=4)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--weights", default=None, help="surrogate weights (default: the packaged network)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB device buffer 10 times (the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters of rocprofv3 to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 64 << 20
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(10):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 10}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: the GI pass is single-GPU by design "
              "(multi-GPU sharding is described in the design notes, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    import radiancecascade_amd as rc

    W, H, L = args.width, args.height, args.levels
    dev = torch.device("cuda:0")
    boxes = rc.scene_generate(args.seed, args.boxes, args.emissive)
    depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    albedo = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    emis = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    irr = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    rc.gbuffer_raster(boxes, W, H, depth, albedo, emis)
    stream = torch.cuda.current_stream()

    stage = rc.RadianceCascadeStage(W, H, levels=L, storage=args.storage)
    tiling = stage.autotune(depth, emis, irr, iters=3, stream=stream) if not args.no_autotune else stage.tiling()
    for _ in range(args.warmup):
        stage.run_device(depth, emis, irr, stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        stage.run_device(depth, emis, irr, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    ms_per_step = ms_total / args.steps
    fps = 1000.0 / ms_per_step

    # Per-kernel launch times: a second context with HIP events around every
    # stage on the same stream (synchronises once per frame, so it is not the
    # throughput figure above).
    prof = rc.RadianceCascadeStage(W, H, levels=L, storage=args.storage, profile=True)
    for lvl, ((tx, ty, kd), rpt) in tiling.items():
        shape = [k for k, v in rc._abi.TILE_SHAPES.items() if v == (tx, ty, kd)][0]
        prof.set_tiling(lvl, shape, rpt)
    split, spec, rpl = stage.march()
    prof.set_march(split, spec, rpl)
    prof.set_jfa_fuse(stage.jfa_fuse())
    for _ in range(5):
        prof.run_device(depth, emis, irr, stream)
    prof.reset_stats()
    n_prof = max(20, min(args.steps, 100))
    for _ in range(n_prof):
        prof.run_device(depth, emis, irr, stream)
    torch.cuda.synchronize()
    st = prof.stats()
    rays = [rc.level_rays(W, H, i) for i in range(L)]
    alg = kernel_alg_bytes(W, H, L, args.storage, rays)
    per_launch_ms = {k: st.stage_ms[k] / st.stage_launches[k] for k in st.stage_ms}
    per_frame_ms = {k: st.stage_ms[k] / st.runs for k in st.stage_ms}
    dominant = max(per_frame_ms, key=per_frame_ms.get)
    dom_ms = per_launch_ms[dominant]
    achieved = alg[dominant] / (dom_ms * 1e-3) / 1e9
    frame_alg = b_alg(W, H, L, args.storage, rays)

    surrogate = None
    if not args.no_surrogate and L >= 3:
        surrogate = surrogate_leg(rc, torch, W, H, L, args, depth, emis, irr, stream, fps)

    batch = None if args.no_batch else batch_leg(rc, torch, args, stream)
    staging = host_staging_leg(rc, W, H, L, args.storage, depth, emis)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(depth.cpu().numpy(), emis.cpu().numpy(), L, args.cpu_seconds)

    line = {
        "metric": "GI frames/s @1080p 6 cascades: surrogate vs plain HIP at <=2% rel-L2 error",
        "value": fps,
        "unit": "GI frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.storage == "f32" else f"fp32 ({args.storage} storage)",
        "data": "synthetic",
        "config": {"workload": f"plain HIP path, {W}x{H} G-buffer, {L} cascades, {args.storage} storage, "
                               f"seeded box scene (seed {args.seed}, {args.boxes} boxes, {args.emissive} emissive)",
                   "path": "plain", "launch": "direct (one stream)", "width": W, "height": H, "levels": L,
                   "storage": args.storage,
                   "rays_per_frame": int(sum(rays)),
                   "tiling": {str(k): f"{v[0][0]}x{v[0][1]}x{v[0][2]} r{v[1]}" for k, v in tiling.items()},
                   "march": {"split": split, "speculation": spec, "rays_per_lane": rpl}, "jfa_fuse": stage.jfa_fuse()},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(dominant, W, H, L, args.storage),
                     "traffic_source": "profiles/frame_1080p_hbm_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                       "copy-calibrated)",
                     "alg_bytes_per_launch": alg[dominant], "mean_launch_ms": dom_ms},
        "frame_roofline": {"alg_bytes": frame_alg, "achieved": frame_alg / (ms_per_step * 1e-3) / 1e9,
                           "frac": frame_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "stage_ms_per_frame": {k: round(v, 5) for k, v in per_frame_ms.items()},
        "surrogate": surrogate,
        "batch": batch,
        "host_staging": staging,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    stage.close()
    prof.close()


if __name__ == "__main__":
    main()
