#!/usr/bin/env python3
"""GI frames/s of the radiance-cascade pass on one MI355X.

Workload (default): 1920x1080 G-buffer rasterised on the GPU from the seeded
box scene (seed 0, 32 boxes, 4 emissive), 6 cascade levels, fp32 storage; a
step is one full GI frame (occupancy + exact distance transform + 6 levels of
march/merge + gather) with the G-buffer already resident in HBM, its kernels
launched one by one on one stream (the library's default; a replayed
hipGraph leaves ~4 us between frames on this GPU).

Prints ONE JSON line.  `value` is the plain path (the form that meets the
2% rel-L2 bound: its error against the FP64 oracle is ~1e-4, pinned by the
parity tests).  `surrogate` reports the learned far field (levels >= 2
replaced by the cone-traced features + MLP, packaged weights) on the same
G-buffer: its frames/s, the ratio to the plain path, and its rel-L2 against
the plain path on the GPU (FP64 reduction) and, in the CPU leg, against the FP64
oracle frame of the same G-buffer (`accuracy`, with the plain frame's).  `batch` reports the
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


# Kernel families of the plain frame and the stages (rc_get_stats) that
# make them up.  "seed" is the occupancy kernel and "jfa" the distance
# transform: the exact transform's column + row kernels (the default), or
# the jump flood as one family whatever the split between global passes and
# the fused LDS tail, so the dominant family cannot change with the
# autotuned fusion depth.
FAMILIES = {"seed": ["seed"], "jfa": ["jfa", "jfa_tail"], "march": ["march"], "levels": None, "level0": ["level0"]}


def family_alg_bytes(W, H, L, storage, rays, split, jfa_passes, top_in_march=False, sdf="exact"):
    """Algorithmic bytes per frame of each kernel family: the bytes its
    kernels must move, charged to the kernel that moves them.
      seed   : depth + emissive read (16 B/px); written: the exact mode's
               occupancy bits (1/8 B/px), or the jump flood's seed field
               (4 B/px);
      jfa    : exact mode: the band-end kernel reads the bits (1/8 B/px)
               and writes 2 x 2 B per band and column (1/16 B/px), the row
               kernel reads the bits and band ends of its row (1/8 + 1/16
               B/px) and writes the 1-byte distance field;
               jump flood: 8 B/px per pass (seed field read and written
               once), the last pass reads 4 B/px and writes the field;
      march  : the 1-byte distance field read once, a 4-byte code per ray of
               levels >= 1 written, and (top_in_march: the march merges the
               top level itself) the top level's merged texels (b R_top / 4);
      levels : level i >= 1 writes its merged texels (b R_i / 4: the stored
               form is the 4-child mean) and reads level i+1's (b R_{i+1} / 4)
               and, after the split march, its 4-byte codes (4 R_i);
      level0 : the distance field (1 B/px), merged level 1 (b R_1 / 4) and
               the 12 B/px irradiance written.
    b = bytes per stored texel (16 fp32, 8 fp16/bf16)."""
    b = texel_bytes_contract(storage)
    n = W * H
    tex = [r // 4 for r in rays]
    levels = 0
    last = L - 2 if top_in_march else L - 1
    for i in range(1, last + 1):
        levels += b * tex[i] + (b * tex[i + 1] if i + 1 < L else 0) + (4 * rays[i] if split else 0)
    march = (n + 4 * sum(rays[1:L]) + (b * tex[L - 1] if top_in_march else 0)) if split else 0
    if sdf == "exact":
        seed, jfa = 16 * n + n / 8, (n / 8 + n / 16) + (n / 8 + n / 16 + n)
    else:
        seed, jfa = 20 * n, (8 * (jfa_passes - 1) + 5) * n if jfa_passes > 0 else 5 * n
    return {"seed": seed,
            "jfa": jfa,
            "march": march,
            "levels": levels,
            "level0": n + (b * tex[1] if L > 1 else 0) + 12 * n}


def b_alg(W, H, L, storage, rays):
    """The compulsory-traffic contract of the whole frame:
    W H (28 G-buffer + 8 distance field + 12 irradiance) + 2 b sum_{i>=1} R_i."""
    return W * H * 48 + 2 * texel_bytes_contract(storage) * sum(rays[1:L])


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "frame_1080p_plain.md")


def profiled_traffic(family, W, H, L, storage):
    """HBM bytes per frame of a kernel family from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-calibrated, per-family table), for
    the workload that summary was taken on; None otherwise."""
    if (W, H, L, storage) != (1920, 1080, 6, "f32") or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] == f"family:{family}":
            try:
                return float(cells[3]) * 1e6
            except ValueError:
                return None
    return None


def _host_cpu():
    """nproc, the cores this process may run on, and the CPU model name."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "usable_cores": usable, "model": model}


def _time_frames(fn, seconds, min_frames=1):
    fn()  # warm-up (page-in, thread pool)
    frames, t0 = 0, time.perf_counter()
    while True:
        fn()
        frames += 1
        el = time.perf_counter() - t0
        if el >= seconds and frames >= min_frames:
            return frames, el


def cpu_baseline(depth, emissive, L, seconds, frames=None):
    """The CPU restatement of the same maths (the oracle, C) on this host:
    FP32 with OpenMP on every core this process may use (the GPU box's share
    for one GPU is given by OMP_NUM_THREADS when it is set), FP32 on one
    core, and the FP64 flavour on one core.  Whole frames of the workload.
    `frames` ({name: GPU irradiance [3, H, W]}) are then checked against one
    FP64 oracle frame of the same G-buffer: rel-L2 per frame under "checks"."""
    from oracle import rco
    host = _host_cpu()
    threads = host["usable_cores"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = max(1, min(threads, int(os.environ["OMP_NUM_THREADS"])))
    rco.set_threads(threads)
    fr, el = _time_frames(lambda: rco.run(depth, emissive, L, np.float32), seconds)
    rco.set_threads(1)
    f1, e1 = _time_frames(lambda: rco.run(depth, emissive, L, np.float32), seconds / 3)
    f64, e64 = _time_frames(lambda: rco.run(depth, emissive, L, np.float64), seconds / 3)
    rco.set_threads(threads)
    checks = None
    if frames:
        ref = rco.run(depth, emissive, L, np.float64)
        den = float(np.linalg.norm(ref.ravel()))
        checks = {name: (float(np.linalg.norm((np.asarray(f, np.float64) - ref).ravel()) / den) if den > 0 else None)
                  for name, f in frames.items()}   # None: an unlit frame has no relative error
    return {"value": fr / el, "unit": "GI frames/s", "cores": threads, "kind": "port", "checks": checks,
            "sample": f"{fr} full frames of the same workload in {el:.1f} s (oracle FP32 flavour, OpenMP, "
                      f"{threads} threads); single core: FP32 {f1} frames in {e1:.1f} s, FP64 {f64} frames "
                      f"in {e64:.1f} s",
            "single_core_fp32": f1 / e1, "single_core_fp64": f64 / e64,
            "host": host}


# MLP FLOP per level-2 probe (64 cone features -> 64 -> 64 -> 64 -> 64: four 64x64
# layers, multiply-adds x 2)
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
    if args.load_tuning:
        apply_tuning(sur, load_tuning(args.load_tuning))
    elif not args.no_autotune:
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
    learned = out.cpu().numpy()
    sur.close()
    prof.close()
    return learned, {"fps": 1000.0 / ms, "ms_per_step": ms, "ratio_vs_plain": (1000.0 / ms) / fps_plain,
            "rel_l2_vs_plain": err["rel_l2"], "max_abs_vs_plain": err["max_abs"],
            "meets_2pct": bool(err["rel_l2"] <= 0.02),
            "fallback": "off (threshold +inf)",
            "stage_ms_per_frame": {k: round(v, 5) for k, v in per_frame.items()},
            "mlp_roofline": {"bound": "mfma", "kernel": "k_surr_mlp (features + MLP + level-2 texels)",
                             "flop_per_launch": flop, "achieved": tflops, "peak": BF16_MFMA_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": tflops / BF16_MFMA_PEAK_TFLOPS}}


def _time_stage(torch, stage, depth, emis, out, stream, warmup, steps):
    """Mean ms per frame of `steps` frames launched back to back on `stream`."""
    for _ in range(warmup):
        stage.run_device(depth, emis, out, stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        stage.run_device(depth, emis, out, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def storage_leg(rc, torch, W, H, L, args, depth, emis, irr_f32, stream):
    """The plain frame with 16-bit storage of the merged levels (fp16, bf16),
    autotuned like the fp32 one: frames/s and irradiance rel-L2 / max-abs
    against the fp32 frame (GPU FP64 reduction)."""
    out = {}
    for st in ("f16", "bf16"):
        stage = rc.RadianceCascadeStage(W, H, levels=L, storage=st)
        o = torch.empty_like(irr_f32)
        stage.autotune(depth, emis, o, iters=2, stream=stream)
        ms = _time_stage(torch, stage, depth, emis, o, stream, args.warmup, args.steps)
        err = stage.error_norms(o, irr_f32, stream)
        out[st] = {"fps": 1000.0 / ms, "ms_per_step": ms, "rel_l2_vs_fp32": err["rel_l2"],
                   "max_abs_vs_fp32": err["max_abs"]}
        stage.close()
    return out


def uhd_leg(rc, torch, args, stream):
    """3840x2160 with 7 levels (the seeded scene of the main workload):
    plain and learned frames/s, both autotuned, and the learned frame's
    rel-L2 against the plain one."""
    W, H, L = 3840, 2160, 7
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    rc.gbuffer_raster(rc.scene_generate(args.seed, args.boxes, args.emissive), W, H, depth, None, emis)
    irr = torch.empty((3, H, W), device=dev)
    out = torch.empty_like(irr)
    plain = rc.RadianceCascadeStage(W, H, levels=L, storage=args.storage)
    plain.autotune(depth, emis, irr, iters=2, stream=stream)
    ms_p = _time_stage(torch, plain, depth, emis, irr, stream, args.warmup, args.steps)
    sur = rc.RadianceCascadeStage(W, H, levels=L, storage=args.storage)
    sur.load_weights(args.weights)
    sur.set_surrogate(True)
    sur.autotune(depth, emis, out, iters=2, stream=stream)
    ms_s = _time_stage(torch, sur, depth, emis, out, stream, args.warmup, args.steps)
    err = sur.error_norms(out, irr, stream)
    plain.close()
    sur.close()
    return {"workload": f"{W}x{H}, {L} cascades, seed {args.seed}", "plain_fps": 1000.0 / ms_p,
            "learned_fps": 1000.0 / ms_s, "ratio": ms_p / ms_s, "learned_rel_l2_vs_plain": err["rel_l2"]}


def load_tuning(arg):
    """A tuning as written by --save-tuning: a JSON file, or the JSON text itself."""
    text = arg.strip()
    return json.loads(text) if text.startswith("{") else json.load(open(arg))


def apply_tuning(stage, tun):
    for lvl, (shape, rpt) in tun["tiling"].items():
        if int(lvl) < stage.levels:
            stage.set_tiling(int(lvl), shape, rpt)
    stage.set_march(*tun["march"])
    stage.set_jfa_fuse(tun["jfa_fuse"])
    stage.set_jfa_rows(tun.get("jfa_rows", 1))
    for lvl, order in tun.get("march_order", {}).items():
        if int(lvl) < stage.levels:
            stage.set_march_order(int(lvl), order)


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


def batch_leg(rc, torch, args, stream, fps_plain):
    """The multi-view workload: 64 seeded views (seeds 1000..1063) of uneven
    sizes (rc.batch_views_spec), plain cascades with the coverage rule's level
    count, one graph launch per batch, in both batch modes: the work list (one
    launch per stage over every view, one shared workspace, tuned as a whole)
    and one context per view on 8 captured streams (tuned per distinct view
    size).  Aggregate megapixels of irradiance per second, and the ratio to the
    single-view 1080p rate of the plain path."""
    dev = torch.device("cuda:0")
    spec = rc.batch_views_spec(1000, 64)
    views = []
    for seed, W, H, nb, ne in spec:
        d = torch.empty((H, W), device=dev)
        e = torch.empty((3, H, W), device=dev)
        rc.gbuffer_raster(rc.scene_generate(seed, nb, ne), W, H, d, None, e)
        views.append((d, e, torch.empty((3, H, W), device=dev)))
    mp = sum(W * H for _, W, H, _, _ in spec) / 1e6
    single_mps = fps_plain * args.width * args.height / 1e6
    out = {"views": len(spec), "megapixels": mp, "single_view_mpix_per_s": single_mps}
    for mode in ("worklist", "streams"):
        b = rc.RadianceCascadeBatch(views, storage=args.storage, mode=mode, streams=8)
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
        info = b.info()
        levels = [b.levels(i) for i in range(len(spec))]
        b.close()
        out[mode] = {"ms_per_batch": ms, "mpix_per_s": mp / (ms * 1e-3), "views_per_s": len(spec) / (ms * 1e-3),
                     "vs_single_view": mp / (ms * 1e-3) / single_mps, "kernel_nodes": info["kernel_nodes"],
                     "workspace_bytes": info["workspace_bytes"]}
        out["levels"] = {str(L): levels.count(L) for L in sorted(set(levels))}
    out["mpix_per_s"] = out["worklist"]["mpix_per_s"]
    out["launches"] = "one graph launch per batch"
    return out


# kernel name prefix -> family (the plain frame's kernels)
KERNEL_FAMILY = [("k_seed_init", "seed"), ("k_occ_bits", "seed"), ("k_edt_ends", "jfa"), ("k_edt_rows", "jfa"),
                 ("k_jfa_pass", "jfa"), ("k_jfa_tail", "jfa"), ("k_jfa_none", "jfa"),
                 ("k_march_all", "march"), ("k_march_persist", "march"), ("k_level0", "level0"),
                 ("k_level<", "levels")]


def _family_of(name):
    short = name.split("rc::", 1)[-1]
    for prefix, fam in KERNEL_FAMILY:
        if short.startswith(prefix):
            return fam
    return None


def _rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _col(row, *names):
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(names)


def _counter_per_kernel(path, counter):
    """{kernel name: (dispatches, mean counter value)} from a rocprofv3
    counter-collection CSV (one row per dispatch and counter)."""
    acc = {}
    for r in _rows(path):
        if _col(r, "Counter_Name") != counter:
            continue
        name = _col(r, "Kernel_Name")
        n, v = acc.get(name, (0, 0.0))
        acc[name] = (n + 1, v + float(_col(r, "Counter_Value")))
    return {k: (n, v / n) for k, (n, v) in acc.items()}


def summarize_profile(trace_stats, fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_bytes, workload):
    """Markdown summary of one profiled bench configuration: per kernel the
    calls and mean duration (kernel trace) and HBM bytes per dispatch
    (FETCH_SIZE and WRITE_SIZE in KiB, each scaled by what the same counter
    reads for a plain device copy of known size), then per kernel family the
    time and bytes per frame (frames = level-0 launches)."""
    stats = {}
    for r in _rows(trace_stats):
        stats[_col(r, "Name", "Kernel_Name")] = (int(_col(r, "Calls")), float(_col(r, "AverageNs")) / 1e3)
    fetch = _counter_per_kernel(fetch_csv, "FETCH_SIZE")
    write = _counter_per_kernel(write_csv, "WRITE_SIZE")
    cf = _counter_per_kernel(copy_fetch_csv, "FETCH_SIZE")
    cw = _counter_per_kernel(copy_write_csv, "WRITE_SIZE")
    copy_name = max(cf, key=lambda k: cf[k][1])
    kf = copy_bytes / (cf[copy_name][1] * 1024.0)   # bytes per FETCH_SIZE KiB unit
    kw = copy_bytes / (cw[copy_name][1] * 1024.0)
    frames = sum(c for n, (c, _) in stats.items() if _family_of(n) == "level0")
    lines = [f"Workload: {workload}", "",
             f"Copy calibration ({copy_name}, {copy_bytes} B read and written per launch): FETCH_SIZE reads "
             f"{cf[copy_name][1]:.0f} KiB, WRITE_SIZE {cw[copy_name][1]:.0f} KiB, so bytes = FETCH_SIZE x {kf * 1024:.3f} "
             f"and WRITE_SIZE x {kw * 1024:.3f} (per KiB).", "",
             f"Frames in the run: {frames}", "",
             "| kernel | family | calls | avg us | read MB / dispatch | written MB / dispatch |",
             "|---|---|---|---|---|---|"]
    fam = {}
    for name, (calls, us) in sorted(stats.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        f = _family_of(name)
        rd = fetch.get(name, (0, float("nan")))[1] * 1024 * kf / 1e6
        wr = write.get(name, (0, float("nan")))[1] * 1024 * kw / 1e6
        short = name if len(name) < 90 else name[:87] + "..."
        lines.append(f"| `{short}` | {f or '-'} | {calls} | {us:.2f} | {rd:.2f} | {wr:.2f} |")
        if f and frames:
            t, b, n = fam.get(f, (0.0, 0.0, 0.0))
            fam[f] = (t + calls * us / frames, b + calls * (rd + wr) / frames, n + calls / frames)
    lines += ["", "Per kernel family and frame (the bench's `traffic` reads the `family:` rows):", "",
              "| family | launches / frame | us / frame | HBM MB / frame |", "|---|---|---|---|"]
    for f, (t, b, n) in fam.items():
        lines.append(f"| `family:{f}` | {n:.2f} | {t:.2f} | {b:.2f} |")
    return "\n".join(lines)


def summarize_counters(csv_paths, title):
    """Markdown table of rocprofv3 counter-collection CSVs (one or more
    --pmc passes of the same command): per kernel, the dispatches and the
    mean per dispatch of every counter (summed over its dimension
    instances), plus, where the counters are present, VALU and SALU
    instructions per wave."""
    per = {}
    for path in csv_paths:
        disp = {}
        for r in _rows(path):
            name = _col(r, "Kernel_Name")
            key = (name, _col(r, "Counter_Name"))
            d = _col(r, "Dispatch_Id")
            disp.setdefault(key, {}).setdefault(d, 0.0)
            disp[key][d] += float(_col(r, "Counter_Value"))
        for (name, counter), by in disp.items():
            per.setdefault(name, {})[counter] = (len(by), sum(by.values()) / len(by))
    counters = sorted({c for v in per.values() for c in v})
    lines = [f"Counters: {title}", "", "| kernel | dispatches | " + " | ".join(counters) +
             " | VALU / wave | SALU / wave |", "|---" * (len(counters) + 4) + "|"]
    for name, v in sorted(per.items(), key=lambda kv: kv[0]):
        short = name.split("(")[0].replace("void ", "")
        n = max(c[0] for c in v.values())
        cells = [f"{v[c][1]:.4g}" if c in v else "-" for c in counters]
        waves = v.get("SQ_WAVES", (0, 0.0))[1]
        valu = f"{v['SQ_INSTS_VALU'][1] / waves:.1f}" if waves and "SQ_INSTS_VALU" in v else "-"
        salu = f"{v['SQ_INSTS_SALU'][1] / waves:.1f}" if waves and "SQ_INSTS_SALU" in v else "-"
        lines.append(f"| `{short}` | {n} | " + " | ".join(cells) + f" | {valu} | {salu} |")
    return "\n".join(lines)


def summarize_trace(stats_csv, title):
    """Markdown table of a rocprofv3 kernel-trace stats CSV: calls, mean, min
    and max duration and share of GPU time per kernel."""
    lines = [f"Kernel trace: {title}", "", "| kernel | calls | avg us | min us | max us | % of GPU time |",
             "|---|---|---|---|---|---|"]
    for r in _rows(stats_csv):
        name = _col(r, "Name", "Kernel_Name")
        short = name if len(name) < 90 else name[:87] + "..."
        lines.append(f"| `{short}` | {_col(r, 'Calls')} | {float(_col(r, 'AverageNs')) / 1e3:.2f} | "
                     f"{float(_col(r, 'MinNs')) / 1e3:.2f} | {float(_col(r, 'MaxNs')) / 1e3:.2f} | "
                     f"{float(_col(r, 'Percentage')):.2f} |")
    return "\n".join(lines)


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
    ap.add_argument("--sdf", default="exact", choices=["exact", "jfa"],
                    help="distance transform of the timed frame (exact: the library default)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-storage", action="store_true", help="skip the fp16 / bf16 storage leg")
    ap.add_argument("--no-uhd", action="store_true", help="skip the 3840x2160, 7-level leg")
    ap.add_argument("--no-host-staging", action="store_true",
                    help="skip the host-staging leg (profiled runs: its contexts use the default configuration)")
    ap.add_argument("--weights", default=None, help="surrogate weights (default: the packaged network)")
    ap.add_argument("--save-tuning", default=None, help="write the autotuned tiling / march / jump-flood choice (JSON)")
    ap.add_argument("--load-tuning", default=None,
                    help="use a saved tuning instead of autotuning (profiled runs): a JSON file as written by "
                         "--save-tuning, or the JSON text itself")
    ap.add_argument("--summarize-profile", nargs=6, default=None,
                    metavar=("TRACE_STATS", "FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE", "WORKLOAD"),
                    help="print the markdown summary of rocprofv3 CSVs of this bench (kernel stats, FETCH_SIZE and "
                         "WRITE_SIZE runs of the bench and of --copy-calibration)")
    ap.add_argument("--summarize-counters", nargs="+", default=None, metavar="CSV_OR_TITLE",
                    help="print the per-kernel markdown table of rocprofv3 counter CSVs (the last argument is the "
                         "title)")
    ap.add_argument("--summarize-trace", nargs=2, default=None, metavar=("TRACE_STATS", "TITLE"),
                    help="print the markdown table of a rocprofv3 kernel-trace stats CSV")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB device buffer 10 times (the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters of rocprofv3 to bytes)")
    args = ap.parse_args()
    if args.summarize_profile:
        print(summarize_profile(*args.summarize_profile[:5], copy_bytes=4 * (64 << 20),
                                workload=args.summarize_profile[5]))
        return
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters[:-1], args.summarize_counters[-1]))
        return
    if args.summarize_trace:
        print(summarize_trace(*args.summarize_trace))
        return
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
    stage.set_sdf_mode(args.sdf)
    if args.load_tuning:
        apply_tuning(stage, load_tuning(args.load_tuning))
        tiling = stage.tiling()
    elif not args.no_autotune:
        tiling = stage.autotune(depth, emis, irr, iters=3, stream=stream)
    else:
        tiling = stage.tiling()
    if args.save_tuning:
        inv = {v: k for k, v in rc._abi.TILE_SHAPES.items()}
        json.dump({"tiling": {str(l): [inv[tuple(v[0])], v[1]] for l, v in tiling.items()},
                   "march": list(stage.march()), "jfa_fuse": stage.jfa_fuse(), "jfa_rows": stage.jfa_rows(),
                   "march_order": {str(k): v for k, v in stage.march_order().items()}}, open(args.save_tuning, "w"))
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
    prof.set_sdf_mode(stage.sdf_mode())
    prof.set_jfa_fuse(stage.jfa_fuse())
    prof.set_jfa_rows(stage.jfa_rows())
    for lvl, order in stage.march_order().items():
        prof.set_march_order(lvl, order)
    for _ in range(5):
        prof.run_device(depth, emis, irr, stream)
    prof.reset_stats()
    n_prof = max(20, min(args.steps, 100))
    for _ in range(n_prof):
        prof.run_device(depth, emis, irr, stream)
    torch.cuda.synchronize()
    st = prof.stats()
    rays = [rc.level_rays(W, H, i) for i in range(L)]
    per_frame_ms = {k: st.stage_ms[k] / st.runs for k in st.stage_ms}
    top_in_march = bool(split) and L > 1 and st.stage_launches.get(f"level{L - 1}", 0) == 0
    fam_alg = family_alg_bytes(W, H, L, args.storage, rays, bool(split), (max(W, H) - 1).bit_length(),  # P = ceil(log2 max(W, H))
                               top_in_march, stage.sdf_mode())
    fam_ms, fam_launches = {}, {}
    for fam, stages in FAMILIES.items():
        stages = stages or [f"level{i}" for i in range(1, L)]
        got = [k for k in stages if k in st.stage_ms]
        if got:
            fam_ms[fam] = sum(st.stage_ms[k] for k in got) / st.runs
            fam_launches[fam] = sum(st.stage_launches[k] for k in got) / st.runs
    dominant = max(fam_ms, key=fam_ms.get)
    dom_ms = fam_ms[dominant] / fam_launches[dominant]          # mean launch
    dom_bytes = fam_alg[dominant] / fam_launches[dominant]      # per launch
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = profiled_traffic(dominant, W, H, L, args.storage)
    frame_alg = b_alg(W, H, L, args.storage, rays)

    surrogate, learned = None, None
    if not args.no_surrogate and L >= 3:
        learned, surrogate = surrogate_leg(rc, torch, W, H, L, args, depth, emis, irr, stream, fps)

    storage = None if args.no_storage else storage_leg(rc, torch, W, H, L, args, depth, emis, irr, stream)
    uhd = None if args.no_uhd else uhd_leg(rc, torch, args, stream)
    batch = None if args.no_batch else batch_leg(rc, torch, args, stream, fps)
    staging = None if args.no_host_staging else host_staging_leg(rc, W, H, L, args.storage, depth, emis)

    cpu, accuracy = None, None
    if not args.no_cpu_baseline:
        frames = {"plain": irr.cpu().numpy()}
        if learned is not None:
            frames["learned"] = learned
        cpu = cpu_baseline(depth.cpu().numpy(), emis.cpu().numpy(), L, args.cpu_seconds, frames)
        accuracy = {"plain_rel_l2_vs_fp64": cpu["checks"]["plain"],
                    "learned_rel_l2_vs_fp64": cpu["checks"].get("learned"),
                    "note": "irradiance rel-L2 of the GPU frames against the FP64 oracle frame of the same "
                            "G-buffer (exact distance transform), computed in the CPU leg"}
        if surrogate is not None and accuracy["learned_rel_l2_vs_fp64"] is not None:
            surrogate["rel_l2_vs_fp64"] = accuracy["learned_rel_l2_vs_fp64"]
            surrogate["meets_2pct_vs_fp64"] = bool(accuracy["learned_rel_l2_vs_fp64"] <= 0.02)

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
                   "march": {"split": split, "speculation": spec, "rays_per_lane": rpl, "top_level_merged": top_in_march,
                             "order": {str(k): v for k, v in stage.march_order().items()}},
                   "sdf": stage.sdf_mode(), "jfa_fuse": stage.jfa_fuse(), "jfa_rows": stage.jfa_rows()},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic / fam_launches[dominant],
                     "traffic_over_alg": None if traffic is None else traffic / fam_alg[dominant],
                     "traffic_source": "profiles/frame_1080p_plain.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                       "copy-calibrated, per kernel family and frame)",
                     "alg_bytes_per_launch": dom_bytes, "mean_launch_ms": dom_ms,
                     "launches_per_frame": fam_launches[dominant]},
        "families": {f: {"ms_per_frame": round(fam_ms[f], 5), "launches": fam_launches[f],
                         "alg_bytes": fam_alg[f],
                         "frac": fam_alg[f] / (fam_ms[f] * 1e-3) / 1e9 / HBM_PEAK_GBS} for f in fam_ms},
        "frame_roofline": {"alg_bytes": frame_alg, "achieved": frame_alg / (ms_per_step * 1e-3) / 1e9,
                           "frac": frame_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "stage_ms_per_frame": {k: round(v, 5) for k, v in per_frame_ms.items()},
        "surrogate": surrogate,
        "accuracy": accuracy,
        "storage": storage,
        "uhd": uhd,
        "batch": batch,
        "host_staging": staging,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    stage.close()
    prof.close()


if __name__ == "__main__":
    main()
