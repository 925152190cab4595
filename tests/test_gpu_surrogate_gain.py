"""What the trained network gains: the packaged weights (trained on the GPU
on 8,000 scenes at 1920x1080, seeds 10000..17999, 6 cascades) evaluated on
held-out 1920x1080 scenes with 6 cascades.

Three forms of the frame replace levels >= 2:
  truncated  levels 0 and 1 only (the method without any far field);
  proxy      the cone-traced far field alone (the network's weights zeroed);
  learned    cone trace + trained network (the shipped surrogate).
Against the FP64 oracle on four scenes, the learned frame must beat the
truncated one on every scene.  Against the plain GPU path (equal to the
oracle within 1e-4 here) on the lit scenes of seeds 100..159 (not used to
choose the weights, which were compared on seeds 20..59), the learned frame
must beat the proxy in median and in mean and on more than half of the
scenes: the gain of the network itself.  (On some scenes it loses: e.g.
seed 4, whose light comes from a 4-px emissive strip along the image's top
edge.)
"""
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W, H, L = 1920, 1080, 6
SEEDS = [0, 1, 3, 9]         # held-out scenes with visible emitters (32 boxes, 4 emissive)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rc():
    import radiancecascade_amd as rc
    if rc.device_count() < 1:
        pytest.skip("no HIP device")
    if not os.path.exists(rc.packaged_weights()):
        pytest.skip("no packaged weights")
    return rc


def test_learned_far_field_gain(rc):
    from oracle import rco
    rco.set_threads(16)
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    learned = rc.RadianceCascadeStage(W, H, levels=L)
    learned.load_weights()
    learned.set_surrogate(True)
    m = learned.mlp_get()
    proxy = rc.RadianceCascadeStage(W, H, levels=L)
    proxy.mlp_set(np.zeros_like(m["params"]), m["in_mean"], m["in_istd"], m["out_scale"])
    proxy.set_surrogate(True)
    trunc = rc.RadianceCascadeStage(W, H, levels=2)
    plain = rc.RadianceCascadeStage(W, H, levels=L)
    rows = []
    for seed in SEEDS:
        rc.gbuffer_raster(rc.scene_generate(seed, 32, 4), W, H, depth, None, emis)
        ref = torch.from_numpy(rco.run(depth.cpu().numpy(), emis.cpu().numpy(), L)).to(dev)
        err = {}
        for name, st in (("plain", plain), ("truncated", trunc), ("proxy", proxy), ("learned", learned)):
            out = torch.empty((3, H, W), device=dev)
            st.run_device(depth, emis, out)
            torch.cuda.synchronize()
            err[name] = st.error_norms(out, ref)["rel_l2"]
        rows.append(err)
        _say(f"  seed {seed}: rel-L2 vs FP64  plain {err['plain']:.2e}  truncated {err['truncated']:.4f}  "
             f"proxy {err['proxy']:.4f}  learned {err['learned']:.4f}")
        assert err["plain"] < 1e-3
        assert err["learned"] < err["truncated"]
    med = {k: float(np.median([r[k] for r in rows])) for k in rows[0]}
    _say(f"  median: truncated {med['truncated']:.4f}  proxy {med['proxy']:.4f}  learned {med['learned']:.4f}")
    assert med["learned"] < 0.5 * med["truncated"]
    # the network against the cones alone, over more scenes
    ep, el = [], []
    for seed in range(100, 160):
        rc.gbuffer_raster(rc.scene_generate(seed, 32, 4), W, H, depth, None, emis)
        ref = torch.empty((3, H, W), device=dev)
        plain.run_device(depth, emis, ref)
        torch.cuda.synchronize()
        if float(ref.norm()) == 0.0:
            continue
        for st, acc in ((proxy, ep), (learned, el)):
            out = torch.empty((3, H, W), device=dev)
            st.run_device(depth, emis, out)
            torch.cuda.synchronize()
            acc.append(st.error_norms(out, ref)["rel_l2"])
    ep, el = np.array(ep), np.array(el)
    wins = int((el < ep).sum())
    _say(f"  seeds 100..159 ({len(ep)} lit): proxy median {np.median(ep):.4f} mean {ep.mean():.4f}; "
         f"learned median {np.median(el):.4f} mean {el.mean():.4f}; learned better on {wins}")
    assert len(ep) >= 20
    assert wins > 0.5 * len(ep)
    assert np.median(el) < np.median(ep) and el.mean() < ep.mean()
    for st in (learned, proxy, trunc, plain):
        st.close()


RANK_SEEDS = [101, 102, 106, 108, 111, 112]   # lit held-out scenes (32 boxes, 4 emissive)


def test_error_head_ranks_tiles_by_their_true_error(rc):
    """The tile fallback trusts output 48: per 8x8 tile of level-2 probes, the
    mean predicted squared error.  On lit held-out 1080p scenes its ranking of
    the tiles must follow the tiles' true error (mean squared difference of
    the network's level 2 from the plain level 2, in the network's output
    units): Spearman rank correlation >= 0.3 on every scene and >= 0.6 in
    median (measured 0.50-0.94, median 0.83), and the top quarter of tiles by
    estimate shares at least 40 % with the top quarter by true error (25 %
    by chance)."""
    from scipy.stats import spearmanr
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    plain = rc.RadianceCascadeStage(W, H, levels=L)
    learned = rc.RadianceCascadeStage(W, H, levels=L)
    learned.load_weights()
    learned.set_surrogate(True)
    scale = learned.mlp_get()["out_scale"]
    sh = learned.surrogate_shapes()
    (qy, qx), (ty, tx) = sh["level2"], sh["tiles"]
    rhos, overlaps = [], []
    for seed in RANK_SEEDS:
        rc.gbuffer_raster(rc.scene_generate(seed, 32, 4), W, H, depth, None, emis)
        out = torch.empty((3, H, W), device=dev)
        a2p = torch.empty(plain.level_shape(2), device=dev)
        a2l = torch.empty_like(a2p)
        terr = torch.empty((ty, tx), device=dev)
        plain.run_device(depth, emis, out)
        plain.debug_level(2, a2p)
        learned.run_device(depth, emis, out)
        learned.debug_level(2, a2l)
        learned.debug_surrogate(tile_err=terr)
        torch.cuda.synchronize()
        assert float(out.norm()) > 0.0, f"seed {seed} is not lit"
        mse = (((a2l - a2p)[..., :3] / scale) ** 2).mean(dim=(2, 3)).cpu().numpy()   # [qy][qx]
        true = np.zeros((ty, tx))
        for j in range(ty):
            for i in range(tx):
                true[j, i] = mse[8 * j:8 * j + 8, 8 * i:8 * i + 8].mean()
        est = terr.cpu().numpy().astype(np.float64)
        rho = spearmanr(est.ravel(), true.ravel()).correlation
        k = est.size // 4
        top_e = set(np.argsort(-est.ravel())[:k].tolist())
        top_t = set(np.argsort(-true.ravel())[:k].tolist())
        rhos.append(rho)
        overlaps.append(len(top_e & top_t) / k)
        _say(f"  seed {seed}: Spearman {rho:.3f}, top-quartile overlap {overlaps[-1]:.2f}")
    assert min(rhos) >= 0.3 and np.median(rhos) >= 0.6
    assert np.median(overlaps) >= 0.4
    plain.close()
    learned.close()


SWEEP_SEEDS = [0, 101, 102, 106, 108, 111, 112]       # the benchmark scene, then lit held-out scenes
SWEEP_THRESHOLDS = [float("inf"), 2.0, 1.0, 0.5, 0.25, 0.1, -1.0]


def _fps(st, depth, emis, out, frames=20):
    for _ in range(5):
        st.run_device(depth, emis, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(frames):
        st.run_device(depth, emis, out)
    e1.record()
    torch.cuda.synchronize()
    return 1000.0 * frames / e0.elapsed_time(e1)


def test_fallback_threshold_sweep(rc):
    """The learned frame with the device-side tile fallback at a range of
    thresholds (mean predicted squared error per tile, network units): frames/s,
    the ratio to the plain frame, the fraction of tiles flagged and the
    irradiance rel-L2 against the FP64 oracle, on the benchmark scene and on
    lit held-out scenes.  The table is printed (the design notes quote it).
    Checked: every tile flagged gives the plain frame bit for bit; fewer tiles
    are flagged as the threshold rises; the fallback lowers the error."""
    from oracle import rco
    rco.set_threads(16)
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    plain = rc.RadianceCascadeStage(W, H, levels=L)
    learned = rc.RadianceCascadeStage(W, H, levels=L)
    learned.load_weights()
    rows = {}
    for seed in SWEEP_SEEDS:
        rc.gbuffer_raster(rc.scene_generate(seed, 32, 4), W, H, depth, None, emis)
        ref = torch.from_numpy(rco.run(depth.cpu().numpy(), emis.cpu().numpy(), L)).to(dev)
        out_p = torch.empty((3, H, W), device=dev)
        fps_p = _fps(plain, depth, emis, out_p)
        e_p = plain.error_norms(out_p, ref)["rel_l2"]
        _say(f"  seed {seed}: plain {fps_p:.0f} frames/s, rel-L2 vs FP64 {e_p:.2e}")
        prev_flag = None
        for thr in SWEEP_THRESHOLDS:
            learned.set_surrogate(True, threshold=thr)
            out = torch.empty((3, H, W), device=dev)
            fps = _fps(learned, depth, emis, out)
            learned.reset_stats()
            learned.run_device(depth, emis, out)
            torch.cuda.synchronize()
            s = learned.stats()
            frac = s.flagged_tiles / max(s.tiles_per_frame, 1)
            err = learned.error_norms(out, ref)["rel_l2"]
            rows.setdefault(thr, []).append((fps / fps_p, err, frac))
            _say(f"    threshold {thr:>5}: {fps:7.0f} frames/s = {fps / fps_p:4.2f}x plain, "
                 f"{100 * frac:5.1f}% tiles flagged, rel-L2 vs FP64 {err:.4f}")
            if thr < 0:
                assert torch.equal(out, out_p)
            if prev_flag is not None:
                assert frac >= prev_flag - 1e-12
            prev_flag = frac
    _say("  median over the scenes (threshold: ratio, rel-L2, flagged):")
    for thr in SWEEP_THRESHOLDS:
        r = np.array(rows[thr])
        _say(f"    {thr:>5}: {np.median(r[:, 0]):4.2f}x  {np.median(r[:, 1]):.4f}  {100 * np.median(r[:, 2]):5.1f}%")
    assert np.median(np.array(rows[0.25])[:, 1]) < np.median(np.array(rows[float("inf")])[:, 1])
    plain.close()
    learned.close()
