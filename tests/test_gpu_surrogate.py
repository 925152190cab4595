"""GPU parity of the learned far field (HIP kernels through the C ABI) against
the FP64 oracle (oracle/surrogate.py, oracle/mlp.py).

Bit-exact: the near distance field (integer d^2, the same sqrt rounding) and
the fallback path with every tile flagged (it is the plain path).  Within
tolerance: the mip pyramid (fp32 means rounded to fp16 per level, as the
oracle emulates), the cone trace
(fp32 bilinear, exp2/log2), the network (bf16 operands; the GPU accumulates
in fp32 in another order, which can move a hidden activation across a bf16
rounding boundary), its gradient and one Adam step, and the whole frame.
"""
import math
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

# Bounds about 10x the errors measured on MI355X (in parentheses).
MIP_RTOL = 1e-5
CONE_REL_L2 = 5e-6         # cone radiance (4.1e-7 at 1080p)
CONE_T_MAX = 1e-4          # cone transmittance, max-abs (1.4e-5)
MLP_REL_L2 = 1.5e-3        # forward outputs vs the bf16-emulating FP64 network (1.6e-4)
GRAD_REL_L2 = 1e-6         # per layer weight and bias gradient (9e-8)
FRAME_REL_L2 = 3e-5        # learned frame: merged level 1 and irradiance vs its FP64 oracle (2.9e-6)
# the fallback's composition also carries the plain path's fp32-vs-FP64
# error on the flagged tiles (8.7e-5 for level 1 at 1080p), bounded like the
# plain path's own (1e-3)
FALLBACK_REL_L2 = 1e-3


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / den if den > 0 else np.linalg.norm(x - ref)


@pytest.fixture(scope="module")
def rc():
    import radiancecascade_amd as rc
    if rc.device_count() < 1:
        pytest.skip("no HIP device")
    return rc


def _net(seed=3, scale=0.5):
    from oracle import mlp
    from oracle import surrogate as so
    rng = np.random.default_rng(seed)
    params = so.layers_to_params(mlp.init_params(seed)).astype(np.float32)
    mean = rng.normal(0.2, 0.1, mlp.IN).astype(np.float32)
    istd = rng.uniform(0.5, 2.0, mlp.IN).astype(np.float32)
    return params, mean, istd, np.float32(scale)


def _gbuffer(rc, W, H, seed, nb, ne):
    dev = torch.device("cuda:0")
    boxes = rc.scene_generate(seed, nb, ne)
    depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    emis = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    rc.gbuffer_raster(boxes, W, H, depth, None, emis)
    torch.cuda.synchronize()
    return depth, emis


CASES = [(256, 256, 5, 9, 8, 2), (200, 136, 4, 9, 8, 2), (1920, 1080, 6, 0, 32, 4)]


@pytest.mark.parametrize("W,H,L,seed,nb,ne", CASES)
def test_surrogate_frame_parity(rc, W, H, L, seed, nb, ne):
    from oracle import surrogate as so
    _say(f"[surrogate frame] {W}x{H} L={L} seed={seed}")
    depth, emis = _gbuffer(rc, W, H, seed, nb, ne)
    params, mean, istd, scale = _net()
    irr = torch.empty((3, H, W), dtype=torch.float32, device="cuda:0")
    with rc.RadianceCascadeStage(W, H, levels=L, debug=True) as st:
        st.mlp_set(params, mean, istd, scale)
        st.set_surrogate(True)
        st.run_device(depth, emis, irr)
        sh = st.surrogate_shapes()
        (qy, qx), (py, px), (ty, tx) = sh["level2"], sh["level1"], sh["tiles"]
        cone = torch.empty((qy, qx, 16, 4), dtype=torch.float32, device="cuda:0")
        a1 = torch.empty((py, px, 4, 4), dtype=torch.float32, device="cuda:0")
        terr = torch.empty((ty, tx), dtype=torch.float32, device="cuda:0")
        sdf = torch.empty((H, W), dtype=torch.float32, device="cuda:0")
        st.debug_surrogate(cone, a1, terr)
        st.debug_sdf(sdf)
        codes = []
        for lvl in (0, 1):
            h = torch.empty(rc.level_rays(W, H, lvl), dtype=torch.int32, device="cuda:0")
            st.debug_march(lvl, h)
            codes.append(h)
        with pytest.raises(rc.RcError):   # the learned frame marches no level >= 2
            st.debug_march(2, torch.empty(rc.level_rays(W, H, 2), dtype=torch.int32, device="cuda:0"))
        mips = []
        top = so.top_level(W, H)
        for m in range(1, top + 1):
            w, h = -(-W // (1 << m)), -(-H // (1 << m))
            e = torch.empty((h, w, 4), dtype=torch.float32, device="cuda:0")
            o = torch.empty((h, w, 4), dtype=torch.float32, device="cuda:0")
            st.debug_mip(m, e, o)
            mips.append((e, o))
        torch.cuda.synchronize()
        s = st.stats()
    d = depth.cpu().numpy()
    e = emis.cpu().numpy()
    ref_irr, ref = so.run(d, e, L, params, mean, istd, float(scale), want=True)
    # near distance field: bit-exact
    assert np.array_equal(sdf.cpu().numpy(), ref["sdf"].astype(np.float32))
    # levels 0 and 1 marched in place on it: every ray code bit-exact
    from oracle import rco
    for lvl, h in enumerate(codes):
        _, hit_ref = rco.march_level(ref["sdf"].astype(np.float32), e, lvl)
        assert np.array_equal(h.cpu().numpy(), hit_ref.ravel()), f"learned frame, level {lvl} ray codes"

    for m, (eg, og) in enumerate(mips, start=1):
        re, ro = ref["levels"][m]
        assert np.allclose(eg.cpu().numpy(), re, rtol=MIP_RTOL, atol=1e-6), m
        assert np.allclose(og.cpu().numpy(), ro, rtol=MIP_RTOL, atol=1e-6), m
    cg = cone.cpu().numpy()
    r_cone = rel_l2(cg[..., :3], ref["cone"][..., :3])
    t_cone = np.abs(cg[..., 3] - ref["cone"][..., 3]).max()
    r_a1 = rel_l2(a1.cpu().numpy(), ref["A1"])
    r_irr = rel_l2(irr.cpu().numpy(), ref_irr)
    _say(f"  cone rel-L2 {r_cone:.2e}  T max {t_cone:.2e}  level1 rel-L2 {r_a1:.2e}  irradiance rel-L2 {r_irr:.2e}")
    assert r_cone <= CONE_REL_L2 and t_cone <= CONE_T_MAX
    assert r_a1 <= FRAME_REL_L2 and r_irr <= FRAME_REL_L2
    assert np.isfinite(terr.cpu().numpy()).all()
    assert s.surrogate_frames == 1 and s.fallback_frames == 0 and s.tiles_per_frame == ty * tx


def test_mip_pyramid_of_a_wide_frame(rc):
    """4096x1200 (13 mip levels, level 6 with 1216 cells, a 19-row last tile
    row): every level against the oracle, on two frames in a row."""
    from oracle import rco
    from oracle import surrogate as so
    W, H, L = 4096, 1200, 7
    _say(f"[mip pyramid] {W}x{H} L={L}")
    depth, emis = _gbuffer(rc, W, H, 5, 32, 4)
    params, mean, istd, scale = _net()
    irr = torch.empty((3, H, W), dtype=torch.float32, device="cuda:0")
    d, e = depth.cpu().numpy(), emis.cpu().numpy()
    ref = so.mip(rco.seed_init(d, e), e)
    with rc.RadianceCascadeStage(W, H, levels=L) as st:
        st.mlp_set(params, mean, istd, scale)
        st.set_surrogate(True)
        for frame in range(2):
            st.run_device(depth, emis, irr)
            for m in range(1, so.top_level(W, H) + 1):
                w, h = -(-W // (1 << m)), -(-H // (1 << m))
                eg = torch.empty((h, w, 4), dtype=torch.float32, device="cuda:0")
                og = torch.empty((h, w, 4), dtype=torch.float32, device="cuda:0")
                st.debug_mip(m, eg, og)
                torch.cuda.synchronize()
                re, ro = ref[m]
                assert np.allclose(eg.cpu().numpy(), re, rtol=MIP_RTOL, atol=1e-6), (frame, m)
                assert np.allclose(og.cpu().numpy(), ro, rtol=MIP_RTOL, atol=1e-6), (frame, m)


def _rand_io(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.gamma(1.0, 0.5, (n, 64)).astype(np.float32)      # log1p features and T are >= 0
    t = rng.normal(0.0, 1.0, (n, 48)).astype(np.float32)
    return x, t


def test_mlp_forward_parity(rc):
    from oracle import mlp
    from oracle import surrogate as so
    params, mean, istd, scale = _net(5)
    x, _ = _rand_io(1000, 1)
    with rc.RadianceCascadeStage(64, 64, levels=3) as st:
        st.mlp_set(params, mean, istd, scale)
        xd = torch.from_numpy(x).cuda()
        out = torch.empty((1000, 64), dtype=torch.float32, device="cuda:0")
        st.mlp_forward(xd, out)
        torch.cuda.synchronize()
    z, _ = mlp.forward(so.params_to_layers(params), (x.astype(np.float64) - mean) * istd, emulate_bf16=True)
    r = rel_l2(out.cpu().numpy()[:, :49], z[:, :49])
    _say(f"  forward rel-L2 {r:.2e}")
    assert r <= MLP_REL_L2
    assert np.all(out.cpu().numpy()[:, 49:] == 0)


def test_mlp_gradient_and_adam_parity(rc):
    from oracle import mlp
    from oracle import surrogate as so
    params, mean, istd, scale = _net(7)
    n = 777                                                  # ragged: 24 full blocks + 9 rows
    x, t = _rand_io(n, 2)
    layers = so.params_to_layers(params)
    xn = (x.astype(np.float64) - mean) * istd
    z, cache = mlp.forward(layers, xn, emulate_bf16=True)
    loss, g4 = mlp.loss_and_grad_out(z, t.astype(np.float64))
    grads = mlp.backward(layers, cache, g4, emulate_bf16=True)
    with rc.RadianceCascadeStage(64, 64, levels=3) as st:
        st.mlp_set(params, mean, istd, scale)
        xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
        grad = torch.empty(rc._abi.MLP_PARAMS, dtype=torch.float32, device="cuda:0")
        lo = torch.empty(1, dtype=torch.float32, device="cuda:0")
        st.mlp_gradient(xd, td, grad, lo)
        g = grad.cpu().numpy()
        st.mlp_adam(grad, 1e-3)
        after = st.mlp_get()["params"]
    gl = so.params_to_layers(g)
    r_loss = abs(float(lo.item()) - loss) / abs(loss)
    _say(f"  loss {lo.item():.6f} vs {loss:.6f} (rel {r_loss:.2e})")
    assert r_loss <= 1e-7          # (9.9e-9)
    for l in range(4):
        rw = rel_l2(gl[l][0], grads[l][0])
        rb = rel_l2(gl[l][1], grads[l][1])
        _say(f"  layer {l}: dW rel-L2 {rw:.2e}  db rel-L2 {rb:.2e}")
        assert rw <= GRAD_REL_L2 and rb <= GRAD_REL_L2
    # one Adam step from the GPU's own gradient (first step: p - lr * g / (|g| + eps))
    opt = mlp.Adam(so.params_to_layers(params), lr=1e-3)
    ref = so.layers_to_params(opt.step(so.params_to_layers(params), gl))
    assert np.allclose(after, ref, rtol=1e-6, atol=1e-7)


def test_training_rows_parity(rc):
    from oracle import rco
    from oracle import surrogate as so
    W, H, L = 160, 128, 5
    depth, emis = _gbuffer(rc, W, H, 9, 8, 2)
    px, py = rc.probe_grid(W, H, 2)
    feats = torch.empty((px * py, 64), dtype=torch.float32, device="cuda:0")
    targ = torch.empty((px * py, 48), dtype=torch.float32, device="cuda:0")
    with rc.RadianceCascadeStage(W, H, levels=L) as st:
        st.surrogate_samples(depth, emis, feats, targ)
        torch.cuda.synchronize()
    d, e = depth.cpu().numpy(), emis.cpu().numpy()
    _, lv = rco.run(d, e, L, want_levels=True)
    s0 = rco.seed_init(d, e)
    cone = so.cone_trace(so.mip(s0, e), W, H)
    x, t = so.samples(cone, lv[2])
    rx, rt = rel_l2(feats.cpu().numpy(), x), rel_l2(targ.cpu().numpy(), t)
    _say(f"  rows: features rel-L2 {rx:.2e}  targets rel-L2 {rt:.2e}")
    assert rx <= 1e-6 and rt <= 1e-5    # (8.6e-8, 7.6e-7)


def test_fallback_everywhere_is_the_plain_path(rc):
    W, H, L = 256, 192, 5
    depth, emis = _gbuffer(rc, W, H, 9, 8, 2)
    params, mean, istd, scale = _net()
    plain = torch.empty((3, H, W), dtype=torch.float32, device="cuda:0")
    surr = torch.empty((3, H, W), dtype=torch.float32, device="cuda:0")
    with rc.RadianceCascadeStage(W, H, levels=L) as st:
        st.run_device(depth, emis, plain)
        st.mlp_set(params, mean, istd, scale)
        st.set_surrogate(True, threshold=-1.0)        # every tile is above the threshold
        st.run_device(depth, emis, surr)
        torch.cuda.synchronize()
        s = st.stats()
    assert s.fallback_frames == 1 and s.flagged_tiles == s.tiles_per_frame > 0
    assert torch.equal(plain, surr)


@pytest.mark.parametrize("W,H,L,seed,nb,ne", [(512, 384, 5, 1, 32, 4), (1920, 1080, 6, 0, 32, 4)])
def test_fallback_of_some_tiles(rc, W, H, L, seed, nb, ne):
    """A finite threshold at the median tile estimate flags about half the
    tiles, decided on the device.  The flagged tiles' level 1 equals the plain
    frame's bit for bit (plain far field on the jump-flood field, masked to the
    probes they need), the others equal the learned frame's bit for bit, and
    the irradiance equals the oracle's composition (plain FP64 level 1 on the
    flagged tiles, the network's FP64 level 1 elsewhere, level 0 on the near
    field) within the frame tolerance."""
    from oracle import rco
    from oracle import surrogate as so
    _say(f"[partial fallback] {W}x{H} L={L} seed={seed}")
    depth, emis = _gbuffer(rc, W, H, seed, nb, ne)
    params, mean, istd, scale = _net()
    dev = "cuda:0"
    irr_p, irr_l, irr_f = (torch.empty((3, H, W), dtype=torch.float32, device=dev) for _ in range(3))
    with rc.RadianceCascadeStage(W, H, levels=L) as plain, rc.RadianceCascadeStage(W, H, levels=L) as st:
        a1_p = torch.empty(plain.level_shape(1), device=dev)
        a1_l, a1_f = torch.empty_like(a1_p), torch.empty_like(a1_p)
        plain.run_device(depth, emis, irr_p)
        plain.debug_level(1, a1_p)
        st.mlp_set(params, mean, istd, scale)
        st.set_surrogate(True)
        st.run_device(depth, emis, irr_l)
        st.debug_level(1, a1_l)
        (ty, tx) = st.surrogate_shapes()["tiles"]
        terr = torch.empty((ty, tx), device=dev)
        st.debug_surrogate(tile_err=terr)
        torch.cuda.synchronize()
        est = terr.cpu().numpy()
        thr = float(np.median(est))
        st.reset_stats()
        st.set_surrogate(True, threshold=thr)
        st.run_device(depth, emis, irr_f)
        st.debug_level(1, a1_f)
        torch.cuda.synchronize()
        s = st.stats()
    flagged = ~(est <= thr)
    n_flag = int(flagged.sum())
    _say(f"  {n_flag} of {flagged.size} tiles flagged")
    assert 0 < n_flag < flagged.size
    assert s.fallback_frames == 1 and s.flagged_tiles == n_flag
    py, px = a1_p.shape[:2]
    probe_flag = np.repeat(np.repeat(flagged, 16, axis=0), 16, axis=1)[:py, :px]
    f, p, l = a1_f.cpu().numpy(), a1_p.cpu().numpy(), a1_l.cpu().numpy()
    assert np.array_equal(f[probe_flag].view(np.uint32), p[probe_flag].view(np.uint32))
    assert np.array_equal(f[~probe_flag].view(np.uint32), l[~probe_flag].view(np.uint32))
    d, e = depth.cpu().numpy(), emis.cpu().numpy()
    _, ref = so.run(d, e, L, params, mean, istd, float(scale), want=True)
    _, lv64 = rco.run(d, e, L, np.float64, want_levels=True)
    a1_mix = np.where(probe_flag[:, :, None, None], lv64[1], ref["A1"])
    irr_ref = rco.gather(W, H, rco.merge_level(W, H, 0, ref["m0"], a1_mix))
    r_a1, r_irr = rel_l2(f, a1_mix), rel_l2(irr_f.cpu().numpy(), irr_ref)
    _say(f"  level 1 rel-L2 {r_a1:.2e}, irradiance rel-L2 {r_irr:.2e} vs the oracle's composition")
    assert r_a1 <= FALLBACK_REL_L2 and r_irr <= FALLBACK_REL_L2


def test_surrogate_graph_replay_matches_direct_launches(rc):
    W, H, L = 240, 160, 5
    depth, emis = _gbuffer(rc, W, H, 9, 8, 2)
    params, mean, istd, scale = _net()
    outs = []
    for graph in (True, False):
        with rc.RadianceCascadeStage(W, H, levels=L, graph=graph) as st:
            st.mlp_set(params, mean, istd, scale)
            st.set_surrogate(True)
            for _ in range(3):
                irr = torch.empty((3, H, W), dtype=torch.float32, device="cuda:0")
                st.run_device(depth, emis, irr)
                torch.cuda.synchronize()
                outs.append(irr.clone())
    assert all(torch.equal(outs[0], o) for o in outs[1:])


def test_autotuned_learned_frame_is_bit_identical(rc):
    """Tuning the learned frame's level-1 tiling changes the launch shape only."""
    W, H, L = 240, 160, 5
    _say(f"[surrogate autotune] {W}x{H} L={L}")
    depth, emis = _gbuffer(rc, W, H, 9, 8, 2)
    params, mean, istd, scale = _net()
    with rc.RadianceCascadeStage(W, H, levels=L) as st:
        st.mlp_set(params, mean, istd, scale)
        st.set_surrogate(True)
        before = torch.empty((3, H, W), dtype=torch.float32, device="cuda:0")
        st.run_device(depth, emis, before)
        scratch = torch.empty_like(before)
        st.autotune(depth, emis, scratch, iters=1)
        after = torch.empty_like(before)
        st.run_device(depth, emis, after)
        torch.cuda.synchronize()
    assert torch.equal(before, after)


def test_training_reduces_loss(rc):
    with rc.RadianceCascadeStage(64, 64, levels=3) as st:
        rep = st.train_surrogate(width=128, height=128, levels=5, scene_seed0=10000, scenes=24, epochs=4,
                                 batch=512, lr=2e-3, lr_decay=0.8, seed=1)
        _say(f"  losses {rep['epoch_loss']}  rows {rep['samples']}  scale {rep['out_scale']:.4f}")
        assert rep["samples"] == 24 * 16 * 16 and rep["steps"] == 4 * (24 * 256 // 512)
        assert all(math.isfinite(v) for v in rep["epoch_loss"])
        assert rep["epoch_loss"][-1] < rep["epoch_loss"][0]
        blob = st.weights_blob()
    with rc.RadianceCascadeStage(64, 64, levels=3) as st2:
        st2.load_weights_blob(blob)
        assert st2.weights_blob() == blob


def test_training_on_dumped_gbuffers_matches_seeded_training(rc, tmp_path):
    """Rows from G-buffers read back from RCDUMP01 files train the same
    network, bit for bit, as the seeded trainer on the same scenes."""
    from radiancecascade_amd import train as tr
    W, H, L, seeds = 256, 192, 5, [500, 501, 502, 503]
    print(f"[train rows] {W}x{H} L={L} seeds {seeds}", flush=True)
    kw = dict(epochs=2, batch=1024, lr=2e-3, lr_decay=0.7, seed=7)
    with rc.RadianceCascadeStage(64, 64, levels=3) as st:
        ra = st.train_surrogate(width=W, height=H, levels=L, scene_seed0=seeds[0], scenes=len(seeds), n_boxes=32,
                                n_emissive=4, **kw)
        ma = st.mlp_get()
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    paths = []
    for s in seeds:
        rc.gbuffer_raster(rc.scene_generate(s, 32, 4), W, H, depth, None, emis)
        p = str(tmp_path / f"scene{s}.rcdump")
        rc.dump_write(p, {"depth": depth.cpu().numpy(), "emissive": emis.cpu().numpy()})
        paths.append(p)
    feats, targets = tr.rows_from_dumps(paths, L)
    qx, qy = rc.probe_grid(W, H, 2)
    assert feats.shape == (len(seeds) * qx * qy, rc._abi.MLP_INPUTS)
    with rc.RadianceCascadeStage(64, 64, levels=3) as st:
        rb = st.train_surrogate_rows(feats, targets, frames=len(seeds), **kw)
        mb = st.mlp_get()
    assert rb["samples"] == ra["samples"] and rb["steps"] == ra["steps"]
    assert rb["epoch_loss"] == ra["epoch_loss"] and rb["out_scale"] == ra["out_scale"]
    for k in ("params", "in_mean", "in_istd"):
        assert np.array_equal(ma[k], mb[k]), k


def test_train_rows_rejects_bad_input(rc):
    n = 512
    feats = torch.zeros((n, rc._abi.MLP_INPUTS), device="cuda:0")
    targets = torch.zeros((n, rc._abi.MLP_RADIANCE), device="cuda:0")
    with rc.RadianceCascadeStage(64, 64, levels=3) as st:
        with pytest.raises(rc.RcError):
            st.train_surrogate_rows(feats, targets, batch=1024)      # fewer rows than one batch
        with pytest.raises(rc.RcError):
            st.train_surrogate_rows(feats, targets, batch=100)       # not a multiple of 32
        with pytest.raises(ValueError):
            st.train_surrogate_rows(feats, targets[:256], batch=256)


def test_train_cli_on_dump_files(rc, tmp_path):
    """`python -m radiancecascade_amd.train --dumps ...` end to end: the blob
    loads, and its sidecar names the dump files with their SHA-256."""
    import json

    from radiancecascade_amd import train as tr
    W, H = 256, 192
    _say(f"[train cli] {W}x{H} two dumps")
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    paths = []
    for s in (600, 601):
        rc.gbuffer_raster(rc.scene_generate(s, 32, 4), W, H, depth, None, emis)
        p = str(tmp_path / f"f{s}.rcdump")
        rc.dump_write(p, {"depth": depth.cpu().numpy(), "emissive": emis.cpu().numpy()})
        paths.append(p)
    out = str(tmp_path / "net.bin")
    assert tr.main(["--dumps", *paths, "--levels", "5", "--epochs", "2", "--batch", "512", "--out", out]) == 0
    meta = json.load(open(out + ".json"))
    assert [d["file"] for d in meta["train"]["dumps"]] == ["f600.rcdump", "f601.rcdump"]
    assert meta["train"]["dumps"][0]["sha256"] == rc.sha256(open(paths[0], "rb").read())
    assert len(meta["report"]["epoch_loss"]) == 2
    with rc.RadianceCascadeStage(W, H, levels=5) as st:
        st.load_weights(out)
        st.set_surrogate(True)
        irr = torch.empty((3, H, W), device=dev)
        st.run_device(depth, emis, irr)
        torch.cuda.synchronize()
        assert torch.isfinite(irr).all()
