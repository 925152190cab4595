"""Batches of views (rc_batch_*): one graph launch over views of uneven
sizes must give, for every view, exactly the bits the single-view pass gives
on the same G-buffer -- before and after autotuning, and on graph replay --
in both batch modes: the work list (one launch per stage over every view,
one shared workspace) and one context per view on captured streams."""
import sys

import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (97, 61), (256, 128), (200, 300), (128, 128), (300, 1), (1, 1), (512, 256)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rc():
    import radiancecascade_amd as rc
    if rc.device_count() < 1:
        pytest.skip("no HIP device")
    return rc


def _views(rc, sizes, levels=None, seed0=9):
    dev = torch.device("cuda:0")
    views = []
    for j, (W, H) in enumerate(sizes):
        d = torch.empty((H, W), device=dev)
        e = torch.empty((3, H, W), device=dev)
        rc.gbuffer_raster(rc.scene_generate(seed0 + j, 12, 3), W, H, d, None, e)
        views.append((d, e, torch.full((3, H, W), -1.0, device=dev)))
    return views


def _single(rc, views, levels, storage="f32"):
    refs = []
    for (d, e, _), L in zip(views, levels):
        H, W = d.shape
        ref = torch.empty((3, H, W), device=d.device)
        with rc.RadianceCascadeStage(W, H, levels=L, storage=storage) as st:
            st.run_device(d, e, ref)
        refs.append(ref)
    torch.cuda.synchronize()
    return refs


@pytest.mark.parametrize("mode", ["worklist", "streams"])
def test_batch_matches_single_views(rc, mode):
    _say(f"[batch] {mode}: {len(SIZES)} views {SIZES}")
    views = _views(rc, SIZES)
    refs = _single(rc, views, [0] * len(SIZES))
    with rc.RadianceCascadeBatch(views, mode=mode, streams=3) as b:
        for rnd in range(3):
            if rnd == 2:
                _say(f"[batch] {mode}: autotune")
                b.autotune(iters=1)
                for _, _, o in views:
                    o.fill_(-1.0)
            b.run()
            torch.cuda.synchronize()
            for j, ((_, _, o), ref) in enumerate(zip(views, refs)):
                assert torch.equal(o, ref), (mode, rnd, j, SIZES[j])
        assert [b.levels(j) for j in range(len(SIZES))] == [rc.levels_for_coverage(W, H) for W, H in SIZES]
        info = b.info()
        if mode == "worklist":
            # distance transform (3) + march + levels below the merged top + level 0
            assert 6 <= info["kernel_nodes"] <= 12, info
            assert info["workspace_bytes"] > 0
        else:
            # per view at least the distance transform (3) and level 0
            assert info["kernel_nodes"] >= 4 * len(SIZES), info


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_worklist_levels_and_storage(rc, storage):
    """Explicit level counts from 1 (no level >= 1: no march, level 0 only)
    to 8 beside coverage-rule views, in half-precision storage: each view
    equals its single-view frame, replayed twice."""
    sizes = [(160, 90), (64, 64), (333, 77), (1, 1), (48, 200), (257, 129)]
    levels = [1, 2, 8, 3, 0, 5]
    _say(f"[batch] worklist {storage}: {list(zip(sizes, levels))}")
    views = _views(rc, sizes, seed0=40)
    refs = _single(rc, views, levels, storage)
    with rc.RadianceCascadeBatch(views, levels=levels, storage=storage) as b:
        for rnd in range(2):
            for _, _, o in views:
                o.fill_(-1.0)
            b.run()
            torch.cuda.synchronize()
            for j, ((_, _, o), ref) in enumerate(zip(views, refs)):
                assert torch.equal(o, ref), (storage, rnd, j, sizes[j], levels[j])


def test_batch_rejects_bad_arguments(rc):
    with pytest.raises(ValueError):
        rc.RadianceCascadeBatch([(None, None, None)])


@pytest.mark.parametrize("mode", ["worklist", "streams"])
def test_batch_workload_matches_single_views_and_oracle(rc, mode):
    """The 64-view batch workload (rc.batch_views_spec(1000, 64): seeds
    1000..1063, 16..64 boxes, W and H from {256, ..., 2048}, 5..7 levels by the
    coverage rule, 86 MP) in one graph launch, autotuned: every view equals its
    single-view frame bit for bit, and the three smallest views equal the
    FP32 oracle bit for bit."""
    import numpy as np
    from oracle import rco
    dev = torch.device("cuda:0")
    spec = rc.batch_views_spec(1000, 64)
    views, gb = [], []
    for seed, W, H, nb, ne in spec:
        d = torch.empty((H, W), device=dev)
        e = torch.empty((3, H, W), device=dev)
        rc.gbuffer_raster(rc.scene_generate(seed, nb, ne), W, H, d, None, e)
        views.append((d, e, torch.full((3, H, W), -1.0, device=dev)))
        gb.append((seed, W, H, nb, ne))
    torch.cuda.synchronize()
    _say(f"[batch workload] {mode}: {len(spec)} views, {sum(W * H for _, W, H, _, _ in spec) / 1e6:.1f} MP")
    with rc.RadianceCascadeBatch(views, mode=mode, streams=8) as b:
        b.autotune(iters=1)
        b.run()
        torch.cuda.synchronize()
        levels = [b.levels(j) for j in range(len(spec))]
    assert sorted(set(levels)) == [5, 6, 7]
    for j, ((d, e, o), (seed, W, H, nb, ne)) in enumerate(zip(views, gb)):
        ref = torch.empty((3, H, W), device=dev)
        with rc.RadianceCascadeStage(W, H, levels=levels[j]) as st:
            st.run_device(d, e, ref)
        torch.cuda.synchronize()
        assert torch.equal(o, ref), (j, seed, W, H)
    smallest = sorted(range(len(spec)), key=lambda j: spec[j][1] * spec[j][2])[:3]
    for j in smallest:
        seed, W, H, nb, ne = spec[j]
        d, e, o = views[j]
        ref32 = rco.run(d.cpu().numpy(), e.cpu().numpy(), levels[j], np.float32)
        _say(f"  view {j} (seed {seed}, {W}x{H}, L={levels[j]}) vs FP32 oracle")
        assert np.array_equal(o.cpu().numpy().view(np.uint32), ref32.view(np.uint32)), (j, W, H)
