"""Batches of views (rc_batch_*): one graph launch over views of uneven
sizes must give, for every view, exactly the bits the single-view pass gives
on the same G-buffer -- before and after autotuning, and on graph replay."""
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


def test_batch_matches_single_views(rc):
    dev = torch.device("cuda:0")
    views, refs = [], []
    for j, (W, H) in enumerate(SIZES):
        _say(f"[batch] view {j}: {W}x{H}")
        d = torch.empty((H, W), device=dev)
        e = torch.empty((3, H, W), device=dev)
        rc.gbuffer_raster(rc.scene_generate(9 + j, 12, 3), W, H, d, None, e)
        o = torch.full((3, H, W), -1.0, device=dev)
        views.append((d, e, o))
        ref = torch.empty((3, H, W), device=dev)
        with rc.RadianceCascadeStage(W, H, levels=0) as st:
            st.run_device(d, e, ref)
        refs.append(ref)
    torch.cuda.synchronize()
    with rc.RadianceCascadeBatch(views, streams=3) as b:
        for rnd in range(3):
            if rnd == 2:
                b.autotune(iters=1)
                for _, _, o in views:
                    o.fill_(-1.0)
            b.run()
            torch.cuda.synchronize()
            for j, ((_, _, o), ref) in enumerate(zip(views, refs)):
                assert torch.equal(o, ref), (rnd, j, SIZES[j])
        assert [b.levels(j) for j in range(len(SIZES))] == [rc.levels_for_coverage(W, H) for W, H in SIZES]


def test_batch_rejects_bad_arguments(rc):
    with pytest.raises(ValueError):
        rc.RadianceCascadeBatch([(None, None, None)])
