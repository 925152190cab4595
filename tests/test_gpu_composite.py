"""GPU composite (rc_composite through the C ABI) against oracle/composite.py:
bit-exact BGRA8 sRGB words, on random inputs that cross every threshold
(negatives, values past 1, NaN) at a ragged size, and on a real 1080p frame
(albedo and emissive from the raster, irradiance from the GI pass), with
and without a direct term."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rc():
    import radiancecascade_amd as rc
    if rc.device_count() < 1:
        pytest.skip("no HIP device")
    return rc


def _run(st, a, e, i, d):
    H, W = a.shape[1:]
    out = torch.empty((H, W), dtype=torch.int32, device="cuda:0")
    st.composite(a, e, i, out, direct=d)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)


def test_composite_random_inputs_bit_exact(rc):
    from oracle import composite as co
    W, H = 97, 61
    _say(f"[composite random] {W}x{H}")
    rng = np.random.default_rng(7)
    a = rng.uniform(-0.2, 1.2, (3, H, W)).astype(np.float32)
    e = rng.uniform(-0.1, 0.6, (3, H, W)).astype(np.float32)
    i = rng.uniform(0.0, 1.5, (3, H, W)).astype(np.float32)
    d = rng.uniform(-0.1, 0.5, (3, H, W)).astype(np.float32)
    i[0, 0, :5] = np.nan
    # values exactly on thresholds
    t = co.thresholds()
    e[1, 1, :20] = t[1:21]
    a[1, 1, :20] = 0.0
    dev = lambda x: torch.from_numpy(x).to("cuda:0")
    with rc.RadianceCascadeStage(W, H, levels=2) as st:
        for direct in (None, d):
            got = _run(st, dev(a), dev(e), dev(i), None if direct is None else dev(direct))
            ref = co.composite(a, e, i, direct)
            assert np.array_equal(got, ref)
    assert np.any((ref >> 8) & 0xFF == 0) and np.any((ref >> 8) & 0xFF == 255)


def test_composite_of_a_gi_frame(rc):
    from oracle import composite as co
    W, H, L = 1920, 1080, 6
    _say(f"[composite frame] {W}x{H} L={L}")
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    alb = torch.empty((3, H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    rc.gbuffer_raster(rc.scene_generate(0, 32, 4), W, H, depth, alb, emis)
    irr = torch.empty((3, H, W), device=dev)
    with rc.RadianceCascadeStage(W, H, levels=L) as st:
        st.run_device(depth, emis, irr)
        got = _run(st, alb, emis, irr, None)
    ref = co.composite(alb.cpu().numpy(), emis.cpu().numpy(), irr.cpu().numpy())
    assert np.array_equal(got, ref)
    assert np.all(got >> 24 == 255)
    # the frame is lit: not all black, not all saturated
    g = (got >> 8) & 0xFF
    assert 0 < g.mean() < 255
