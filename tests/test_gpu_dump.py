"""A captured frame through an RCDUMP01 file: the G-buffer and irradiance of a
GPU frame are written, read back, and the frame rerun from the read-back
G-buffer reproduces the stored irradiance bit for bit (the use as a
cross-machine fixture)."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rc():
    import radiancecascade_amd as rc
    if rc.device_count() < 1:
        pytest.skip("no HIP device")
    return rc


def test_frame_round_trip_through_a_dump(rc, tmp_path):
    W, H, L = 640, 360, 6
    print(f"[dump frame] {W}x{H} L={L}", flush=True)
    sys.stdout.flush()
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), device=dev)
    alb = torch.empty((3, H, W), device=dev)
    emis = torch.empty((3, H, W), device=dev)
    rc.gbuffer_raster(rc.scene_generate(3, 32, 4), W, H, depth, alb, emis)
    irr = torch.empty((3, H, W), device=dev)
    with rc.RadianceCascadeStage(W, H, levels=L) as st:
        st.run_device(depth, emis, irr)
        torch.cuda.synchronize()
        p = str(tmp_path / "frame.rcdump")
        rc.dump_write(p, {"depth": depth.cpu().numpy(), "albedo": alb.cpu().numpy(),
                          "emissive": emis.cpu().numpy(), "irradiance": irr.cpu().numpy()})
        back = rc.dump_read(p)
        d2 = torch.from_numpy(back["depth"][0]).to(dev)
        e2 = torch.from_numpy(back["emissive"]).to(dev)
        irr2 = torch.empty_like(irr)
        st.run_device(d2, e2, irr2)
        torch.cuda.synchronize()
    assert np.array_equal(back["irradiance"], irr.cpu().numpy())
    assert torch.equal(irr2, irr)
