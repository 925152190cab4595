"""The allocation-free entry point (rc_run_device_noalloc) through the C ABI:
its irradiance equals rc_run_device's bit for bit, on two calls in a row (the
second on a smaller frame, in the scratch the first left behind), and a size
beyond the static scratch is refused with a clear error."""
import ctypes
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


def _frame(rc, W, H, seed):
    dev = torch.device("cuda:0")
    depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    emis = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    rc.gbuffer_raster(rc.scene_generate(seed, 32, 4), W, H, depth, None, emis)
    return depth, emis


def test_noalloc_equals_rc_run_device(rc):
    from radiancecascade_amd import _abi
    lib = _abi.load()
    stream = torch.cuda.current_stream()
    free0 = None
    for W, H, L, seed in [(1920, 1080, 6, 0), (640, 360, 5, 1), (97, 61, 6, 3)]:
        _say(f"[noalloc] {W}x{H} L={L} seed={seed}")
        depth, emis = _frame(rc, W, H, seed)
        ref = torch.empty((3, H, W), dtype=torch.float32, device="cuda:0")
        with rc.RadianceCascadeStage(W, H, levels=L) as st:
            st.run_device(depth, emis, ref)
        torch.cuda.synchronize()
        out = torch.full((3, H, W), -1.0, dtype=torch.float32, device="cuda:0")
        g = _abi.GBuffer(depth.data_ptr(), 0, emis.data_ptr())
        assert lib.rc_noalloc_scratch_bytes(W, H, L) <= 64 << 20
        st = lib.rc_run_device_noalloc(W, H, L, ctypes.byref(g), out.data_ptr(), ctypes.c_void_p(stream.cuda_stream))
        assert st == _abi.OK, lib.rc_noalloc_last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (W, H, L)
        # no device memory taken by the calls after the first
        free = torch.cuda.mem_get_info()[0]
        if free0 is not None:
            assert free >= free0 - (2 << 20), (free0, free)
        free0 = free


def test_noalloc_refuses_sizes_beyond_its_scratch(rc):
    from radiancecascade_amd import _abi
    lib = _abi.load()
    depth, emis = _frame(rc, 64, 64, 1)
    g = _abi.GBuffer(depth.data_ptr(), 0, emis.data_ptr())
    out = torch.empty((3, 64, 64), dtype=torch.float32, device="cuda:0")
    assert lib.rc_noalloc_scratch_bytes(4096, 4096, 7) > 64 << 20
    st = lib.rc_run_device_noalloc(4096, 4096, 7, ctypes.byref(g), out.data_ptr(), None)
    assert st == _abi.ERR_INVALID
    assert b"scratch" in lib.rc_noalloc_last_error()
    assert lib.rc_run_device_noalloc(64, 64, 8, ctypes.byref(g), out.data_ptr(), None) == _abi.ERR_INVALID
