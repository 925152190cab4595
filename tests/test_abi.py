"""CPU checks of the C ABI library: it loads, exports every symbol the header
declares, validates arguments before touching a device, agrees with the
oracle on the host-side integer work (scene PRNG, probe grids, coverage), and
reports a missing device as a status instead of crashing.  No kernel runs
here."""
import ctypes
import os
import re

import numpy as np
import pytest

import radiancecascade_amd as rc
from radiancecascade_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "rc_gi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = _header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.SIGNATURES), "ctypes table out of sync with the header"


def test_abi_version_and_status_strings():
    lib = _abi.load()
    assert lib.rc_abi_version() == _abi.ABI_VERSION
    for s in range(7):
        assert lib.rc_status_string(s)


def test_create_argument_validation():
    lib = _abi.load()
    ctx = ctypes.c_void_p()
    assert lib.rc_create(None, ctypes.byref(ctx)) == _abi.ERR_NULL
    bad = [dict(version=99), dict(version=1), dict(width=0), dict(height=_abi.MAX_DIM + 1), dict(levels=9),
           dict(storage=3), dict(flags=0x80), dict(depth_mode=2),
           dict(depth_mode=_abi.DEPTH_NDC, z_near=0.0, z_far=100.0),
           dict(depth_mode=_abi.DEPTH_NDC, z_near=1.0, z_far=1.0),
           dict(depth_mode=_abi.DEPTH_NDC, z_near=0.1, z_far=float("inf")),
           dict(depth_mode=_abi.DEPTH_NDC, z_near=float("nan"), z_far=100.0)]
    for b in bad:
        kw = dict(version=_abi.ABI_VERSION, width=64, height=64, levels=3, storage=0, flags=0, device=0,
                  depth_mode=0, z_near=0.0, z_far=0.0)
        kw.update(b)
        cfg = _abi.Config(**kw)
        assert lib.rc_create(ctypes.byref(cfg), ctypes.byref(ctx)) == _abi.ERR_INVALID, b
        assert lib.rc_last_error(None)


def test_null_context_calls():
    lib = _abi.load()
    assert lib.rc_resize(None, 8, 8, 0) == _abi.ERR_NULL
    assert lib.rc_render_host(None, None) == _abi.ERR_NULL
    assert lib.rc_run_device(None, None, None, None) == _abi.ERR_NULL
    assert lib.rc_autotune(None, None, None, 1, None) == _abi.ERR_NULL
    lib.rc_destroy(None)  # no-op
    assert lib.rc_host_alloc(8, None) == _abi.ERR_NULL
    ptr = ctypes.c_void_p(1)
    assert lib.rc_host_alloc(0, ctypes.byref(ptr)) == _abi.ERR_INVALID and not ptr.value
    lib.rc_host_free(None)  # no-op


@pytest.mark.skipif(rc.device_count() > 0, reason="a HIP device is present")
def test_no_device_is_a_status():
    with pytest.raises(rc.RcError) as ei:
        rc.RadianceCascadeStage(64, 64)
    assert ei.value.status == _abi.ERR_NO_DEVICE
    n = ctypes.c_int(7)
    assert _abi.load().rc_device_count(ctypes.byref(n)) == _abi.ERR_NO_DEVICE and n.value == 0


def test_geometry_matches_oracle(oracle):
    for W, H in [(1, 1), (3, 2), (97, 61), (256, 256), (1920, 1080), (3840, 2160), (16384, 7)]:
        assert rc.levels_for_coverage(W, H) == oracle.levels_for_coverage(W, H)
        for i in range(8):
            assert rc.probe_grid(W, H, i) == oracle.probe_grid(W, H, i)
            assert rc.level_rays(W, H, i) == oracle.level_rays(W, H, i)
    with pytest.raises(rc.RcError):
        rc.probe_grid(0, 5, 0)
    assert rc.level_rays(5, 5, 8) == 0


def test_scene_generator_matches_oracle(oracle):
    for seed in [0, 1, 7, 12345, 2 ** 63 + 11]:
        for n, k in [(8, 2), (32, 4), (64, 0), (256, 256), (0, 0)]:
            assert np.array_equal(rc.scene_generate(seed, n, k), oracle.scene(seed, n, k))
    with pytest.raises(rc.RcError):
        rc.scene_generate(0, 3, 4)
    with pytest.raises(rc.RcError):
        rc.scene_generate(0, 257, 0)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "radiancecascade_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f
                assert "librco" not in text and "rco.h" not in text, f
