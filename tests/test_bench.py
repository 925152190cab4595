"""bench.py's contract logic on the CPU: the one-GPU refusal, the traffic
models the roofline figures are built on, the reader of the committed
profile, and the FP64 checks of the CPU leg."""
import os
import subprocess
import sys

import numpy as np
import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# rays per level at 1920x1080, 6 levels: Px_i * Py_i * N_i with Px_i = ceil(W / s_i)
RAYS_1080P = [2073600, 2073600, 2073600, 2088960, 2088960, 2088960]


def test_rays_per_level_from_the_grid(oracle):
    assert [oracle.level_rays(1920, 1080, i) for i in range(6)] == RAYS_1080P
    assert [oracle.level_rays(256, 256, i) for i in range(4)] == [65536] * 4


def test_gpus_above_one_exits_before_touching_a_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_frame_contract_bytes():
    """W H (28 + 8 + 12) + 2 b sum_{i >= 1} R_i: 432,783,360 B at 1080p fp32,
    266,158,080 B with 16-bit storage, 9,437,184 B at 256^2 with 4 levels."""
    assert bench.b_alg(1920, 1080, 6, "f32", RAYS_1080P) == 432_783_360
    assert bench.b_alg(1920, 1080, 6, "f16", RAYS_1080P) == 266_158_080
    assert bench.b_alg(1920, 1080, 6, "bf16", RAYS_1080P) == 266_158_080
    assert bench.b_alg(256, 256, 4, "f32", [65536] * 4) == 9_437_184


def test_family_bytes_of_the_split_frame():
    n = 1920 * 1080
    fam = bench.family_alg_bytes(1920, 1080, 6, "f32", RAYS_1080P, True, 11, top_in_march=True, sdf="exact")
    # march: the 1-byte field, a 4-byte code per ray of levels 1..5, the top level's texels
    assert fam["march"] == n + 4 * sum(RAYS_1080P[1:]) + 16 * (RAYS_1080P[5] // 4) == 52_085_760
    # levels 1..4 (the march merged level 5): own texels + upper texels + codes
    lv = sum(16 * (RAYS_1080P[i] // 4) + 16 * (RAYS_1080P[i + 1] // 4) + 4 * RAYS_1080P[i] for i in range(1, 5))
    assert fam["levels"] == lv == 99_962_880
    assert fam["level0"] == n + 16 * (RAYS_1080P[1] // 4) + 12 * n == 35_251_200
    assert fam["seed"] == 16 * n + n / 8
    assert fam["jfa"] == (n / 8 + n / 16) + (n / 8 + n / 16 + n)
    # the jump flood: 8 B/px per pass but the last, which reads 4 and writes the field
    jf = bench.family_alg_bytes(1920, 1080, 6, "f32", RAYS_1080P, True, 11, top_in_march=True, sdf="jfa")
    assert jf["jfa"] == (8 * 10 + 5) * n and jf["seed"] == 20 * n
    # without the split march the level kernels march in place: no codes, no march family
    ip = bench.family_alg_bytes(1920, 1080, 6, "f32", RAYS_1080P, False, 11)
    assert ip["march"] == 0
    assert ip["levels"] == sum(16 * (RAYS_1080P[i] // 4) + (16 * (RAYS_1080P[i + 1] // 4) if i + 1 < 6 else 0)
                               for i in range(1, 6))


def test_profiled_traffic_reads_the_committed_profile():
    got = bench.profiled_traffic("march", 1920, 1080, 6, "f32")
    want = None
    for line in open(os.path.join(ROOT, "profiles", "frame_1080p_plain.md")):
        if line.startswith("| `family:march`"):
            want = float(line.split("|")[4]) * 1e6
    assert want is not None and got == want
    assert 1.0 < got / 52_085_760 < 4.0   # the march moves more than its algorithmic bytes, not wildly more
    assert bench.profiled_traffic("march", 3840, 2160, 7, "f32") is None
    assert bench.profiled_traffic("march", 1920, 1080, 6, "f16") is None


def test_cpu_leg_checks_frames_against_the_fp64_oracle(oracle):
    d, _, e = oracle.gbuffer(0, 240, 136, 32, 4)
    assert (e != 0).any()
    f32 = oracle.run(d, e, 4, np.float32)
    r = bench.cpu_baseline(d, e, 4, 0.2, {"fp32": f32, "zero": np.zeros_like(f32), "twice": 2 * f32})
    assert r["value"] > 0 and r["cores"] >= 1 and r["kind"] == "port"
    c = r["checks"]
    assert c["fp32"] < 1e-5           # the FP32 flavour against the FP64 one
    assert c["zero"] == pytest.approx(1.0)
    assert c["twice"] == pytest.approx(1.0, rel=1e-5)
    assert bench.cpu_baseline(d, e, 4, 0.05)["checks"] is None


def test_cpu_leg_check_of_an_unlit_frame_is_null(oracle):
    d, _, e = oracle.gbuffer(0, 64, 64, 4, 0)
    assert not (e != 0).any()
    r = bench.cpu_baseline(d, e, 3, 0.05, {"plain": np.zeros((3, 64, 64), np.float32)})
    assert r["checks"] == {"plain": None}
