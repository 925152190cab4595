"""Pins for the oracle's cascade geometry, march, merge and gather.

Geometry is pinned by the closed-form tables of the workload model (ray counts
per level at 256^2, 1080p and 4K, intervals, coverage levels); the march by
an empty scene and an analytic emissive disc; the merge by constant and
linear upper fields (bilinear weights 3/4, 1/4 reproduce a linear field
exactly, clamping at the border); the gather by uniform and linear fluence.
"""
import math

import numpy as np
import pytest


# ---- geometry ------------------------------------------------------------------
def test_intervals(oracle):
    assert [oracle.interval(i) for i in range(7)] == [
        (0, 2), (2, 10), (10, 42), (42, 170), (170, 682), (682, 2730), (2730, 10922)]
    for i in range(7):
        assert oracle.spacing(i) == 2 * 2 ** i and oracle.num_dirs(i) == 4 * 4 ** i


def test_level_ray_tables(oracle):
    r1080 = [oracle.level_rays(1920, 1080, i) for i in range(6)]
    assert r1080 == [2073600] * 3 + [2088960] * 3
    assert sum(r1080) == 12487680
    assert [oracle.probe_grid(1920, 1080, i) for i in range(6)] == [
        (960, 540), (480, 270), (240, 135), (120, 68), (60, 34), (30, 17)]
    assert [oracle.level_rays(256, 256, i) for i in range(4)] == [65536] * 4
    r4k = [oracle.level_rays(3840, 2160, i) for i in range(7)]
    assert r4k == [8294400] * 4 + [8355840] * 3


def test_coverage_levels(oracle):
    assert oracle.levels_for_coverage(256, 256) == 5
    assert oracle.levels_for_coverage(1920, 1080) == 6
    assert oracle.levels_for_coverage(2048, 2048) == 7
    for W, H in [(1, 1), (17, 3), (512, 768), (2048, 256)]:
        L = oracle.levels_for_coverage(W, H)
        assert oracle.interval(L - 1)[1] >= math.hypot(W, H)
        assert L == 1 or oracle.interval(L - 2)[1] < math.hypot(W, H)


def test_ray_index_bijection(oracle):
    W, H, i = 37, 23, 2
    PX, PY = oracle.probe_grid(W, H, i)
    N = oracle.num_dirs(i)
    seen = set()
    for py in range(PY):
        for px in range(PX):
            for k in (0, 1, N - 1):
                r = oracle.ray_index(W, H, i, px, py, k)
                assert (r // N % PX, r // N // PX, r % N) == (px, py, k)
                seen.add(r)
    assert max(seen) == PX * PY * N - 1


def test_direction_table_and_children(oracle):
    for i in range(5):
        d = oracle.dir_table_f32(i).astype(np.float64)
        N = oracle.num_dirs(i)
        th = 2 * np.pi * (np.arange(N) + 0.5) / N
        assert np.allclose(d[:, 0], np.cos(th), atol=6e-8) and np.allclose(d[:, 1], np.sin(th), atol=6e-8)
        # children 4k..4k+3 of level i+1 lie in the parent's sector [2pi k/N, 2pi (k+1)/N)
        ch = np.arctan2(oracle.dir_table_f32(i + 1)[:, 1], oracle.dir_table_f32(i + 1)[:, 0]) % (2 * np.pi)
        parent = np.floor(ch / (2 * np.pi / N)).astype(int)
        assert np.array_equal(parent, np.arange(4 * N) // 4)


# ---- march -------------------------------------------------------------------------
def test_march_empty_scene(oracle):
    W, H = 40, 30
    sdf = np.full((H, W), np.inf, np.float32)
    emis = np.zeros((3, H, W), np.float32)
    for i in range(4):
        for s in (sdf, sdf.astype(np.float64)):
            rgba, hit = oracle.march_level(s, emis, i)
            assert np.all(rgba[..., :3] == 0) and np.all(rgba[..., 3] == 1)
            assert np.all((hit == oracle.RAY_EXIT_SCREEN) | (hit == oracle.RAY_REACHED_T1))


def _disc_scene(W, H, cx, cy, r, E=(2.0, 3.0, 5.0)):
    yy, xx = np.mgrid[0:H, 0:W]
    inside = (xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2 <= r * r
    emis = np.zeros((3, H, W), np.float32)
    for c in range(3):
        emis[c][inside] = E[c]
    depth = np.ones((H, W), np.float32)
    return depth, emis, inside


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_march_disc_closed_form(oracle, level):
    """A ray hits a filled emissive disc iff it passes through it within its
    interval; rays within 1.5 px of the tangency or the interval ends are not
    judged (pixel discretisation)."""
    W, H, cx, cy, r = 160, 120, 97.0, 61.0, 9.0
    depth, emis, inside = _disc_scene(W, H, cx, cy, r)
    seeds = oracle.jfa(oracle.seed_init(depth, emis))
    for sdf in (oracle.sdf_f32(seeds), oracle.sdf_f64(seeds)):
        rgba, hit = oracle.march_level(sdf, emis, level)
        t0, t1 = oracle.interval(level)
        s = oracle.spacing(level)
        PX, PY = oracle.probe_grid(W, H, level)
        N = oracle.num_dirs(level)
        th = 2 * np.pi * (np.arange(N) + 0.5) / N
        dx, dy = np.cos(th), np.sin(th)
        judged = 0
        for py in range(PY):
            for px in range(PX):
                ox, oy = (px + 0.5) * s, (py + 0.5) * s
                if (ox - cx) ** 2 + (oy - cy) ** 2 <= (r + 2) ** 2:
                    continue
                # entry distance of the ray into the disc of radius rr
                bx, by = cx - ox, cy - oy
                proj = bx * dx + by * dy
                perp2 = bx * bx + by * by - proj ** 2

                def chord(rr):
                    half = np.sqrt(np.maximum(rr * rr - perp2, 0))
                    return proj - half, proj + half

                t_in, _ = chord(r - 1.5)
                sure_hit = (perp2 < (r - 1.5) ** 2) & (t_in > t0 + 1.5) & (t_in < t1 - 1.5)
                big_in, big_out = chord(r + 1.5)
                sure_miss = (perp2 > (r + 1.5) ** 2) | (big_in > t1 + 1.5) | (big_out < t0 - 1.5)
                h = hit[py, px]
                got_hit = h >= 0
                assert np.all(got_hit[sure_hit]), (py, px)
                assert not np.any(got_hit[sure_miss]), (py, px)
                hit_px = h[got_hit]
                assert np.all(inside.ravel()[hit_px])
                assert np.all(rgba[py, px][got_hit][:, :3] == np.array([2.0, 3.0, 5.0]))
                assert np.all(rgba[py, px][got_hit][:, 3] == 0)
                judged += int(sure_hit.sum() + sure_miss.sum())
        assert judged > 0


# ---- merge -----------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_merge_constant_upper(oracle, dtype):
    W, H, i = 24, 20, 1   # a 6x5 upper grid exercises clamping on every side
    PX, PY = oracle.probe_grid(W, H, i)
    UX, UY = oracle.probe_grid(W, H, i + 1)
    assert (UX, UY) == (3, 3)
    N = oracle.num_dirs(i)
    march = np.zeros((PY, PX, N, 4), dtype)
    march[..., 3] = 1
    c = np.array([0.3, 1.7, 4.25], dtype)
    upper = np.zeros((UY, UX, N, 4), dtype)
    upper[..., :3] = c
    A = oracle.merge_level(W, H, i, march, upper)
    assert np.array_equal(A[..., :3], np.broadcast_to(c, A[..., :3].shape))
    assert np.all(A[..., 3] == 1)


def test_merge_linear_field_and_clamp(oracle):
    """Bilinear weights (3/4, 1/4) reproduce a field linear in the upper probe
    index at the lower probe's coordinate p/2 - 1/4, with indices clamped."""
    W, H, i = 48, 40, 1
    PX, PY = oracle.probe_grid(W, H, i)
    UX, UY = oracle.probe_grid(W, H, i + 1)
    N = oracle.num_dirs(i)
    march = np.zeros((PY, PX, N, 4))
    march[..., 3] = 1
    uy, ux = np.mgrid[0:UY, 0:UX]
    upper = np.zeros((UY, UX, N, 4))
    upper[..., 0] = (ux + 10.0 * uy)[..., None]
    upper[..., 1] = np.arange(N)[None, None, :]      # per-direction value passes through
    A = oracle.merge_level(W, H, i, march, upper)
    py, px = np.mgrid[0:PY, 0:PX]
    fx = np.clip(px / 2 - 0.25, 0, UX - 1)
    fy = np.clip(py / 2 - 0.25, 0, UY - 1)
    expect = (fx + 10 * fy)
    assert np.allclose(A[..., 0], expect[..., None], atol=1e-12)
    # 4-child average of direction values 4k..4k+3 is 4k + 1.5
    assert np.allclose(A[..., 1], (4 * np.arange(N // 4) + 1.5)[None, None, :])


def test_merge_hits_ignore_upper(oracle):
    W, H, i = 16, 16, 2
    PX, PY = oracle.probe_grid(W, H, i)
    UX, UY = oracle.probe_grid(W, H, i + 1)
    N = oracle.num_dirs(i)
    march = np.zeros((PY, PX, N, 4))
    march[..., 0] = 2.0           # every ray hit an emitter of radiance (2, 0, 0)
    upper = np.full((UY, UX, N, 4), 100.0)
    A = oracle.merge_level(W, H, i, march, upper)
    assert np.all(A[..., 0] == 2.0) and np.all(A[..., 1:] == 0)


def test_merge_linearity(oracle):
    rng = np.random.default_rng(0)
    W, H, i = 30, 22, 1
    PX, PY = oracle.probe_grid(W, H, i)
    UX, UY = oracle.probe_grid(W, H, i + 1)
    N = oracle.num_dirs(i)
    march = np.zeros((PY, PX, N, 4))
    march[..., 3] = 1
    u1, u2 = rng.random((UY, UX, N, 4)), rng.random((UY, UX, N, 4))
    a = oracle.merge_level(W, H, i, march, 2.0 * u1 + 3.0 * u2)
    b = 2.0 * oracle.merge_level(W, H, i, march, u1) + 3.0 * oracle.merge_level(W, H, i, march, u2)
    assert np.allclose(a[..., :3], b[..., :3], rtol=1e-13, atol=1e-13)


# ---- gather and whole path ----------------------------------------------------------
def test_gather_uniform_and_linear(oracle):
    W, H = 21, 14
    PX, PY = oracle.probe_grid(W, H, 0)
    A0 = np.zeros((PY, PX, 1, 4))
    A0[..., :3] = [1.5, 2.5, 3.5]
    irr = oracle.gather(W, H, A0)
    assert np.array_equal(irr, np.broadcast_to(np.array([1.5, 2.5, 3.5])[:, None, None], irr.shape))
    py, px = np.mgrid[0:PY, 0:PX]
    A0[..., 0, 0] = px + 100 * py
    irr = oracle.gather(W, H, A0)
    y, x = np.mgrid[0:H, 0:W]
    fx = np.clip(x / 2 - 0.25, 0, PX - 1)
    fy = np.clip(y / 2 - 0.25, 0, PY - 1)
    assert np.allclose(irr[0], fx + 100 * fy, atol=1e-12)


def test_run_no_emitters_is_black(oracle):
    d, a, e = oracle.gbuffer(2, 96, 64, 8, 0)
    assert oracle.run(d, e, 4).max() == 0


def test_run_f32_vs_f64(oracle):
    d, a, e = oracle.gbuffer(7, 256, 256, 8, 2)
    i32 = oracle.run(d, e, 4, np.float32)
    i64 = oracle.run(d, e, 4, np.float64)
    assert i64.max() > 0
    assert np.linalg.norm(i32 - i64) / np.linalg.norm(i64) < 1e-5


def test_run_levels_consistent(oracle):
    """The whole-path driver equals chaining march/merge/gather by hand."""
    W, H, L = 64, 48, 4
    d, a, e = oracle.gbuffer(9, W, H, 16, 4)
    init = oracle.seed_init(d, e)
    irr, lv, seeds = oracle.run(d, e, L, np.float64, want_levels=True, want_seeds=True, sdf="jfa")
    assert np.array_equal(seeds, oracle.jfa(init))
    irr_x, lv_x = oracle.run(d, e, L, np.float64, want_levels=True)
    for sdf, irr, lv in ((oracle.sdf_f64(seeds), irr, lv),
                         (oracle.sdf_from_d2(oracle.edt_d2(init), np.float64), irr_x, lv_x)):
        up = None
        for i in range(L - 1, -1, -1):
            m, _ = oracle.march_level(sdf, e, i)
            up = oracle.merge_level(W, H, i, m, up)
            assert np.array_equal(up, lv[i])
        assert np.array_equal(oracle.gather(W, H, up), irr)


def test_uniform_sky_irradiance(oracle):
    """All rays miss and the top level holds radiance c: every level passes c
    through the merge and the irradiance is c everywhere."""
    W, H, L = 40, 28, 3
    sdf = np.full((H, W), np.inf)
    emis = np.zeros((3, H, W), np.float32)
    c = np.array([0.25, 0.5, 2.0])
    PX, PY = oracle.probe_grid(W, H, L - 1)
    up = np.zeros((PY, PX, oracle.num_dirs(L - 1) // 4, 4))
    up[..., :3] = c
    for i in range(L - 2, -1, -1):
        m, _ = oracle.march_level(sdf, emis, i)
        up = oracle.merge_level(W, H, i, m, up)
    irr = oracle.gather(W, H, up)
    assert np.allclose(irr, c[:, None, None], rtol=0, atol=1e-15)
