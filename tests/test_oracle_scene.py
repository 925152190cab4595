"""Pins for the oracle's scene generator and analytic G-buffer raster.

The PRNG is pinned by SplitMix64's published output for seed 0; the raster
by closed forms (room-only depth, a box straight ahead, the look_to_rh
handedness of src/camera.rs:43-52) that do not go through the slab code.
"""
import numpy as np
import pytest


def test_splitmix64_published_vector(oracle):
    # Vigna's splitmix64.c, seed 0
    assert oracle.splitmix64_stream(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_scene_ranges_and_clear_zone(oracle):
    for seed in range(20):
        for n, k in [(8, 2), (32, 4), (64, 64), (0, 0)]:
            b = oracle.scene(seed, n, k)
            assert b.shape == (n, 12)
            if n == 0:
                continue
            lo, hi, alb, em = b[:, 0:3], b[:, 3:6], b[:, 6:9], b[:, 9:12]
            e = hi - lo
            c = 0.5 * (hi + lo)
            assert np.all(e >= 0.1 - 1e-12) and np.all(e <= 0.4 + 1e-12)
            assert np.all(c >= -1e-12) and np.all(c <= 1 + 1e-12)
            assert np.all(alb >= 0.2) and np.all(alb <= 0.9)
            assert np.all(em[:k] >= 1.0) and np.all(em[:k] <= 10.0)
            assert np.all(em[k:] == 0.0)
            # no box overlaps the open clear zone (0.35, 0.65)^3 around the camera
            overlap = np.all((lo < 0.65) & (hi > 0.35), axis=1)
            assert not overlap.any()


def test_scene_invalid(oracle):
    with pytest.raises(ValueError):
        oracle.scene(0, 4, 5)
    with pytest.raises(ValueError):
        oracle.scene(0, oracle.MAX_BOXES + 1, 0)


def _camera_dirs(W, H):
    th = 1.0 / np.sqrt(3.0)
    x = np.arange(W, dtype=np.float64)
    y = np.arange(H, dtype=np.float64)
    ndc_x = (2.0 * (x + 0.5)) / W - 1.0
    ndc_y = 1.0 - (2.0 * (y + 0.5)) / H
    u = (ndc_x * th) * (W / H)
    v = ndc_y * th
    return np.broadcast_to(-u[None, :], (H, W)), np.broadcast_to(v[:, None], (H, W))


@pytest.mark.parametrize("W,H", [(64, 48), (31, 17), (1, 1), (200, 3)])
def test_room_only_depth_closed_form(oracle, W, H):
    """Inside the unit room the view ray (dx, dy, 1) from the centre leaves
    through the nearest wall: depth = min(0.5, 0.5/|dx|, 0.5/|dy|)."""
    d, a, e = oracle.raster(np.zeros((0, 12)), W, H)
    dx, dy = _camera_dirs(W, H)
    with np.errstate(divide="ignore"):
        ref = np.minimum(0.5, np.minimum(0.5 / np.abs(dx), 0.5 / np.abs(dy))).astype(np.float32)
    assert np.array_equal(d, ref)
    assert np.all(a == np.float32(0.75)) and np.all(e == 0)


def test_box_ahead_and_handedness(oracle):
    W, H = 96, 64
    # box straight ahead: front face z = 0.8 -> depth 0.3 at the image centre
    ahead = np.array([[0.45, 0.45, 0.8, 0.55, 0.55, 0.9, 0.3, 0.4, 0.5, 2.0, 3.0, 4.0]])
    d, a, e = oracle.raster(ahead, W, H)
    cy, cx = H // 2, W // 2
    assert d[cy, cx] == np.float32(0.3 / 1.0) or abs(float(d[cy, cx]) - 0.3) < 1e-7
    assert tuple(a[:, cy, cx]) == (np.float32(0.3), np.float32(0.4), np.float32(0.5))
    assert tuple(e[:, cy, cx]) == (np.float32(2.0), np.float32(3.0), np.float32(4.0))
    assert e[:, 0, 0].sum() == 0  # corners see the room
    # a box on the +x side of the room appears in the LEFT half of the image:
    # look_to_rh with forward +z and up +y has right = forward x up = -x
    side = np.array([[0.75, 0.45, 0.8, 0.95, 0.55, 0.9, 0.5, 0.5, 0.5, 1.0, 1.0, 1.0]])
    d, a, e = oracle.raster(side, W, H)
    cols = np.nonzero(e[0].sum(axis=0) > 0)[0]
    assert cols.size > 0 and cols.max() < W // 2


def test_depth_matches_slab_distance(oracle):
    """Depth of every box pixel equals the distance to the box face the ray
    enters, from an independent per-face intersection."""
    W, H = 40, 30
    boxes = oracle.scene(3, 16, 4)
    d, a, e = oracle.raster(boxes, W, H)
    dx, dy = _camera_dirs(W, H)
    best = np.minimum(0.5, np.minimum(0.5 / np.abs(dx), 0.5 / np.abs(dy)))
    for b in boxes:
        lo, hi = b[0:3], b[3:6]
        for axis, comp in ((0, dx), (1, dy), (2, np.ones_like(dx))):
            for plane in (lo[axis], hi[axis]):
                with np.errstate(divide="ignore", invalid="ignore"):
                    t = (plane - 0.5) / comp
                p = [0.5 + t * dx, 0.5 + t * dy, 0.5 + t * 1.0]
                inside = np.ones_like(t, dtype=bool)
                for j in range(3):
                    if j != axis:
                        inside &= (p[j] >= lo[j]) & (p[j] <= hi[j])
                ok = inside & (t > 0)
                best = np.where(ok & (t < best), t, best)
    assert np.allclose(d, best.astype(np.float32), rtol=1e-6, atol=0)


def test_ndc_depth_conversion(oracle):
    """perspective_rh NDC depth (src/camera.rs:77-79) and its inverse as the
    pass evaluates it: the near and far planes map to 0 and 1, the inverse
    is monotonic, and the fp32 round trip of depths in [0.15, 1] (the room's
    are within [0.15, 0.87]) stays within 1e-6 relative."""
    n, f = 0.1, 100.0
    assert oracle.ndc_from_linear(np.array([n]), n, f)[0] == 0.0
    assert oracle.ndc_from_linear(np.array([f]), n, f)[0] == 1.0
    back = oracle.linear_from_ndc(np.array([0.0, 1.0], np.float32), n, f)
    # at the far plane f - d (f - n) cancels to f - fp32(f - n): 1.5e-5 relative
    assert abs(back[0] - n) <= 1e-6 * n and abs(back[1] - f) <= 2e-5 * f
    D = np.linspace(0.15, 1.0, 20001)
    d = oracle.ndc_from_linear(D, n, f)
    lin = oracle.linear_from_ndc(d, n, f).astype(np.float64)
    assert np.all(np.diff(lin) >= 0)
    assert np.max(np.abs(lin - D) / D) <= 1e-6
    # the G-buffer of a seeded scene keeps every occupancy bit
    dep, _, e = oracle.gbuffer(7, 256, 256, 8, 2)
    s_lin = oracle.seed_init(dep, e)
    s_ndc = oracle.seed_init(oracle.linear_from_ndc(oracle.ndc_from_linear(dep, n, f), n, f), e)
    assert np.array_equal(s_lin, s_ndc)
