"""Pins for the composite reference (oracle/composite.py) and the library's
threshold table (a host function: no GPU needed)."""
import numpy as np

from oracle import composite as co


def test_thresholds_match_the_library():
    import radiancecascade_amd as rc
    t = rc.srgb_thresholds()
    ref = co.thresholds()
    assert t[0] == -np.inf and np.array_equal(t[1:], ref[1:])
    assert np.all(np.diff(t[1:]) > 0)


def test_srgb_closed_forms():
    # 0 -> 0, 1 -> 255, values past the ends clamp, NaN -> 0
    assert list(co.encode8([0.0, -1.0, 1.0, 7.5, np.nan])) == [0, 0, 255, 255, 0]
    # linear 0.5 has sRGB code 0.7354 -> 187.5.. -> 188 (the textbook value)
    assert co.encode8([0.5])[0] == 188
    # linear 0.2140 is sRGB code 0.5 (127.5/255 rounds up to 128 just above)
    assert co.encode8([0.2141])[0] == 128 and co.encode8([0.2138])[0] == 127
    # the linear segment: c = 0.001 -> 12.92e-3 * 255 = 3.29 -> 3
    assert co.encode8([0.001])[0] == 3
    # decode(encode(x)) is the identity
    x = np.linspace(0, 1, 1001)
    assert np.allclose(co.srgb_decode(co.srgb_encode(x)), x, atol=1e-12)


def test_threshold_rule_is_round_to_nearest_away_from_boundaries():
    rng = np.random.default_rng(0)
    c = rng.random(200000).astype(np.float32)
    code = co.srgb_encode(c.astype(np.float64)) * 255.0
    far = np.abs(code - np.floor(code) - 0.5) > 1e-4
    assert far.mean() > 0.99
    assert np.array_equal(co.encode8(c)[far], np.floor(code[far] + 0.5).astype(np.uint8))


def test_composite_layout_and_order():
    H, W = 2, 3
    a = np.full((3, H, W), 0.5, np.float32)
    i = np.zeros((3, H, W), np.float32)
    e = np.zeros((3, H, W), np.float32)
    i[0] = 2.0            # red: 0.5 * 2 = 1 -> 255
    e[1, 0, 0] = 0.5      # green at (0, 0): 0.5 -> 188
    d = np.zeros((3, H, W), np.float32)
    d[2] = 1.0            # blue: direct only -> 255
    out = co.composite(a, e, i, d)
    assert out.dtype == np.uint32 and out.shape == (H, W)
    px = out[0, 0]
    b, g, r, al = px & 0xFF, (px >> 8) & 0xFF, (px >> 16) & 0xFF, px >> 24
    assert (b, g, r, al) == (255, 188, 255, 255)
    assert (out[1, 2] >> 8) & 0xFF == 0
    # without direct the blue channel is dark
    assert np.all(co.composite(a, e, i) & 0xFF == 0)
