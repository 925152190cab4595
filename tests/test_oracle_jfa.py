"""Pins for the oracle's occupancy rule, jump flood and distance field.

The brute-force nearest seed is checked against scipy's exact Euclidean
distance transform; the jump flood is exact for one seed and never nearer
than the exact answer otherwise (Rong & Tan 2006 error behaviour).
"""
import numpy as np
import pytest

from scipy import ndimage


def _seed_field(W, H, pts):
    s = np.full((H, W), 0xFFFFFFFF, np.uint32)
    for x, y in pts:
        s[y, x] = (y << 16) | x
    return s


def _d2(seeds):
    H, W = seeds.shape
    yy, xx = np.mgrid[0:H, 0:W]
    sx = (seeds & 0xFFFF).astype(np.int64)
    sy = (seeds >> 16).astype(np.int64)
    d2 = (sx - xx) ** 2 + (sy - yy) ** 2
    return np.where(seeds == 0xFFFFFFFF, -1, d2)


def test_pass_count(oracle):
    assert [oracle.jfa_num_passes(w, w) for w in (1, 2, 3, 4, 5, 256, 257, 1920)] == [0, 1, 2, 2, 3, 8, 9, 11]
    assert oracle.jfa_num_passes(1920, 1080) == 11


def test_occupancy_rule(oracle):
    W, H = 8, 6
    depth = np.ones((H, W), np.float32)
    depth[:, 4:] = 1.06  # right half 6 % farther: column 3 is the near side
    depth[0, :] = np.float32(1.0)
    depth[0, 4:] = 1.04  # row 0: only 4 % farther -> not a silhouette
    emis = np.zeros((3, H, W), np.float32)
    emis[2, 5, 0] = 0.5  # blue-only emitter still occupies
    s = oracle.seed_init(depth, emis)
    occ = s != oracle.SEED_NONE
    expect = np.zeros((H, W), bool)
    expect[1:, 3] = True
    expect[5, 0] = True
    assert np.array_equal(occ, expect)
    y, x = 5, 0
    assert s[y, x] == (y << 16) | x


def test_single_seed_exact(oracle):
    for (W, H, x, y) in [(64, 64, 10, 50), (37, 5, 36, 0), (1, 1, 0, 0), (100, 3, 0, 2)]:
        s = oracle.jfa(_seed_field(W, H, [(x, y)]))
        assert np.all(s == ((y << 16) | x))
        sdf = oracle.sdf_f32(s)
        yy, xx = np.mgrid[0:H, 0:W]
        d = np.sqrt((xx - x) ** 2 + (yy - y) ** 2)
        assert np.array_equal(sdf, _coded(d).astype(np.float32))


def test_empty_field(oracle):
    s = oracle.jfa(_seed_field(33, 17, []))
    assert np.all(s == oracle.SEED_NONE)
    assert np.all(oracle.sdf_f32(s) == oracle.SDF_FAR)


def test_bruteforce_matches_scipy_edt(oracle):
    rng = np.random.default_rng(1)
    for (W, H, n) in [(48, 40, 5), (64, 64, 60), (31, 57, 300)]:
        pts = {(int(rng.integers(W)), int(rng.integers(H))) for _ in range(n)}
        init = _seed_field(W, H, pts)
        near = oracle.nearest_seed_bruteforce(init)
        edt = ndimage.distance_transform_edt(init == 0xFFFFFFFF)
        assert np.allclose(np.sqrt(_d2(near)), edt, rtol=0, atol=1e-12)


def test_tie_break(oracle):
    # pixel (5,5) is equidistant from (3,5) and (7,5) and from (5,3),(5,7):
    # (d2, y, x) lexicographic picks (5,3) (smallest y)
    init = _seed_field(11, 11, [(3, 5), (7, 5), (5, 3), (5, 7)])
    near = oracle.nearest_seed_bruteforce(init)
    assert near[5, 5] == (3 << 16) | 5
    assert oracle.jfa(init)[5, 5] == (3 << 16) | 5


def test_jfa_vs_exact(oracle):
    rng = np.random.default_rng(7)
    for (W, H, n) in [(64, 64, 3), (64, 64, 40), (50, 30, 200)]:
        pts = {(int(rng.integers(W)), int(rng.integers(H))) for _ in range(n)}
        init = _seed_field(W, H, pts)
        jfa = oracle.jfa(init)
        exact = oracle.nearest_seed_bruteforce(init)
        dj, de = _d2(jfa), _d2(exact)
        assert np.all(dj >= de)                      # never nearer than exact
        assert np.all(np.isin(jfa, list(init[init != 0xFFFFFFFF])))  # only real seeds
        assert (dj != de).mean() < 0.01              # JFA's rare misses
        for x, y in pts:                             # seeds keep themselves
            assert jfa[y, x] == (y << 16) | x


def test_pass_is_order_independent(oracle):
    """The min-key rule makes a pass independent of neighbour order: the
    result equals the brute-force minimum over the 9 candidates."""
    rng = np.random.default_rng(3)
    W, H = 23, 19
    pts = {(int(rng.integers(W)), int(rng.integers(H))) for _ in range(30)}
    s = oracle.jfa_pass(_seed_field(W, H, pts), 4)
    s = oracle.jfa_pass(s, 2)
    out = oracle.jfa_pass(s, 1)
    for y in range(H):
        for x in range(W):
            cands = []
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    qx, qy = x + dx, y + dy
                    if 0 <= qx < W and 0 <= qy < H and s[qy, qx] != 0xFFFFFFFF:
                        v = int(s[qy, qx])
                        cands.append(((v & 0xFFFF) - x) ** 2 + ((v >> 16) - y) ** 2 << 32 | v)
            assert out[y, x] == (min(cands) & 0xFFFFFFFF if cands else 0xFFFFFFFF)


def _coded(d):
    """Independent restatement of the conservative distance code (float input)."""
    f = np.floor(np.asarray(d, np.float64) + 1e-9)
    return np.where(f < 128, f, np.where(f < 2160, 128 + 16 * np.floor((f - 128) / 16), 2160.0))


def test_sdf_values(oracle):
    init = _seed_field(9, 9, [(0, 0)])
    s = oracle.jfa(init)
    f32 = oracle.sdf_f32(s)
    f64 = oracle.sdf_f64(s)
    assert f32[0, 0] == 0 and f32[3, 4] == np.float32(5.0) and f64[8, 8] == 11.0   # floor(8 sqrt 2)
    assert f32.dtype == np.float32 and f64.dtype == np.float64


def test_sdf_code_rule(oracle):
    # exact squares, one below a square, the band edges and the cap
    for d2 in [0, 1, 2, 3, 4, 99, 100, 101, 127 ** 2, 128 ** 2 - 1, 128 ** 2, 143 ** 2, 144 ** 2 - 1, 144 ** 2,
               2159 ** 2, 2160 ** 2, 5000 ** 2, 2 * 16383 ** 2]:
        v = oracle.sdf_value_d2(d2)
        d = d2 ** 0.5
        assert v <= d + 1e-12                                   # conservative: never past a surface
        assert (v == 0) == (d2 == 0)                            # hits are exact
        assert v == _coded(np.sqrt(float(d2)))
    assert oracle.sdf_value_d2(127 ** 2) == 127 and oracle.sdf_value_d2(128 ** 2 - 1) == 127
    assert oracle.sdf_value_d2(143 ** 2) == 128 and oracle.sdf_value_d2(144 ** 2) == 144
    assert oracle.sdf_value_d2(10 ** 8) == 2160


# ---- exact transform (rco_edt_d2): the default distance field -------------------
def test_edt_matches_bruteforce_and_scipy(oracle):
    """rco_edt_d2 (Felzenszwalb & Huttenlocher, exact integer envelope) equals
    the brute-force nearest seed's d^2 and scipy's exact transform (a library)
    on random fields of every density, ragged and degenerate shapes."""
    rng = np.random.default_rng(11)
    shapes = [(1, 1), (1, 9), (9, 1), (2, 2), (64, 1), (1, 64), (17, 23), (64, 64), (65, 63), (130, 70)]
    for (W, H) in shapes:
        for p in (0.0, 0.001, 0.02, 0.2, 0.7, 1.0):
            occ = rng.random((H, W)) < p
            init = _seed_field(W, H, [(int(x), int(y)) for y, x in zip(*np.nonzero(occ))])
            d2 = oracle.edt_d2(init)
            assert np.array_equal(d2, _d2(oracle.nearest_seed_bruteforce(init))), (W, H, p)
            if occ.any():
                edt = ndimage.distance_transform_edt(~occ)
                assert np.array_equal(d2, np.rint(edt ** 2).astype(np.int64)), (W, H, p)
            else:
                assert np.all(d2 == -1)


def test_edt_closed_forms(oracle):
    # one seed: d^2 is the squared distance to it everywhere
    W, H = 37, 29
    d2 = oracle.edt_d2(_seed_field(W, H, [(30, 4)]))
    yy, xx = np.mgrid[0:H, 0:W]
    assert np.array_equal(d2, (xx - 30) ** 2 + (yy - 4) ** 2)
    # a full column of seeds at x = 0: d^2 = x^2 (the row envelope's first site only)
    d2 = oracle.edt_d2(_seed_field(W, H, [(0, y) for y in range(H)]))
    assert np.array_equal(d2, np.broadcast_to(np.arange(W) ** 2, (H, W)))
    # the four corners of a large image: the farthest pixel is the centre
    W, H = 16384, 3
    d2 = oracle.edt_d2(_seed_field(W, H, [(0, 0), (W - 1, 0), (0, H - 1), (W - 1, H - 1)]))
    assert d2.max() == (W // 2 - 1) ** 2 + 1 and d2[1, W // 2 - 1] == (W // 2 - 1) ** 2 + 1


def test_edt_is_never_farther_than_the_jump_flood(oracle):
    """The exact field is a lower bound of the jump flood's, equal except at
    the flood's rare misses; the SDF codes follow."""
    rng = np.random.default_rng(5)
    for (W, H, n) in [(64, 64, 3), (64, 64, 40), (90, 50, 200), (200, 120, 25)]:
        pts = {(int(rng.integers(W)), int(rng.integers(H))) for _ in range(n)}
        init = _seed_field(W, H, pts)
        de, dj = oracle.edt_d2(init), _d2(oracle.jfa(init))
        assert np.all(de <= dj) and (de != dj).mean() < 0.01
        se, sj = oracle.sdf_from_d2(de), oracle.sdf_f32(oracle.jfa(init))
        assert np.all(se <= sj)
        assert np.array_equal(oracle.sdf_from_d2(de, np.float64), np.vectorize(oracle.sdf_value_d2)(de))


def test_whole_path_modes(oracle):
    """rco_run's exact mode equals the stage chain on the exact field; its jump-flood
    mode equals the chain on the flood's field; with no flood miss they agree."""
    d, _, e = oracle.gbuffer(3, 96, 64, 8, 2)
    L = oracle.levels_for_coverage(96, 64)
    init = oracle.seed_init(d, e)
    for dt in (np.float32, np.float64):
        ex = oracle.run(d, e, L, dt)
        assert np.array_equal(ex, oracle.run_on_field(oracle.sdf_from_d2(oracle.edt_d2(init), dt), e, L))
        jf = oracle.run(d, e, L, dt, sdf="jfa")
        sj = oracle.sdf_f32(oracle.jfa(init)) if dt == np.float32 else oracle.sdf_f64(oracle.jfa(init))
        assert np.array_equal(jf, oracle.run_on_field(sj, e, L))
    with pytest.raises(ValueError):
        oracle.run(d, e, L, np.float32, want_seeds=True)
