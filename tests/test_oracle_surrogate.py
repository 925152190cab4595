"""Pins for the learned-far-field oracle (oracle/surrogate.py): a library
EDT, block means taken directly from pixels, closed forms and symmetries of
the cone trace, and the frame's reduction to the plain path when the
network's level 2 is the plain one."""
import numpy as np
import pytest

from oracle import rco
from oracle import surrogate as so


def _occ_field(H, W, p, seed):
    rng = np.random.default_rng(seed)
    occ = rng.random((H, W)) < p
    yy, xx = np.mgrid[0:H, 0:W]
    return np.where(occ, (yy.astype(np.uint32) << 16) | xx.astype(np.uint32), np.uint32(rco.SEED_NONE)).astype(np.uint32)


@pytest.mark.parametrize("H,W,p,seed", [(48, 64, 0.01, 1), (37, 53, 0.002, 2), (64, 64, 0.0005, 3)])
def test_near_sdf_matches_library_edt(H, W, p, seed):
    ndimage = pytest.importorskip("scipy.ndimage")
    field = _occ_field(H, W, p, seed)
    occ = field != rco.SEED_NONE
    if not occ.any():
        occ[H // 2, W // 3] = True
        field[H // 2, W // 3] = ((H // 2) << 16) | (W // 3)
    edt = ndimage.distance_transform_edt(~occ)
    got = so.near_sdf(field)
    near = edt <= so.NEAR_R
    assert np.array_equal(got[near], np.floor(edt[near] + 1e-9).astype(np.float32))
    assert (got[~near] == rco.SDF_FAR).all()


def test_near_sdf_empty_and_jfa_agreement():
    assert (so.near_sdf(np.full((20, 30), rco.SEED_NONE, np.uint32)) == rco.SDF_FAR).all()
    d, _, e = rco.gbuffer(9, 128, 96, 8, 2)
    s0 = rco.seed_init(d, e)
    jfa = rco.sdf_f32(rco.jfa(s0))
    got = so.near_sdf(s0)
    m = jfa < so.NEAR_R
    # the jump flood is exact almost everywhere; the near field is exact below
    # 10 px, and both fields are >= 10 elsewhere (which ends a level-0/1 step)
    assert (got[m] == jfa[m]).mean() > 0.999
    assert np.all(got[m] <= jfa[m])
    assert np.all(got[~m] >= so.NEAR_R) and np.all(jfa[~m] >= so.NEAR_R)


def _blocks(a, b, H, W):
    """[.., h, b, w, b] view of a zero-padded to whole b x b blocks."""
    h, w = -(-H // b), -(-W // b)
    pad = np.zeros(a.shape[:-2] + (h * b, w * b), a.dtype)
    pad[..., :H, :W] = a
    return pad.reshape(a.shape[:-2] + (h, b, w, b))


def test_mip_block_means_and_row_column_fractions():
    d, _, e = rco.gbuffer(9, 200, 136, 8, 2)
    s0 = rco.seed_init(d, e)
    EM, em, oc = so.pixel_fields(s0, e)
    levels = so.mip(s0, e)
    H, W = s0.shape
    assert len(levels) - 1 == so.top_level(W, H) == 8
    for m in range(1, len(levels)):
        b = 1 << m
        emi, cov = levels[m]
        h, w = -(-H // b), -(-W // b)
        assert emi.shape == (h, w, 4) and cov.shape == (h, w, 4)
        # means: one fp16 rounding per level on top of the exact block mean
        ref = np.concatenate([_blocks(EM.astype(np.float64), b, H, W).mean(axis=(-3, -1)),
                              _blocks(em.astype(np.float64), b, H, W).mean(axis=(-3, -1))[None]])
        assert np.allclose(emi, np.moveaxis(ref, 0, -1), rtol=m * 2.0 ** -11, atol=m * 2.0 ** -24), m
        # row / column fractions: the max of the children never exceeds the
        # true fraction of held columns (rows), and equals it at level 1
        for k, (mask, ax) in enumerate(((em, "x"), (em, "y"), (oc, "x"), (oc, "y"))):
            blk = _blocks(mask, b, H, W)
            true = (blk.any(axis=-3).sum(axis=-1) if ax == "x" else blk.any(axis=-1).sum(axis=-2)) / b
            got = cov[..., k]
            assert np.all(got <= true + 1e-12), (m, k)
            if m == 1:
                assert np.array_equal(got, true)
    # mass: sum over the top cell times its area is the image total (fp16 means)
    top_em, _ = levels[-1]
    assert np.isclose(top_em[..., 0].sum() * 4 ** (len(levels) - 1), EM[0].sum(), rtol=1e-2)


def test_mip_fractions_exact_for_lines_and_pixels():
    H = W = 256
    e = np.zeros((3, H, W), np.float32)
    e[:, 100, :] = 2.0                      # an emissive row across the image
    e[:, :, 37] = 3.0                       # an emissive column
    depth = np.ones((H, W), np.float32)
    s0 = rco.seed_init(depth, e)
    levels = so.mip(s0, e)
    for m in range(1, len(levels)):
        b = 1 << m
        emi, cov = levels[m]
        ry, cx = 100 // b, 37 // b
        row = cov[ry, :, :]                 # cells on the row
        assert np.all(row[:, 0] == 1.0)     # every column of the cell is held
        others = np.ones(W // b, bool)
        others[cx] = False
        assert np.all(row[others, 1] == 1.0 / b)   # one row of b
        col = cov[:, cx, :]
        assert np.all(col[:, 1] == 1.0)
        rows = np.ones(H // b, bool)
        rows[ry] = False
        assert np.all(col[rows, 0] == 1.0 / b)
        assert np.all(cov[..., 2:] == 0)    # no dark occluder pixels
    # a single occluder pixel: 1/b of the columns and rows, area-free
    depth = np.ones((64, 64), np.float32)
    depth[20:, :] = 2.0                     # depth edge: row 19 is the near side
    depth[:, :] = 1.0
    depth[33, 40] = 0.5                     # one nearer pixel: it is the near side of its 4 edges
    s0 = rco.seed_init(depth, np.zeros((3, 64, 64), np.float32))
    occ = s0 != rco.SEED_NONE
    assert occ.sum() == 1 and occ[33, 40]
    levels = so.mip(s0, np.zeros((3, 64, 64), np.float32))
    for m in range(1, len(levels)):
        b = 1 << m
        _, cov = levels[m]
        assert cov[33 // b, 40 // b, 2] == cov[33 // b, 40 // b, 3] == 1.0 / b
        assert cov[..., 2:].sum() == 2.0 / b


def test_cone_schedule_shape():
    s = so.cone_schedule(1920, 1080)
    ts = [t for t, *_ in s]
    assert ts[0] == 10.0 and all(b > a for a, b in zip(ts, ts[1:]))
    assert ts[-1] < np.hypot(1920, 1080) + 2 <= ts[-1] + 0.75 * max(2 * ts[-1] * np.tan(np.pi / 16), 1)
    assert all(1 <= m <= so.top_level(1920, 1080) for _, m, _, _ in s)
    # exponents: the step's share of a cell, at least 3/4 (the step is 3/4 of the diameter)
    assert all(0.75 <= re <= 1.0 and np.isclose(ro, so.CONE_OCC_GAIN * re, rtol=1e-6) for _, _, re, ro in s)
    assert len(s) < 64


def test_cone_empty_scene():
    H = W = 64
    field = np.full((H, W), rco.SEED_NONE, np.uint32)
    cone = so.cone_trace(so.mip(field, np.zeros((3, H, W), np.float32)), W, H)
    assert np.all(cone[..., :3] == 0) and np.all(cone[..., 3] == 1)


def test_cone_fully_emissive_scene():
    H = W = 128
    e = np.full((3, H, W), 2.5, np.float32)
    e[1] = 4.0
    field = rco.seed_init(np.ones((H, W), np.float32), e)
    cone = so.cone_trace(so.mip(field, e), W, H)
    # interior probes: every sample lies in fully emissive cells, so
    # L = E (1 - T) exactly and T falls below the stop threshold at once
    inner = cone[3:-3, 3:-3]
    T = inner[..., 3]
    assert np.all(T < so.T_MIN)
    assert np.allclose(inner[..., 0], 2.5 * (1 - T), rtol=1e-12) and np.allclose(inner[..., 1], 4.0 * (1 - T), rtol=1e-12)


def test_cone_line_across_the_cone():
    """A 3-px emissive line across the whole image lights every cone that
    crosses it with most of its radiance (its area share of a cell is small;
    its column share is 1), and a dark line stops those cones."""
    H = W = 256
    for dark in (False, True):
        e = np.zeros((3, H, W), np.float32)
        depth = np.ones((H, W), np.float32)
        if dark:
            depth[60:63, :] = 0.5           # a nearer band: its outline rows are occluders
            e[:, 200:, :] = 1.0             # light below
        else:
            e[:, 60:63, :] = 5.0
        s0 = rco.seed_init(depth, e)
        cone = so.cone_trace(so.mip(s0, e), W, H)
        dirs = rco.dir_table_f32(1)
        up = [k for k in range(16) if dirs[k][1] < -0.95]      # the two cones nearest to straight up
        # level-2 probes 14..126 px below the line (rows 9..23), away from the sides
        sub = cone[9:24, 4:-4][:, :, up]
        if dark:
            assert np.all(sub[..., 3] < 0.05), sub[..., 3].max()
        else:
            # area share of the line in a 32-px cell is 3/32; its column share is 1
            assert np.all(sub[..., 0] > 0.6 * 5.0), sub[..., 0].min()


def test_cone_linear_in_emission_and_mirror_symmetric():
    W = H = 256
    d, _, e = rco.gbuffer(9, W, H, 8, 2)
    e = e.copy()
    d = d.copy()
    # mirror the left half onto the right: a scene symmetric under x -> W - x
    e[:, :, W // 2:] = e[:, :, :W // 2][:, :, ::-1]
    d[:, W // 2:] = d[:, :W // 2][:, ::-1]
    s0 = rco.seed_init(d, e)
    c1 = so.cone_trace(so.mip(s0, e), W, H)
    c2 = so.cone_trace(so.mip(s0, 2 * e), W, H)
    assert np.allclose(c2[..., :3], 2 * c1[..., :3], rtol=1e-12, atol=1e-12)
    assert np.array_equal(c2[..., 3], c1[..., 3])
    occ = s0 != rco.SEED_NONE
    assert np.array_equal(occ, occ[:, ::-1])
    mk = [(7 - k) % 16 for k in range(16)]
    mir = c1[:, ::-1][:, :, mk]
    assert np.allclose(mir, c1, rtol=1e-9, atol=1e-9)
    assert c1[..., :3].sum() > 0


def test_exact_level2_reproduces_the_plain_frame():
    # with the network's output equal to the plain level 2 (c + s z = A2), the
    # frame is the plain frame's levels 1 and 0
    W, H, L = 96, 80, 5
    d, _, e = rco.gbuffer(9, W, H, 8, 2)
    irr, lv = rco.run(d, e, L, want_levels=True)
    QX, QY = rco.probe_grid(W, H, 2)
    c = np.zeros((QX * QY, 48))
    A2 = so.predicted_level2(c, lv[2][..., :3].reshape(-1, 48), 1.0, QX, QY)
    assert np.array_equal(A2[..., :3], lv[2][..., :3])
    sdf = rco.sdf_from_d2(rco.edt_d2(rco.seed_init(d, e)), np.float64)
    m1, _ = rco.march_level(sdf, e, 1)
    m0, _ = rco.march_level(sdf, e, 0)
    A1 = rco.merge_level(W, H, 1, m1, A2)
    assert np.allclose(A1, lv[1], rtol=1e-12, atol=1e-12)
    assert np.allclose(rco.gather(W, H, rco.merge_level(W, H, 0, m0, A1)), irr, rtol=1e-12, atol=1e-12)


def test_samples_and_params_roundtrip():
    rng = np.random.default_rng(0)
    cone = rng.random((5, 7, 16, 4))
    A2 = rng.random((5, 7, 16, 4))
    x, t = so.samples(cone, A2)
    c, _ = so.features(cone)
    xd = x.reshape(35, 16, 4)
    assert x.shape == (35, 64) and np.allclose(np.expm1(xd[..., :3]).reshape(-1, 48) + t, A2[..., :3].reshape(-1, 48))
    assert np.array_equal(xd[..., 3], cone[..., 3].reshape(35, 16))
    assert np.array_equal(c, cone[..., :3].reshape(-1, 48))
    from oracle import mlp
    layers = mlp.init_params(3)
    p = so.layers_to_params(layers)
    assert p.size == 16640
    back = so.params_to_layers(p)
    assert all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(layers, back))
