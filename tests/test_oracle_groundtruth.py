"""Ground truth for the distance-field choice and for the whole path (CPU).

The oracle's FP64 flavour marches the same one-byte conservative distance
code the kernels march.  Two independent references bound what that choice
and the method itself cost:

* the exact-field flavour: the same march / merge / gather on sqrt(d^2) of
  scipy's exact Euclidean distance transform (oracle.rco.sdf_exact_f64);
* the brute-force transport: from every pixel centre, N_{L-1} rays walked
  through every pixel they cross to the first occupied pixel within
  t1_{L-1} (oracle.rco.bruteforce_irradiance), itself pinned here against a
  fine-step march written independently in numpy.

The bounds are the measured errors with a margin; the differences between
the two fields are an order of magnitude below the method's own error
against the brute force, at every size measured.
"""
import numpy as np
import pytest

from oracle import rco


def rel_l2(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def _fine_step_march(occ, emis, n_dirs, t_max, step=1e-3):
    """Independent reference for the grid walk: sample every `step` px along
    each ray from each pixel centre; the first sample inside an occupied
    pixel is the hit."""
    H, W = occ.shape
    out = np.zeros((3, H, W))
    ts = np.arange(0.0, t_max + step, step)
    for k in range(n_dirs):
        th = 2.0 * np.pi * (k + 0.5) / n_dirs
        dx, dy = np.cos(th), np.sin(th)
        for y in range(H):
            for x in range(W):
                ix = np.floor(x + 0.5 + ts * dx).astype(np.int64)
                iy = np.floor(y + 0.5 + ts * dy).astype(np.int64)
                inside = (ix >= 0) & (iy >= 0) & (ix < W) & (iy < H)
                stop = np.argmin(inside) if not inside.all() else ts.size
                ixs, iys = ix[:stop], iy[:stop]
                hit = np.nonzero(occ[iys, ixs])[0]
                if hit.size:
                    out[:, y, x] += emis[:, iys[hit[0]], ixs[hit[0]]]
    return out / n_dirs


def test_grid_walk_matches_a_fine_step_march():
    rng = np.random.default_rng(5)
    H, W = 12, 14
    occ = rng.random((H, W)) < 0.12
    emis = np.where(occ[None], rng.uniform(0.0, 5.0, (3, H, W)), 0.0).astype(np.float32)
    seed = np.where(occ, (np.arange(H)[:, None] << 16) | np.arange(W)[None, :], rco.SEED_NONE).astype(np.uint32)
    n_dirs, t_max = 16, 9.0
    got = np.empty((H * W, 3))
    rco.lib().rco_bruteforce_irradiance(W, H, rco._p(seed), rco._p(np.ascontiguousarray(emis)), n_dirs, t_max,
                                        None, H * W, rco._p(got))
    ref = _fine_step_march(occ, emis.astype(np.float64), n_dirs, t_max)
    # a 1e-3 px step could only differ where a ray clips a pixel corner by
    # less than a step; none does here
    assert np.abs(got.T.reshape(3, H, W) - ref).max() <= 1e-12


def test_bruteforce_closed_forms():
    H, W = 9, 11
    e = np.zeros((3, H, W), np.float32)
    none = np.full((H, W), rco.SEED_NONE, np.uint32)
    assert np.all(rco.bruteforce_irradiance(none, e, 3) == 0.0)         # nothing to hit
    full = ((np.arange(H)[:, None] << 16) | np.arange(W)[None, :]).astype(np.uint32)
    e[:] = np.array([1.0, 2.0, 3.0], np.float32)[:, None, None]
    irr = rco.bruteforce_irradiance(full, e, 3)                          # every ray starts in an emitter
    assert np.array_equal(irr, np.broadcast_to(e.astype(np.float64), irr.shape))


CASES = [
    # (W, H, L, seed, boxes, emissive, max byte-vs-truth, max |byte - exact| vs truth, max byte-vs-exact)
    (64, 64, 3, 7, 8, 2, 0.27, 0.01, 1.2e-2),
    (128, 128, 4, 7, 8, 2, 0.13, 0.01, 5e-3),
    (256, 256, 4, 7, 8, 2, 0.115, 0.01, 6e-3),
    (256, 256, 4, 19, 8, 2, 0.065, 0.01, 4e-3),
]


@pytest.mark.parametrize("W,H,L,seed,nb,ne,b_truth,b_gap,b_fields", CASES)
def test_distance_field_choice_against_ground_truth(oracle, W, H, L, seed, nb, ne, b_truth, b_gap, b_fields):
    d, _, e = oracle.gbuffer(seed, W, H, nb, ne)
    s0 = oracle.seed_init(d, e)
    truth = oracle.bruteforce_irradiance(s0, e, L)
    byte = oracle.run(d, e, L, np.float64)
    exact = oracle.run_on_field(oracle.sdf_exact_f64(s0), e, L)
    r_byte, r_exact, r_fields = rel_l2(byte, truth), rel_l2(exact, truth), rel_l2(byte, exact)
    print(f"{W}x{H} L={L} seed {seed}: byte field vs truth {r_byte:.4f}, exact field vs truth {r_exact:.4f}, "
          f"byte vs exact {r_fields:.2e}")
    assert r_byte <= b_truth
    assert abs(r_byte - r_exact) <= b_gap
    assert r_fields <= b_fields


@pytest.mark.parametrize("seed,b_truth,b_fields", [(0, 0.025, 2.5e-3), (1, 0.04, 5e-3)])
def test_full_size_frame_against_ground_truth(oracle, seed, b_truth, b_fields):
    """1920x1080, 6 levels (the benchmark workload): the brute force on 4096
    random pixels (4096 directions each, 2730 px)."""
    W, H, L = 1920, 1080, 6
    d, _, e = oracle.gbuffer(seed, W, H, 32, 4)
    s0 = oracle.seed_init(d, e)
    pix = np.random.default_rng(1234).choice(W * H, 4096, replace=False).astype(np.int32)
    truth = oracle.bruteforce_irradiance(s0, e, L, pix)
    byte = oracle.run(d, e, L, np.float64)
    exact = oracle.run_on_field(oracle.sdf_exact_f64(s0), e, L)
    at = lambda x: x.reshape(3, -1)[:, pix]
    r_byte, r_exact, r_fields = rel_l2(at(byte), truth), rel_l2(at(exact), truth), rel_l2(byte, exact)
    print(f"1080p seed {seed}: byte field vs truth {r_byte:.4f}, exact field vs truth {r_exact:.4f}, "
          f"byte vs exact (all pixels) {r_fields:.2e}")
    assert r_byte <= b_truth and r_exact <= b_truth
    assert abs(r_byte - r_exact) <= 0.005
    assert r_fields <= b_fields


def test_run_on_field_is_the_driver(oracle):
    """The stage composition on the jump-flood field equals rco_run's
    jump-flood mode, and on the exact transform's field its default mode."""
    W, H, L = 97, 61, 5
    d, _, e = oracle.gbuffer(3, W, H, 16, 6)
    init = oracle.seed_init(d, e)
    sdf = oracle.sdf_f64(oracle.jfa(init))
    assert np.array_equal(oracle.run_on_field(sdf, e, L), oracle.run(d, e, L, np.float64, sdf="jfa"))
    sdf = oracle.sdf_from_d2(oracle.edt_d2(init), np.float64)
    assert np.array_equal(oracle.run_on_field(sdf, e, L), oracle.run(d, e, L, np.float64))


def test_exact_field_against_brute_force_edt(oracle):
    d, _, e = oracle.gbuffer(7, 64, 48, 8, 2)
    s0 = oracle.seed_init(d, e)
    near = oracle.nearest_seed_bruteforce(s0)
    yy, xx = np.mgrid[0:48, 0:64]
    d2 = (((near >> 16).astype(np.int64) - yy) ** 2 + ((near & 0xFFFF).astype(np.int64) - xx) ** 2)
    assert np.array_equal(oracle.edt_d2_exact(s0), d2)
