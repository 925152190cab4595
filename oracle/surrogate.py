"""FP64 reference of the learned far field (near distance field, emitter /
occluder mip pyramid, cone trace, MLP features and the predicted level 2).

TEST INFRASTRUCTURE ONLY: the checker for the GPU kernels of
radiancecascade_amd/csrc/rc_surrogate.hip; the product never imports it.

The reference repository has no global illumination and no learned part
(its only pass is the forward raster, src/renderer.rs:559-613); the rules
restated here are the project's own and are pinned in
tests/test_oracle_surrogate.py by a library EDT (scipy), block means and
row / column fractions computed directly from pixels (exact for lines and
single pixels, a lower bound elsewhere), closed forms of the cone trace
(empty scene, fully emissive scene, a line across the cone, linearity,
mirror symmetry); the composition is the plain merge of oracle/rco.c itself.
"""
from __future__ import annotations

import math

import numpy as np

from . import mlp, rco

NEAR_R = 10                      # exact distances up to the level-1 interval end
N_DIRS = 16                      # level-1 directions = cones per level-2 probe
CONE_T0 = 10.0                   # the start of level 2's interval
CONE_GROW = 0.75
CONE_SPLIT = 60.0                # near segment from level-2 probes, far from level-3
CONE_OCC_GAIN = 3.0              # occluder exponent relative to the emitter one
T_MIN = 1e-4


# ---- near distance field (rc_surrogate.hip k_near_sdf) -------------------------
def near_sdf(seed_init_field: np.ndarray) -> np.ndarray:
    """float32 [H][W]: the distance value (floor of the distance, rco's
    conservative code) of the nearest occupied pixel where d^2 <= 100, the far
    value 2160 elsewhere (brute force over the 21x21 window)."""
    occ = seed_init_field != rco.SEED_NONE
    H, W = occ.shape
    best = np.full((H, W), NEAR_R * NEAR_R + 1, np.int64)
    pad = np.zeros((H + 2 * NEAR_R, W + 2 * NEAR_R), bool)
    pad[NEAR_R:NEAR_R + H, NEAR_R:NEAR_R + W] = occ
    for dy in range(-NEAR_R, NEAR_R + 1):
        for dx in range(-NEAR_R, NEAR_R + 1):
            d2 = dx * dx + dy * dy
            if d2 > NEAR_R * NEAR_R:
                continue
            hit = pad[NEAR_R + dy:NEAR_R + dy + H, NEAR_R + dx:NEAR_R + dx + W]
            best = np.where(hit & (d2 < best), d2, best)
    out = np.full((H, W), rco.SDF_FAR, np.float32)
    ok = best <= NEAR_R * NEAR_R
    out[ok] = np.floor(np.sqrt(best[ok].astype(np.float64)) + 1e-9).astype(np.float32)
    return out


# ---- mip pyramid (k_seed_surr, k_mip_mid, k_mip_top) -------------------------------------------
def top_level(W: int, H: int) -> int:
    top = 1
    while (1 << top) < W or (1 << top) < H:
        top += 1
    return min(top, 14)


def pixel_fields(seed_init_field: np.ndarray, emissive: np.ndarray):
    """(E*[emitter] [3,H,W] fp32, [emitter] [H,W] bool, [occupied and not emitter] [H,W] bool)."""
    e = np.asarray(emissive, np.float32)
    em = (e > 0).any(axis=0)
    occ = seed_init_field != rco.SEED_NONE
    return e * em[None], em, occ & ~em


def _f16(a: np.ndarray) -> np.ndarray:
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)


def _quads(a: np.ndarray):
    """The 2x2 children (x0y0, x1y0, x0y1, x1y1) of every parent cell, zero outside."""
    h, w = a.shape[-2:]
    h2, w2 = (h + 1) // 2, (w + 1) // 2
    a = np.pad(a, [(0, 0)] * (a.ndim - 2) + [(0, 2 * h2 - h), (0, 2 * w2 - w)])
    return a[..., 0::2, 0::2], a[..., 0::2, 1::2], a[..., 1::2, 0::2], a[..., 1::2, 1::2]


def mip(seed_init_field: np.ndarray, emissive: np.ndarray):
    """levels[m] for m = 1..top: (emission [h][w][4], coverage [h][w][4]).

    Per cell of 2^m x 2^m pixels: emission = (mean E*[emitter] rgb, emitter
    area fraction); coverage = (ex, ey, ox, oy), the fraction of the cell's
    pixel columns (x) and rows (y) holding an emitter (e) or a dark occluder
    (o) pixel.  A pixel is a level-0 cell with all of these 0 or 1.  A parent
    is computed in fp32 from its 4 stored children TL, TR, BL, BR:
        means  ((TL + TR) + (BL + BR)) / 4
        x      (max(TL, BL) + max(TR, BR)) / 2   (a column of the left half
                                                  is held if either child holds it)
        y      (max(TL, TR) + max(BL, BR)) / 2
    and every value is stored as fp16 (round to nearest even).  The max is a
    lower bound of the union, exact when the children's held columns nest
    (one line, one block)."""
    H, W = seed_init_field.shape
    EM, em, oc = pixel_fields(seed_init_field, emissive)
    f = np.concatenate([EM, em[None], em[None], em[None], oc[None], oc[None]]).astype(np.float32)
    q, h = np.float32(0.25), np.float32(0.5)
    levels = [None]
    for _ in range(1, top_level(W, H) + 1):
        TL, TR, BL, BR = _quads(f)
        mean = q * ((TL[:4] + TR[:4]) + (BL[:4] + BR[:4]))
        x = h * (np.maximum(TL[[4, 6]], BL[[4, 6]]) + np.maximum(TR[[4, 6]], BR[[4, 6]]))
        y = h * (np.maximum(TL[[5, 7]], TR[[5, 7]]) + np.maximum(BL[[5, 7]], BR[[5, 7]]))
        f = _f16(np.concatenate([mean, x[0:1], y[0:1], x[1:2], y[1:2]]))
        levels.append((np.moveaxis(f[:4], 0, -1).astype(np.float64), np.moveaxis(f[4:], 0, -1).astype(np.float64)))
    return levels


# ---- cone trace (k_cone) ------------------------------------------------------
def cone_schedule(W: int, H: int):
    """[(t, m, re, ro)] as the host builds it (double, then rounded to fp32):
    re = len / 2^m and ro = CONE_OCC_GAIN * len / 2^m are the emitter and
    occluder exponents of the step, len = min(step, 2^m)."""
    top = top_level(W, H)
    tmax = math.sqrt(float(W) * W + float(H) * H) + 2.0
    tn = math.tan(math.pi / 16.0)
    t, out = CONE_T0, []
    while t < tmax and len(out) < 64:
        dia = max(2.0 * t * tn, 1.0)
        m = min(max(int(math.floor(math.log2(dia))), 1), top)
        step = CONE_GROW * dia
        cell = float(1 << m)
        r = min(step, cell) / cell
        out.append((float(np.float32(t)), m, float(np.float32(r)), float(np.float32(CONE_OCC_GAIN * r))))
        t += step
    return out


def _sample(level, x, y, m):
    """Bilinear sample of level m at pixel positions (x, y); cells outside the
    level are zero (the stored one-cell border).  Returns (emission [..,4],
    coverage [..,4], inside) where `inside` is false once the footprint has
    left the image (lower cell index outside [-1, w-1] x [-1, h-1]; the cone
    then stops)."""
    em, cov = level
    h, w = em.shape[:2]
    pe = np.zeros((h + 2, w + 2, 4))
    pe[1:-1, 1:-1] = em
    pc = np.zeros((h + 2, w + 2, 4))
    pc[1:-1, 1:-1] = cov
    s = float(1 << m)
    fx = x / s - 0.5
    fy = y / s - 0.5
    x0 = np.floor(fx).astype(np.int64)
    y0 = np.floor(fy).astype(np.int64)
    ax, ay = fx - x0, fy - y0
    inside = (x0 >= -1) & (x0 < w) & (y0 >= -1) & (y0 < h)
    e = np.zeros(x.shape + (4,))
    c = np.zeros(x.shape + (4,))
    for dy, wy in ((0, 1 - ay), (1, ay)):
        for dx, wx in ((0, 1 - ax), (1, ax)):
            xc, yc = np.clip(x0 + 1 + dx, 0, w + 1), np.clip(y0 + 1 + dy, 0, h + 1)
            wt = (wx * wy)[..., None]
            e += pe[yc, xc] * wt
            c += pc[yc, xc] * wt
    return e, c, inside


def cone_trace_grid(levels, W: int, H: int, level: int, schedule) -> np.ndarray:
    """[QY][QX][16][4] cones from the probes of `level` (spacing 2^(level+1))
    along the given schedule: radiance rgb and transmittance.

    Per sample (bilinear in mip level m): the fraction of the cone's parallel
    rays that a cell stops is taken from its column and row fractions with
    the direction's weights, c = dy^2 x + dx^2 y (a line across the cone stops
    it whatever its area); emitters and occluders give opacities
    ae = 1 - (1 - min(ce, .999))^re and ao = 1 - (1 - min(co, .999))^ro, a =
    min(ae + ao, 1); L += T ae Ebar with Ebar = E-mean / emitter area (the
    mean radiance of the footprint's emitter pixels), T *= 1 - a.  A cone
    stops once T < 1e-4 or once its footprint has left the image."""
    QX, QY = rco.probe_grid(W, H, level)
    sp = float(2 << level)
    dirs = rco.dir_table_f32(1).astype(np.float64)
    qy, qx = np.mgrid[0:QY, 0:QX]
    cx = (qx + 0.5) * sp
    cy = (qy + 0.5) * sp
    out = np.zeros((QY, QX, N_DIRS, 4))
    for k in range(N_DIRS):
        dx, dy = dirs[k]
        T = np.ones(cx.shape)
        L = np.zeros(cx.shape + (3,))
        live = np.ones(cx.shape, bool)
        for t, m, re, ro in schedule:
            e, c, inside = _sample(levels[m], t * dx + cx, t * dy + cy, m)
            live = live & inside
            ce = dy * dy * c[..., 0] + dx * dx * c[..., 1]
            co = dy * dy * c[..., 2] + dx * dx * c[..., 3]
            ae = 1.0 - (1.0 - np.minimum(ce, 0.999)) ** re
            ao = 1.0 - (1.0 - np.minimum(co, 0.999)) ** ro
            a = np.minimum(ae + ao, 1.0)
            ebar = np.where(e[..., 3:4] > 0, e[..., :3] / np.where(e[..., 3:4] > 0, e[..., 3:4], 1.0), 0.0)
            L = np.where(live[..., None], L + (T * ae)[..., None] * ebar, L)
            T = np.where(live, T * (1.0 - a), T)
            live = live & (T >= T_MIN)
        out[:, :, k, :3] = L
        out[:, :, k, 3] = T
    return out


def cone_trace(levels, W: int, H: int, schedule=None) -> np.ndarray:
    """[QY][QX][16][4] at the level-2 probes: the near segment (t < 60) traced
    from them, then the far segment traced from the level-3 probes and
    appended with the merge's bilinear weights: L += T * F.rgb, T *= F.T."""
    schedule = schedule if schedule is not None else cone_schedule(W, H)
    near = [s for s in schedule if s[0] < CONE_SPLIT]
    far = [s for s in schedule if s[0] >= CONE_SPLIT]
    c2 = cone_trace_grid(levels, W, H, 2, near)
    c3 = cone_trace_grid(levels, W, H, 3, far)
    QX, QY = rco.probe_grid(W, H, 2)
    F = upper_bilinear(c3, QX, QY)
    FT = upper_bilinear(np.repeat(c3[..., 3:4], 3, -1), QX, QY)[..., 0]
    out = c2.copy()
    out[..., :3] = c2[..., :3] + c2[..., 3:4] * F
    out[..., 3] = c2[..., 3] * FT
    return out


# ---- features and the predicted level 2 (k_surr_mlp, k_surr_sample) ------------
def _axis(n: int, un: int):
    p = np.arange(n)
    base = (p - 1) >> 1
    w1 = np.where(p & 1, 0.25, 0.75)
    return np.clip(base, 0, un - 1), np.clip(base + 1, 0, un - 1), 1.0 - w1, w1


def upper_bilinear(U: np.ndarray, PX: int, PY: int) -> np.ndarray:
    """[PY][PX][16][3]: entries of the level-2 grid array U [QY][QX][16][>=3]
    interpolated at the level-1 probes with the merge's weights."""
    QY, QX = U.shape[:2]
    x0, x1, wx0, wx1 = _axis(PX, QX)
    y0, y1, wy0, wy1 = _axis(PY, QY)
    A = U[..., :3]
    r0, r1 = A[y0], A[y1]
    top = r0[:, x0] * wx0[None, :, None, None] + r0[:, x1] * wx1[None, :, None, None]
    bot = r1[:, x0] * wx0[None, :, None, None] + r1[:, x1] * wx1[None, :, None, None]
    return top * wy0[:, None, None, None] + bot * wy1[:, None, None, None]


def features(cone: np.ndarray):
    """(c [Q][48] cone radiance at the level-2 probes, x [Q][64] = per
    direction (log(1 + c) rgb, cone transmittance T))."""
    QY, QX = cone.shape[:2]
    c = cone[..., :3].reshape(QY * QX, 48)
    x = np.concatenate([np.log1p(cone[..., :3]), cone[..., 3:4]], -1).reshape(QY * QX, 64)
    return c, x


def samples(cone: np.ndarray, A2: np.ndarray):
    """Training rows per level-2 probe: (features, A2.rgb - c) with A2 the
    plain path's merged level 2 (16 entries per probe)."""
    c, x = features(cone)
    return x, A2[..., :3].reshape(c.shape[0], 48) - c


def predicted_level2(c: np.ndarray, z: np.ndarray, out_scale: float, QX: int, QY: int) -> np.ndarray:
    """The network's merged level 2 [QY][QX][16][4]: rgb = max(c + s z, 0),
    4th channel 1 (level 1 reads only rgb of its upper level)."""
    A = np.maximum(c + out_scale * z[:, :48], 0.0).reshape(QY, QX, 16, 3)
    return np.concatenate([A, np.ones((QY, QX, 16, 1))], -1)


def params_to_layers(params: np.ndarray):
    """Flat RC_MLP_PARAMS vector -> [[W_l [K][64], b_l [64]]] (oracle/mlp.py layout)."""
    p = np.asarray(params, np.float64).reshape(-1)
    out, off = [], 0
    for k, n in mlp.SIZES:
        W = p[off:off + k * n].reshape(k, n)
        off += k * n
        b = p[off:off + n]
        off += n
        out.append([W, b])
    assert off == p.size
    return out


def layers_to_params(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b.reshape(-1)]) for W, b in layers])


def run(depth, emissive, L, params, in_mean, in_istd, out_scale, want=False):
    """Whole learned-far-field frame in FP64: the network's level 2 from the
    cone features, then the plain merge of levels 1 and 0 (marched on the
    near distance field) and the gather; the network evaluated with the GPU's
    bf16 rounding points.  Returns irradiance [3,H,W] (and intermediates)."""
    H, W = depth.shape
    seed0 = rco.seed_init(depth, emissive)
    sdf = near_sdf(seed0).astype(np.float64)
    m1, _ = rco.march_level(sdf, emissive, 1)
    m0, _ = rco.march_level(sdf, emissive, 0)
    levels = mip(seed0, emissive)
    cone = cone_trace(levels, W, H)
    c, x = features(cone)
    xn = (x - np.asarray(in_mean, np.float64)) * np.asarray(in_istd, np.float64)
    z, _ = mlp.forward(params_to_layers(params), xn, emulate_bf16=True)
    QX, QY = rco.probe_grid(W, H, 2)
    A2 = predicted_level2(c, z, out_scale, QX, QY)
    A1 = rco.merge_level(W, H, 1, m1, A2)
    A0 = rco.merge_level(W, H, 0, m0, A1)
    irr = rco.gather(W, H, A0)
    if want:
        return irr, dict(sdf=sdf, m1=m1, m0=m0, levels=levels, cone=cone, c=c, x=x, z=z, A2=A2, A1=A1)
    return irr
