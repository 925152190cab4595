"""Reference of the composite after the GI pass (rc_composite.hip).

TEST INFRASTRUCTURE ONLY: the checker for the GPU composite kernel; the
product never imports it.

The presented colour per pixel and channel is, in fp32 with this order,
c = (albedo * irradiance + emissive) + direct, encoded to 8-bit sRGB by the
threshold rule u8 = max{k : c >= T_k} (0 if none), T_k the linear value whose
IEC 61966-2-1 code is (k - 1/2)/255, and stored as the bytes B, G, R, A=255
(the reference presents into a Bgra8UnormSrgb surface, src/window/app.rs:59;
its direct term is src/shader.wgsl:76-101).  Pinned in
tests/test_oracle_composite.py by the sRGB closed forms (0, 1, the 0.5 code,
the linear-segment knee) and by round-to-nearest of the float64 encoding
away from the rounding boundaries.
"""
from __future__ import annotations

import numpy as np


def srgb_decode(s):
    """IEC 61966-2-1: sRGB code in [0, 1] -> linear (float64)."""
    s = np.asarray(s, np.float64)
    return np.where(s <= 0.04045, s / 12.92, ((s + 0.055) / 1.055) ** 2.4)


def srgb_encode(c):
    """Linear -> sRGB code (float64, clamped to [0, 1])."""
    c = np.clip(np.asarray(c, np.float64), 0.0, 1.0)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * c ** (1 / 2.4) - 0.055)


def thresholds() -> np.ndarray:
    """float32 [256]: T_0 = -inf, T_k = decode((k - 0.5) / 255) rounded to fp32."""
    k = np.arange(256, dtype=np.float64)
    t = srgb_decode((k - 0.5) / 255.0).astype(np.float32)
    t[0] = -np.inf
    return t


def encode8(c, t=None) -> np.ndarray:
    """uint8: the largest k with c >= T_k (searchsorted on the sorted table)."""
    t = thresholds() if t is None else t
    c = np.asarray(c, np.float32)
    u = np.searchsorted(t[1:], c, side="right")     # number of T_k (k >= 1) <= c
    return np.where(np.isnan(c), 0, u).astype(np.uint8)


def composite(albedo, emissive, irradiance, direct=None) -> np.ndarray:
    """uint32 [H][W]: B | G << 8 | R << 16 | 255 << 24 from planar [3][H][W] inputs."""
    a = np.asarray(albedo, np.float32)
    e = np.asarray(emissive, np.float32)
    i = np.asarray(irradiance, np.float32)
    c = a * i + e                                      # fp32: product rounded, then sum
    if direct is not None:
        c = c + np.asarray(direct, np.float32)
    u = encode8(c).astype(np.uint32)
    return u[2] | (u[1] << 8) | (u[0] << 16) | np.uint32(0xFF000000)
