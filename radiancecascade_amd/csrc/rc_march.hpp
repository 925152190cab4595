// rc_march.hpp -- device helpers shared by the cascade kernels and the
// surrogate path: lane-quad sums, the distance-field load, the interval
// march and the bilinear upper-probe interpolation.
#pragma once

#include "rc_internal.hpp"

namespace rc {

// Sum over a lane quad in the order (v0+v1)+(v2+v3).  Every lane of the quad
// ends with the same bits (fp addition is commutative).
__device__ inline float quad_sum(float v)
{
    v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));  // lane ^ 1
    v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));  // lane ^ 2
    return v;
}

// The distance field is read through a buffer resource: the strip index is a
// 32-bit lane offset (no 64-bit address arithmetic per step), and a lane whose
// sample is off-image may compute any offset -- the load is range-checked and
// its value unused.
constexpr int kBufferDword3 = 0x00020000;  // gfx9-family raw buffer, 32-bit data

__device__ inline __amdgpu_buffer_rsrc_t sdf_rsrc(const uint8_t *sdf, int W, int H)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(sdf), (short)0, (int)sdf_elems(W, H),
                                             kBufferDword3);
}

// Row-major pixel index q = y*W + x (24-bit multiply: W, y < 2^14) -- the
// hit code and the emissive index.
__device__ inline int pixel_index(int x, int y, int W) { return (int)__umul24((uint32_t)y, (uint32_t)W) + x; }

// Byte offset of (x, y) in the strip layout, sdf_index's value for an
// in-image pixel (x >> 4 < 2^10 and skip < 2^18 fit the 24-bit multiplier):
// a shift, a 24-bit multiply-add and a shift-add.  For an off-image (x, y)
// the offset may be anything; the range-checked load is then unused.
__device__ inline uint32_t sdf_offset(int x, int y, int skip)
{
    const uint32_t ux = (uint32_t)x;
    const uint32_t m = __umul24(ux >> 4, (uint32_t)skip) + ux;  // v_mad_u32_u24
    uint32_t r;                                                  // (y << 4) + m
    asm("v_lshl_add_u32 %0, %1, 4, %2" : "=v"(r) : "v"(y), "v"(m));
    return r;
}

// Distance load (one byte, decoded).
__device__ inline float sdf_load(__amdgpu_buffer_rsrc_t r, int x, int y, int skip)
{
    return sdf_value(__builtin_amdgcn_raw_buffer_load_b8(r, (int)sdf_offset(x, y, skip), 0, 0));
}

// The raw byte code at (x, y) (same offset and range check as sdf_load).
__device__ inline uint32_t sdf_code_load(__amdgpu_buffer_rsrc_t r, int x, int y, int skip)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, (int)sdf_offset(x, y, skip), 0, 0);
}

// march_step(sdf_value(c)) for a code c >= 1 (c = 0 is a hit and takes no
// step): sdf_value is c below 128 and 16 c - 1920 from 128 on, and
// 16 c - 1920 < c exactly when c < 128, so the value is max(c, 16 c - 1920)
// (exact in fp32), already >= 1.
__device__ inline float step_of_code(uint32_t c)
{
    const float f = (float)c;
    return fmaxf(f, fmaf(f, 16.0f, -1920.0f));
}

// floor(v) as an int in one VALU instruction (v_cvt_flr_i32_f32; v is a
// finite sample position here, far inside the int range).
__device__ inline int floor_i32(float v)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// max(d, 1) as one v_med3 (no NaN canonicalisation: d is never NaN).  For
// d = +inf it gives FLT_MAX instead of inf; either ends the interval.
__device__ inline float march_step(float d) { return __builtin_amdgcn_fmed3f(d, 1.0f, 3.4028235e38f); }

// Ray codes: >= 0 hit pixel y*W+x, -1 left the image, -2 reached t1.
constexpr int kRayExit = -1, kRayT1 = -2;

// RPT rays per lane marched in lockstep: each iteration issues the RPT
// distance-field loads back to back, so a lane keeps RPT independent load
// chains in flight (the march is a chain of dependent gathers and is bound by
// their latency, not by bandwidth).  A finished ray's slot keeps issuing a
// (cached, ignored) load until every ray of the lane is done.  The body is
// branch-free on the raw byte code: t advances only on an in-image sample,
// by step_of_code (0 on a hit), so where t stops tells how the ray ended --
// at t1 if t >= t1, else at the sample at t, off-image (exit) or a hit.  The
// samples, and so the codes, are those of the one-sample-per-step march.
template <int RPT>
__device__ inline void march_rays(__amdgpu_buffer_rsrc_t sdf, int W, int H, const float *cx, const float *cy,
                                  float dx, float dy, float t0, float t1, const bool *active, int *code)
{
    const int skip = sdf_skip(H);
    float t[RPT];
    bool alive[RPT];
    bool any = false;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        t[j] = t0;
        alive[j] = active[j] && t0 < t1;
        any = any || alive[j];
    }
    while (any) {
        uint32_t c[RPT];
        bool ins[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            // positions are finite, so the integer test equals 0 <= x < W, 0 <= y < H
            const int ix = floor_i32(fmaf(t[j], dx, cx[j]));
            const int iy = floor_i32(fmaf(t[j], dy, cy[j]));
            ins[j] = (uint32_t)ix < (uint32_t)W && (uint32_t)iy < (uint32_t)H;
            c[j] = sdf_code_load(sdf, ix, iy, skip);
        }
        any = false;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const float tn = t[j] + step_of_code(c[j]);
            if (alive[j] && ins[j]) t[j] = tn;
            alive[j] = alive[j] && ins[j] && c[j] != 0 && tn < t1;
            any = any || alive[j];
        }
    }
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        code[j] = kRayT1;
        if (active[j] && t[j] < t1) {
            const int ix = floor_i32(fmaf(t[j], dx, cx[j]));
            const int iy = floor_i32(fmaf(t[j], dy, cy[j]));
            code[j] = ((uint32_t)ix < (uint32_t)W && (uint32_t)iy < (uint32_t)H) ? pixel_index(ix, iy, W) : kRayExit;
        }
    }
}

// Bilinear of four staged texels with the lower probe's parity weights.
__device__ inline float3 bilerp4(float4 u00, float4 u10, float4 u01, float4 u11, int x, int y)
{
    const float wx1 = (x & 1) ? 0.25f : 0.75f, wx0 = 1.0f - wx1;
    const float wy1 = (y & 1) ? 0.25f : 0.75f, wy0 = 1.0f - wy1;
    float3 r;
    r.x = fmaf(wy1, fmaf(wx1, u11.x, wx0 * u01.x), wy0 * fmaf(wx1, u10.x, wx0 * u00.x));
    r.y = fmaf(wy1, fmaf(wx1, u11.y, wx0 * u01.y), wy0 * fmaf(wx1, u10.y, wx0 * u00.y));
    r.z = fmaf(wy1, fmaf(wx1, u11.z, wx0 * u01.z), wy0 * fmaf(wx1, u10.z, wx0 * u00.z));
    return r;
}

}  // namespace rc
