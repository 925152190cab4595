// rc_common.hpp -- shared constants, cascade geometry and storage types for
// the gfx950 radiance-cascade kernels.
//
// Geometry (all integer): level i has probe spacing s_i = 2*2^i px,
// N_i = 4*4^i directions, a ceil-sized probe grid Px_i x Py_i and the ray
// interval [2(4^i-1)/3, 2(4^(i+1)-1)/3) px.  Probe (px,py) is centred at
// ((px+1/2) s_i, (py+1/2) s_i); direction k is theta = 2 pi (k+1/2)/N_i and its
// children at level i+1 are 4k..4k+3.
//
// Materialised merged level i (i >= 1) is stored probe-major with the
// direction innermost: A_i[py][px][k'][texel] for k' < N_{i-1} = N_i/4, where
// A_i(p,k') = 1/4 sum_j R_i(p, 4k'+j) is the merged radiance already averaged
// over the 4 children the level below reads.  A probe's texels are one
// contiguous run, so a wave's reads and writes over consecutive directions
// coalesce.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rc {

constexpr uint32_t kSeedNone = 0xFFFFFFFFu;
constexpr int kBlock = 256;  // threads per workgroup for every kernel (4 wave64)

__host__ __device__ inline int probe_spacing(int level) { return 2 << level; }
__host__ __device__ inline int num_dirs(int level) { return 4 << (2 * level); }
__host__ __device__ inline int grid_dim(int pixels, int level)
{
    int s = probe_spacing(level);
    return (pixels + s - 1) / s;
}
__host__ __device__ inline int64_t interval_t0(int level) { return (2 * ((1ll << (2 * level)) - 1)) / 3; }
__host__ __device__ inline int64_t interval_t1(int level) { return (2 * ((1ll << (2 * level + 2)) - 1)) / 3; }

// ---- distance field: one byte per pixel, 16-pixel-wide strips --------------
// The value stored is a conservative code of the distance d = sqrt(d^2) (in
// px) to the nearest occupied pixel:
//   code = floor(d)                        d < 128
//   code = 128 + floor((d - 128) / 16)     128 <= d < 2160
//   code = 255                             d >= 2160, or no occupied pixel
// and the march reads value(code) = code < 128 ? code : 128 + 16 (code - 128)
// <= d, so a step never passes a surface; value 0 (a hit) is exactly d = 0.
// At 1 byte per pixel the field of a 1920x1080 frame is 2 MB and stays in
// each XCD's 4 MB L2 while the long rays of the upper levels sample it.  The
// bytes are stored in vertical strips 16 pixels wide, each padded to a
// multiple of 8 rows, one row of a strip every 16 bytes: an aligned 128-byte
// line holds a 16x8-pixel block, so the samples of a ray and of neighbouring
// rays share lines, and the byte of (x, y) is at x + 16 y + (x >> 4) * skip
// with skip = 16 * rows - 16 (three instructions per sample in the march).
constexpr uint32_t kSdfFar = 255;
constexpr float kSdfFarValue = 2160.0f;
__host__ __device__ inline uint32_t sdf_code_of_distance(uint32_t d)  // d = floor of the distance
{
    return d < 128 ? d : (d < 2160 ? 128 + ((d - 128) >> 4) : kSdfFar);
}
__host__ __device__ inline float sdf_value(uint32_t code)
{
    return code < 128 ? (float)code : (float)(128 + 16 * (code - 128));
}
// floor(sqrt(d2)) for d2 < 2^31: fp32 estimate, corrected to the exact integer
__host__ __device__ inline uint32_t isqrt_u32(uint32_t d2)
{
    uint32_t r = (uint32_t)sqrtf((float)d2);
    while ((uint64_t)r * r > d2) --r;
    while ((uint64_t)(r + 1) * (r + 1) <= d2) ++r;
    return r;
}
__host__ __device__ inline int sdf_strips(int W) { return (W + 15) >> 4; }
__host__ __device__ inline int sdf_rows(int H) { return (H + 7) & ~7; }
__host__ __device__ inline size_t sdf_elems(int W, int H) { return (size_t)sdf_strips(W) * 16 * (size_t)sdf_rows(H); }
// the strip-to-strip step less the 16 bytes x itself advances per strip
__host__ __device__ inline int sdf_skip(int H) { return 16 * sdf_rows(H) - 16; }
__host__ __device__ inline int sdf_index(int x, int y, int skip) { return x + 16 * y + (x >> 4) * skip; }

// ---- storage texels ---------------------------------------------------------
// fp32: float4.  fp16 / bf16: four 16-bit lanes packed in a uint2 (8 bytes).
enum class Storage : int { F32 = 0, F16 = 1, BF16 = 2 };

template <Storage S> struct Texel;

template <> struct Texel<Storage::F32> {
    using T = float4;
    __device__ static inline T pack(float4 v) { return v; }
    __device__ static inline float4 unpack(T v) { return v; }
};

// fp32 -> bf16, round to nearest even: one v_cvt_pk_bf16_f32 on gfx950.
__device__ inline uint32_t f32_to_bf16_rne(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ inline float bf16_to_f32(uint32_t h) { return __uint_as_float(h << 16); }

__device__ inline uint32_t f32_to_f16_rne(float f)
{
    _Float16 h = (_Float16)f;  // round to nearest even
    return (uint32_t)__builtin_bit_cast(uint16_t, h);
}
__device__ inline float f16_to_f32(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(h & 0xFFFFu)); }

template <> struct Texel<Storage::F16> {
    using T = uint2;
    __device__ static inline T pack(float4 v)
    {
        return make_uint2(f32_to_f16_rne(v.x) | (f32_to_f16_rne(v.y) << 16),
                          f32_to_f16_rne(v.z) | (f32_to_f16_rne(v.w) << 16));
    }
    __device__ static inline float4 unpack(T v)
    {
        return make_float4(f16_to_f32(v.x), f16_to_f32(v.x >> 16), f16_to_f32(v.y), f16_to_f32(v.y >> 16));
    }
};

template <> struct Texel<Storage::BF16> {
    using T = uint2;
    __device__ static inline T pack(float4 v)
    {
        return make_uint2(f32_to_bf16_rne(v.x) | (f32_to_bf16_rne(v.y) << 16),
                          f32_to_bf16_rne(v.z) | (f32_to_bf16_rne(v.w) << 16));
    }
    __device__ static inline float4 unpack(T v)
    {
        return make_float4(bf16_to_f32(v.x & 0xFFFFu), bf16_to_f32(v.x >> 16), bf16_to_f32(v.y & 0xFFFFu),
                           bf16_to_f32(v.y >> 16));
    }
};

inline size_t texel_bytes(Storage s) { return s == Storage::F32 ? 16 : 8; }

}  // namespace rc
