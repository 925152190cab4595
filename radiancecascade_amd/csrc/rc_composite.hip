// rc_composite.hip -- the composite after the GI pass: the frame the
// reference renderer presents.  Its forward pass writes the direct term
// (Blinn-Phong times surface colour, src/shader.wgsl:76-101) into a
// Bgra8UnormSrgb surface (src/window/app.rs:59); with GI the presented colour
// is, per channel, in linear fp32 with this operation order,
//   c = (albedo * irradiance + emissive) + direct
// (direct optional), encoded to 8-bit sRGB and stored as B, G, R, A = 255.
//
// Encoding: u8(c) = the largest k in 1..255 with c >= T_k (0 if none), where
// T_k is the linear value whose sRGB code is (k - 1/2) / 255 (IEC 61966-2-1
// decode, in double, rounded to fp32).  That is round-to-nearest of the sRGB
// code, with c <= 0 -> 0 and c >= 1 -> 255, and it is exact (no pow on the
// GPU): an 8-step binary search over the 256-entry table in LDS.
#include <math.h>

#include "rc_internal.hpp"

namespace rc {

void srgb_thresholds(float *t)
{
    t[0] = -INFINITY;  // unused: the search starts at 0 and only tests 1..255
    for (int k = 1; k < 256; ++k) {
        const double s = ((double)k - 0.5) / 255.0;
        const double lin = s <= 0.04045 ? s / 12.92 : pow((s + 0.055) / 1.055, 2.4);
        t[k] = (float)lin;
    }
}

__device__ inline uint32_t srgb8(float c, const float *__restrict__ t)
{
    uint32_t u = 0;
#pragma unroll
    for (uint32_t s = 128; s >= 1; s >>= 1) u = c >= t[u + s] ? u + s : u;
    return u;
}

// One pixel per lane; the table is loaded into LDS per workgroup.
__global__ __launch_bounds__(kBlock) void k_composite(const float *__restrict__ albedo, const float *__restrict__ emissive,
                                                      const float *__restrict__ irr, const float *__restrict__ direct,
                                                      const float *__restrict__ table, uint32_t *__restrict__ out,
                                                      size_t n)
{
    __shared__ float t[256];
    t[threadIdx.x] = table[threadIdx.x];
    __syncthreads();
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    uint32_t u[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const size_t q = (size_t)ch * n + p;
        float c = albedo[q] * irr[q] + emissive[q];
        if (direct) c = c + direct[q];
        u[ch] = srgb8(c, t);
    }
    out[p] = u[2] | (u[1] << 8) | (u[0] << 16) | 0xFF000000u;  // bytes B, G, R, A
}

hipError_t launch_composite(const float *albedo, const float *emissive, const float *irr, const float *direct,
                            const float *table, uint32_t *out, size_t n, hipStream_t s)
{
    static_assert(kBlock == 256, "one table entry per thread");
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_composite, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, albedo, emissive,
                       irr, direct, table, out, n);
    return hipGetLastError();
}

}  // namespace rc
