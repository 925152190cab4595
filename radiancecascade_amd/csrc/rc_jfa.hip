// rc_jfa.hip -- occupancy seeds, jump flood and the distance field.
//
// The reference has no distance field (its only pass is the forward raster,
// src/renderer.rs:559-613); this follows Rong & Tan, "Jump Flooding in GPU"
// (I3D 2006) with a total-order tie-break so the result is independent of
// evaluation order and therefore bit-exact across implementations:
//   seed(p)  = (y << 16) | x if p is occupied, else kSeedNone;
//   occupied = any emissive channel > 0, or a 4-neighbour q with
//              D(q) - D(p) > 0.05 D(p) in fp32 (near side of a silhouette);
//   pass(s)  : p takes the candidate with the smallest (d^2, seed_y, seed_x)
//              among its own seed and the in-bounds seeds at p + s*(dx,dy);
//   steps    : 2^(P-1), ..., 2, 1 with P = ceil(log2 max(W,H));
//   SDF(p)   = (float)sqrt((double)d^2), +inf when no pixel is occupied.
// Passes with step > 8 run as one global kernel each (every pixel gathers 9
// seeds, rows coalesce); steps 8,4,2,1 and the SDF are fused in one LDS-tiled
// kernel whose tile carries a 15-pixel halo (8+4+2+1), so the short passes
// never round-trip through HBM.
#include "rc_internal.hpp"

namespace rc {

// Workgroups are handed to the 8 XCDs round-robin; give each XCD one
// contiguous run of logical block ids so neighbouring tiles share its L2.
__device__ inline int xcd_remap(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return xcd * q + min(xcd, r) + idx;
}

__device__ inline uint64_t jfa_key(uint32_t s, int x, int y)
{
    if (s == kSeedNone) return ~0ull;
    const int dx = (int)(s & 0xFFFFu) - x, dy = (int)(s >> 16) - y;
    return ((uint64_t)(uint32_t)(dx * dx + dy * dy) << 32) | s;
}

int jfa_num_passes(int W, int H)
{
    int m = W > H ? W : H, P = 0;
    while ((1 << P) < m) ++P;
    return P;
}

int jfa_tail_step(int W, int H)
{
    int P = jfa_num_passes(W, H);
    if (P == 0) return 0;
    int top = 1 << (P - 1);
    return top < 8 ? top : 8;
}

// 64 x 4 pixel blocks, x fastest so each wave covers 64 consecutive pixels.
__global__ __launch_bounds__(kBlock) void k_seed_init(const float *__restrict__ depth,
                                                      const float *__restrict__ emissive,
                                                      uint32_t *__restrict__ seed, int W, int H, int bx_count)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int x = (b % bx_count) * 64 + (threadIdx.x & 63);
    const int y = (b / bx_count) * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const size_t plane = (size_t)W * H, p = (size_t)y * W + x;
    bool occ = emissive[p] > 0.0f || emissive[plane + p] > 0.0f || emissive[2 * plane + p] > 0.0f;
    const float dp = depth[p];
    const float thr = 0.05f * dp;
    if (!occ && x > 0) occ = depth[p - 1] - dp > thr;
    if (!occ && x + 1 < W) occ = depth[p + 1] - dp > thr;
    if (!occ && y > 0) occ = depth[p - W] - dp > thr;
    if (!occ && y + 1 < H) occ = depth[p + W] - dp > thr;
    seed[p] = occ ? (((uint32_t)y << 16) | (uint32_t)x) : kSeedNone;
}

__global__ __launch_bounds__(kBlock) void k_jfa_pass(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                     int W, int H, int step, int bx_count)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int x = (b % bx_count) * 64 + (threadIdx.x & 63);
    const int y = (b / bx_count) * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    uint64_t best = jfa_key(in[(size_t)y * W + x], x, y);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
        const int qy = y + dy * step;
        if (qy < 0 || qy >= H) continue;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            const int qx = x + dx * step;
            if ((dx == 0 && dy == 0) || qx < 0 || qx >= W) continue;
            const uint64_t k = jfa_key(in[(size_t)qy * W + qx], x, y);
            best = k < best ? k : best;
        }
    }
    out[(size_t)y * W + x] = best == ~0ull ? kSeedNone : (uint32_t)best;
}

// Fused passes S, S/2, ..., 1 over a 32x32 output tile.  The tile is loaded
// with a halo of 2S-1 pixels; each pass shrinks the valid region by its step.
// Cells outside the image stay kSeedNone in every pass, which is exactly the
// global kernel's "out-of-bounds neighbours are skipped".
template <int S>
__global__ __launch_bounds__(kBlock) void k_jfa_tail(const uint32_t *__restrict__ in,
                                                     uint32_t *__restrict__ seeds_out, float *__restrict__ sdf,
                                                     int W, int H, int bx_count)
{
    constexpr int T = 32;
    constexpr int HALO = S > 0 ? 2 * S - 1 : 0;
    constexpr int R = T + 2 * HALO;
    __shared__ uint32_t buf[2][R * R];

    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int x0 = (b % bx_count) * T - HALO;
    const int y0 = (b / bx_count) * T - HALO;
    for (int i = threadIdx.x; i < R * R; i += kBlock) {
        const int gx = x0 + i % R, gy = y0 + i / R;
        buf[0][i] = (gx >= 0 && gy >= 0 && gx < W && gy < H) ? in[(size_t)gy * W + gx] : kSeedNone;
    }
    __syncthreads();

    int cur = 0;
    int off = 0;
#pragma unroll
    for (int s = S; s >= 1; s >>= 1) {
        off += s;
        const int n = R - 2 * off;
        for (int i = threadIdx.x; i < n * n; i += kBlock) {
            const int rx = off + i % n, ry = off + i / n;
            const int gx = x0 + rx, gy = y0 + ry;
            uint32_t res = kSeedNone;
            if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
                uint64_t best = jfa_key(buf[cur][ry * R + rx], gx, gy);
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (dx == 0 && dy == 0) continue;
                        const uint64_t k = jfa_key(buf[cur][(ry + dy * s) * R + rx + dx * s], gx, gy);
                        best = k < best ? k : best;
                    }
                res = best == ~0ull ? kSeedNone : (uint32_t)best;
            }
            buf[cur ^ 1][ry * R + rx] = res;
        }
        __syncthreads();
        cur ^= 1;
    }

    for (int i = threadIdx.x; i < T * T; i += kBlock) {
        const int gx = x0 + HALO + i % T, gy = y0 + HALO + i / T;
        if (gx >= W || gy >= H) continue;
        const uint32_t sd = buf[cur][(HALO + i / T) * R + HALO + i % T];
        float d = __builtin_huge_valf();
        if (sd != kSeedNone) {
            const int dx = (int)(sd & 0xFFFFu) - gx, dy = (int)(sd >> 16) - gy;
            d = (float)sqrt((double)(uint32_t)(dx * dx + dy * dy));
        }
        sdf[(size_t)gy * W + gx] = d;
        if (seeds_out) seeds_out[(size_t)gy * W + gx] = sd;
    }
}

hipError_t launch_seed_init(const float *depth, const float *emissive, uint32_t *seed, int W, int H, hipStream_t s)
{
    const int bxc = (W + 63) / 64, byc = (H + 3) / 4;
    hipLaunchKernelGGL(k_seed_init, dim3(bxc * byc), dim3(kBlock), 0, s, depth, emissive, seed, W, H, bxc);
    return hipGetLastError();
}

hipError_t launch_jfa_pass(const uint32_t *in, uint32_t *out, int W, int H, int step, hipStream_t s)
{
    const int bxc = (W + 63) / 64, byc = (H + 3) / 4;
    hipLaunchKernelGGL(k_jfa_pass, dim3(bxc * byc), dim3(kBlock), 0, s, in, out, W, H, step, bxc);
    return hipGetLastError();
}

hipError_t launch_jfa_tail(const uint32_t *in, uint32_t *seeds_out, float *sdf, int W, int H, hipStream_t s)
{
    const int bxc = (W + 31) / 32, byc = (H + 31) / 32;
    const dim3 grid(bxc * byc);
    switch (jfa_tail_step(W, H)) {
    case 0: hipLaunchKernelGGL(k_jfa_tail<0>, grid, dim3(kBlock), 0, s, in, seeds_out, sdf, W, H, bxc); break;
    case 1: hipLaunchKernelGGL(k_jfa_tail<1>, grid, dim3(kBlock), 0, s, in, seeds_out, sdf, W, H, bxc); break;
    case 2: hipLaunchKernelGGL(k_jfa_tail<2>, grid, dim3(kBlock), 0, s, in, seeds_out, sdf, W, H, bxc); break;
    case 4: hipLaunchKernelGGL(k_jfa_tail<4>, grid, dim3(kBlock), 0, s, in, seeds_out, sdf, W, H, bxc); break;
    default: hipLaunchKernelGGL(k_jfa_tail<8>, grid, dim3(kBlock), 0, s, in, seeds_out, sdf, W, H, bxc); break;
    }
    return hipGetLastError();
}

}  // namespace rc
