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

// Inside the kernels "no seed" is a fake seed at (40000, 40000): for any
// pixel of an image up to 16384 px wide its squared distance exceeds 2^30,
// above every real candidate's (< 2 * 16383^2 < 2^30), and stays below 2^32.
// So the (d^2, seed) minimum needs no special case; the external value
// kSeedNone is restored on output.
constexpr uint32_t kFarSeed = (40000u << 16) | 40000u;

// Order key (d^2, seed_y, seed_x): the packed seed already orders (y, x).
__device__ inline uint64_t jfa_key(uint32_t s, int x, int y)
{
    const int dx = (int)(s & 0xFFFFu) - x, dy = (int)(s >> 16) - y;
    const uint32_t d2 = (uint32_t)__mul24(dx, dx) + (uint32_t)__mul24(dy, dy);
    return ((uint64_t)d2 << 32) | s;
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
    // all loads unconditional (clamped addresses) so they are in flight together
    const float e0 = emissive[p], e1 = emissive[plane + p], e2 = emissive[2 * plane + p];
    const float dp = depth[p];
    const float dl = depth[p - (x > 0)], dr = depth[p + (x + 1 < W)];
    const float du = depth[p - (y > 0 ? (size_t)W : 0)], dd = depth[p + (y + 1 < H ? (size_t)W : 0)];
    const float thr = 0.05f * dp;
    // an out-of-image neighbour reads p itself and is masked out
    const bool occ = e0 > 0.0f || e1 > 0.0f || e2 > 0.0f || (x > 0 && dl - dp > thr) ||
                     (x + 1 < W && dr - dp > thr) || (y > 0 && du - dp > thr) || (y + 1 < H && dd - dp > thr);
    seed[p] = occ ? (((uint32_t)y << 16) | (uint32_t)x) : kFarSeed;
}

__global__ __launch_bounds__(kBlock) void k_jfa_pass(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                     int W, int H, int step, int bx_count)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int x = (b % bx_count) * 64 + (threadIdx.x & 63);
    const int y = (b / bx_count) * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    // 9 unconditional loads at clamped coordinates (in flight together); a
    // candidate whose true coordinate is off-image is replaced by the far seed
    const int xs[3] = {x - step, x, x + step}, ys[3] = {y - step, y, y + step};
    uint32_t cand[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int qx = min(max(xs[j % 3], 0), W - 1), qy = min(max(ys[j / 3], 0), H - 1);
        cand[j] = in[(size_t)qy * W + qx];
    }
    uint64_t best = jfa_key(cand[4], x, y);
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        if (j == 4) continue;
        const bool inb = xs[j % 3] >= 0 && xs[j % 3] < W && ys[j / 3] >= 0 && ys[j / 3] < H;
        const uint64_t k = jfa_key(inb ? cand[j] : kFarSeed, x, y);
        best = k < best ? k : best;
    }
    out[(size_t)y * W + x] = (uint32_t)best;
}

// Fused passes S, S/2, ..., 1 over a 32x32 output tile.  The tile is loaded
// with a halo of 2S-1 pixels (R <= 62 cells a side); each pass shrinks the
// valid region by its step.  Threads are laid out 64 x 4: lane = column, so
// every LDS access of a wave is one row of consecutive words.  Cells outside
// the image stay the far seed in every pass, which is exactly the global
// kernel's "out-of-bounds neighbours are skipped".
template <int S>
__global__ __launch_bounds__(kBlock) void k_jfa_tail(const uint32_t *__restrict__ in,
                                                     uint32_t *__restrict__ seeds_out, float *__restrict__ sdf,
                                                     int W, int H, int bx_count)
{
    constexpr int T = 32;
    constexpr int HALO = S > 0 ? 2 * S - 1 : 0;
    constexpr int R = T + 2 * HALO;
    static_assert(R <= 64, "one column per lane");
    __shared__ uint32_t buf[2][R * R];

    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int x0 = (b % bx_count) * T - HALO;
    const int y0 = (b / bx_count) * T - HALO;
    const int cx = threadIdx.x & 63, cy = threadIdx.x >> 6;
    {
        // all rows of this lane's column loaded (clamped, unconditional) before
        // any LDS write, so the global loads overlap
        constexpr int ROWS = (R + 3) / 4;
        const int gx = x0 + cx, qx = min(max(gx, 0), W - 1);
        uint32_t v[ROWS];
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            const int gy = y0 + cy + 4 * j, qy = min(max(gy, 0), H - 1);
            v[j] = in[(size_t)qy * W + qx];
        }
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            const int ry = cy + 4 * j, gy = y0 + ry;
            if (cx < R && ry < R)
                buf[0][ry * R + cx] = (gx >= 0 && gy >= 0 && gx < W && gy < H) ? v[j] : kFarSeed;
        }
    }
    __syncthreads();

    int cur = 0;
    int off = 0;
#pragma unroll
    for (int s = S; s >= 1; s >>= 1) {
        off += s;
        const int rx = off + cx;
        if (rx < R - off) {
            const int gx = x0 + rx;
            const bool xin = gx >= 0 && gx < W;
            for (int ry = off + cy; ry < R - off; ry += 4) {
                const int gy = y0 + ry;
                const uint32_t *src = buf[cur];
                uint32_t res = kFarSeed;
                if (xin && gy >= 0 && gy < H) {
                    uint64_t best = jfa_key(src[ry * R + rx], gx, gy);
#pragma unroll
                    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (dx == 0 && dy == 0) continue;
                            const uint64_t k = jfa_key(src[(ry + dy * s) * R + rx + dx * s], gx, gy);
                            best = k < best ? k : best;
                        }
                    res = (uint32_t)best;
                }
                buf[cur ^ 1][ry * R + rx] = res;
            }
        }
        __syncthreads();
        cur ^= 1;
    }

    if (cx < T) {
        const int gx = x0 + HALO + cx;
        for (int ty = cy; ty < T; ty += 4) {
            const int gy = y0 + HALO + ty;
            if (gx >= W || gy >= H) continue;
            const uint32_t sd = buf[cur][(HALO + ty) * R + HALO + cx];
            const int dx = (int)(sd & 0xFFFFu) - gx, dy = (int)(sd >> 16) - gy;
            const float d = sd == kFarSeed ? __builtin_huge_valf() : (float)sqrt((double)(uint32_t)(dx * dx + dy * dy));
            sdf[(size_t)gy * W + gx] = d;
            if (seeds_out) seeds_out[(size_t)gy * W + gx] = sd == kFarSeed ? kSeedNone : sd;
        }
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
