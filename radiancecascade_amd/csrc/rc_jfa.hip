// rc_jfa.hip -- occupancy seeds, jump flood and the distance field.
//
// The reference has no distance field (its only pass is the forward raster,
// src/renderer.rs:559-613); this follows Rong & Tan, "Jump Flooding in GPU"
// (I3D 2006) with a total-order tie-break so the result is independent of
// evaluation order and therefore bit-exact across implementations:
//   seed(p)  = (y << 16) | x if p is occupied, else "none";
//   occupied = any emissive channel > 0, or a 4-neighbour q with
//              D(q) - D(p) > 0.05 D(p) in fp32 (near side of a silhouette);
//   pass(s)  : p takes the candidate with the smallest (d^2, seed_y, seed_x)
//              among its own seed and the in-bounds seeds at p + s*(dx,dy);
//   steps    : 2^(P-1), ..., 2, 1 with P = ceil(log2 max(W,H));
//   SDF(p)   = one-byte conservative code of sqrt(d^2) (rc_common.hpp), 255
//              when no pixel is occupied.
//
// The passes are VALU-bound on CDNA4 (a wave64 VALU op occupies a 16-lane
// SIMD for 4 cycles; 8 candidates per pixel per pass), so the candidate test
// is kept to a packed 16-bit subtract, one dot2 and a 32-bit lexicographic
// select, and each lane owns two horizontally adjacent pixels so address
// arithmetic is shared.  Passes run as global kernels; the last pass writes
// the SDF directly.  Optionally the last passes (step <= S) are fused in one
// LDS-tiled kernel (halo 2S-1), trading halo recomputation for HBM traffic.
#include "rc_internal.hpp"

namespace rc {

typedef short short2v __attribute__((ext_vector_type(2)));

// Inside the kernels "no seed" is a fake seed at (32767, 32767).  For every
// pixel of an image up to 16384 px a side its d^2 >= 2 * 16384^2, above any
// real candidate's (<= 2 * 16383^2), and every d^2 fits in int32, so the
// (d^2, seed) minimum needs no special case.  The external value kSeedNone is
// restored on output.
constexpr uint32_t kFarSeed = 0x7FFF7FFFu;

// Buffer resource word 3 for gfx9-family raw buffer loads (32-bit data,
// range-checked against num_records).
constexpr int kBufferDword3 = 0x00020000;

__device__ inline uint32_t seed_d2(uint32_t s, uint32_t pix)
{
    const short2v d = __builtin_bit_cast(short2v, s) - __builtin_bit_cast(short2v, pix);
    return (uint32_t)__builtin_amdgcn_sdot2(d, d, 0, false);
}

__device__ inline uint8_t seed_sdf_code(uint32_t bs, uint32_t bd2)
{
    return (uint8_t)(bs == kFarSeed ? kSdfFar : sdf_code_of_distance(isqrt_u32(bd2)));
}

int jfa_num_passes(int W, int H)
{
    int m = W > H ? W : H, P = 0;
    while ((1 << P) < m) ++P;
    return P;
}

int jfa_tail_step(int W, int H, int fuse)
{
    const int P = jfa_num_passes(W, H);
    if (P == 0 || fuse <= 0) return 0;
    const int top = 1 << (P - 1);
    return top < fuse ? top : fuse;
}

// 64 x 16 pixel blocks: wave w covers rows 4w..4w+3 of 64 consecutive
// pixels (occupancy4: a thread's 4 pixels have their loads in flight at once).
constexpr int kSeedRows = 16;

// With emis4 set, also writes the emission interleaved (r, g, b, 0) of every
// occupied pixel: the merge kernels then gather a hit pixel's radiance as one
// 16-byte load instead of one 4-byte load from each of three planes 8 MB
// apart.  Only occupied pixels are written: a ray hits where the distance is
// 0, i.e. on a seed, and seeds are the occupied pixels.
__device__ __forceinline__ void seed_init_body(const float *__restrict__ depth, const DepthConv &dc,
                                               const float *__restrict__ emissive, uint32_t *__restrict__ seed,
                                               float4 *__restrict__ emis4, int W, int H, int bx_count, int b)
{
    const int x = (b % bx_count) * 64 + (threadIdx.x & 63);
    const int y0 = (b / bx_count) * kSeedRows + 4 * (threadIdx.x >> 6);
    if (x >= W) return;
    bool occ[4];
    float4 e[4];
    occupancy4(depth, dc, emissive, W, H, x, y0, occ, e);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = y0 + j;
        if (y >= H) break;
        const size_t p = (size_t)y * W + x;
        seed[p] = occ[j] ? (((uint32_t)y << 16) | (uint32_t)x) : kFarSeed;
        if (emis4 && occ[j]) emis4[p] = e[j];
    }
}

__global__ __launch_bounds__(kBlock) void k_seed_init(const float *__restrict__ depth, DepthConv dc,
                                                      const float *__restrict__ emissive,
                                                      uint32_t *__restrict__ seed, float4 *__restrict__ emis4, int W,
                                                      int H, int bx_count, const int *gate)
{
    if (gate && *gate == 0) return;
    seed_init_body(depth, dc, emissive, seed, emis4, W, H, bx_count, xcd_band(blockIdx.x, gridDim.x));
}

// One global pass.  A wave owns 64 consecutive pixels of R consecutive rows
// (blocks of 64 x 4R): the row setup is shared and every lane keeps 9R loads
// in flight.  The row index is wave-uniform, so the source rows are scalar
// base offsets and each load is a 32-bit lane offset.  FINAL: write the SDF
// (and, when `out` is set, the external seed field) instead of the internal
// seed field.  Every R computes the same field.
template <bool FINAL, int R>
__device__ __forceinline__ void jfa_pass_body(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                              uint8_t *__restrict__ sdf, int W, int H, int step, int bx_count, int b)
{
    const int x = (b % bx_count) * 64 + (threadIdx.x & 63);
    const int y0 = __builtin_amdgcn_readfirstlane((b / bx_count) * (4 * R) + (int)(threadIdx.x >> 6) * R);
    if (y0 >= H) return;
    const bool xin = x < W;
    const int xc = min(x, W - 1);
    const int xl = x - step, xr = x + step;
    // unsigned 32-bit lane offsets let the loads use the scalar-base form
    const uint32_t ol = (uint32_t)max(xl, 0), oc = (uint32_t)xc, orr = (uint32_t)min(xr, W - 1);
    const bool lf = xl >= 0, rt = xr < W;

    // 9R unconditional buffer loads at clamped coordinates, all in flight
    // together: the row offset is a scalar, the column offset a lane VGPR
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t *>(in), (short)0, W * H * 4, kBufferDword3);
    uint32_t cand[R][9];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = min(y0 + r, H - 1);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int qy = min(max(y + (k - 1) * step, 0), H - 1);
            const int rowb = qy * W * 4;
            cand[r][3 * k + 0] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, ol * 4, rowb, 0);
            cand[r][3 * k + 1] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, oc * 4, rowb, 0);
            cand[r][3 * k + 2] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, orr * 4, rowb, 0);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = y0 + r;
        if (y >= H) break;
        const uint32_t pix = ((uint32_t)y << 16) | (uint32_t)xc;
        const bool up = y - step >= 0, dn = y + step < H;
        const bool valid[9] = {up && lf, up, up && rt, lf, true, rt, dn && lf, dn, dn && rt};
        // An out-of-bounds neighbour (read at a clamped coordinate) enters
        // with 2^31 added to its d^2: every in-bounds candidate has d^2 < 2^31
        // (the far seed's is at most 2 * 32767^2), so it never wins, and no
        // sum wraps.
        uint64_t best = ((uint64_t)seed_d2(cand[r][4], pix) << 32) | cand[r][4];
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            if (j == 4) continue;
            const short2v d = __builtin_bit_cast(short2v, cand[r][j]) - __builtin_bit_cast(short2v, pix);
            const uint32_t d2 = (uint32_t)__builtin_amdgcn_sdot2(d, d, valid[j] ? 0 : INT32_MIN, false);
            const uint64_t key = ((uint64_t)d2 << 32) | cand[r][j];
            best = key < best ? key : best;
        }
        if (xin) {
            const uint32_t bs = (uint32_t)best, bd2 = (uint32_t)(best >> 32);
            const size_t p = (size_t)y * W + x;
            if (FINAL) {
                sdf[sdf_index(x, y, sdf_skip(H))] = seed_sdf_code(bs, bd2);
                if (out) out[p] = bs == kFarSeed ? kSeedNone : bs;
            } else {
                out[p] = bs;
            }
        }
    }
}

template <bool FINAL, int R>
__global__ __launch_bounds__(kBlock) void k_jfa_pass(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                     uint8_t *__restrict__ sdf, int W, int H, int step, int bx_count,
                                                     const int *gate)
{
    if (gate && *gate == 0) return;
    jfa_pass_body<FINAL, R>(in, out, sdf, W, H, step, bx_count, xcd_band(blockIdx.x, gridDim.x));
}

// Fused passes S, S/2, ..., 1 over a T x T output tile, loaded with a halo of
// 2S-1 pixels: T = 64 - 2 (2S - 1), so the loaded region is 64 x 64 cells and
// every row of it fills a wave (T = 62, 58, 50, 34 for S = 1, 2, 4, 8).  Each
// pass shrinks the valid region by its step.  A workgroup is 16 waves (there
// are few tiles, and the LDS passes need the waves to hide their latency),
// laid out 64 x 16: lane = column, so every LDS access of a wave is one row of
// consecutive words.  Cells outside the image stay the far seed in every
// pass, which is exactly "out-of-bounds neighbours are skipped"; the minimum
// is over the 64-bit keys (d^2, seed) as in the global pass.
template <int S> constexpr int jfa_tail_tile() { return 64 - 2 * (2 * S - 1); }

constexpr int kTailBlock = 1024, kTailRowsStep = kTailBlock / 64;

template <int S>
__device__ __forceinline__ void jfa_tail_body(const uint32_t *__restrict__ in, uint32_t *__restrict__ seeds_out,
                                              uint8_t *__restrict__ sdf, int W, int H, int bx_count, int b)
{
    constexpr int T = jfa_tail_tile<S>();
    constexpr int HALO = 2 * S - 1;
    constexpr int R = T + 2 * HALO;
    static_assert(S >= 1 && R == 64, "one column per lane");
    __shared__ uint32_t buf[2][R * R];

    const int x0 = (b % bx_count) * T - HALO;
    const int y0 = (b / bx_count) * T - HALO;
    const int cx = threadIdx.x & 63, cy = threadIdx.x >> 6;
    constexpr int RS = kTailRowsStep;
    {
        constexpr int ROWS = R / RS;
        const int gx = x0 + cx, qx = min(max(gx, 0), W - 1);
        uint32_t v[ROWS];
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            const int gy = y0 + cy + RS * j, qy = min(max(gy, 0), H - 1);
            v[j] = in[(size_t)qy * W + qx];
        }
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            const int ry = cy + RS * j, gy = y0 + ry;
            buf[0][ry * R + cx] = (gx >= 0 && gy >= 0 && gx < W && gy < H) ? v[j] : kFarSeed;
        }
    }
    __syncthreads();

    int cur = 0;
    int off = 0;
    uint32_t last_d2[(T + RS - 1) / RS] = {};
#pragma unroll
    for (int s = S; s >= 1; s >>= 1) {
        off += s;
        const int rx = off + cx;
        const int gx = x0 + rx;
        const bool xin = rx < R - off && gx >= 0 && gx < W;
#pragma unroll
        for (int j = 0; j < (R - 2 * off + RS - 1) / RS; ++j) {
            const int ry = off + cy + RS * j;
            const int gy = y0 + ry;
            const uint32_t *src = buf[cur];
            if (rx < R - off && ry < R - off) {
                uint32_t bs = kFarSeed, bd2 = 0xFFFFFFFFu;
                if (xin && gy >= 0 && gy < H) {
                    const uint32_t pix = ((uint32_t)gy << 16) | (uint32_t)gx;
                    const uint32_t c0 = src[ry * R + rx];
                    uint64_t best = ((uint64_t)seed_d2(c0, pix) << 32) | c0;
#pragma unroll
                    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (dx == 0 && dy == 0) continue;
                            const uint32_t cs = src[(ry + dy * s) * R + rx + dx * s];
                            const uint64_t key = ((uint64_t)seed_d2(cs, pix) << 32) | cs;
                            best = key < best ? key : best;
                        }
                    bs = (uint32_t)best;
                    bd2 = (uint32_t)(best >> 32);
                }
                buf[cur ^ 1][ry * R + rx] = bs;
                if (s == 1) last_d2[j] = bd2;
            }
        }
        __syncthreads();
        cur ^= 1;
    }

    if (cx < T) {
        const int gx = x0 + HALO + cx;
#pragma unroll
        for (int j = 0; j < (T + RS - 1) / RS; ++j) {
            const int ty = cy + RS * j;
            const int gy = y0 + HALO + ty;
            if (ty >= T || gx >= W || gy >= H) continue;
            const uint32_t sd = buf[cur][(HALO + ty) * R + HALO + cx];
            sdf[sdf_index(gx, gy, sdf_skip(H))] = seed_sdf_code(sd, last_d2[j]);
            if (seeds_out) seeds_out[(size_t)gy * W + gx] = sd == kFarSeed ? kSeedNone : sd;
        }
    }
}

template <int S>
__global__ __launch_bounds__(kTailBlock) void k_jfa_tail(const uint32_t *__restrict__ in,
                                                     uint32_t *__restrict__ seeds_out, uint8_t *__restrict__ sdf,
                                                     int W, int H, int bx_count, const int *gate)
{
    if (gate && *gate == 0) return;
    jfa_tail_body<S>(in, seeds_out, sdf, W, H, bx_count, xcd_band(blockIdx.x, gridDim.x));
}

// An image with no pass at all (1x1): only the SDF conversion.
__global__ void k_jfa_none(const uint32_t *__restrict__ in, uint32_t *__restrict__ seeds_out,
                           uint8_t *__restrict__ sdf, int W, int n, const int *gate)
{
    if (gate && *gate == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = in[i];
    sdf[sdf_index(i % W, i / W, sdf_skip(n / W))] = (uint8_t)(s == kFarSeed ? kSdfFar : 0u);
    if (seeds_out) seeds_out[i] = s == kFarSeed ? kSeedNone : s;
}

// Row-major fp32 copy of the distance values the march reads (parity readback).
__global__ void k_sdf_untile(const uint8_t *__restrict__ sdf, float *__restrict__ out, int W, int H)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    out[i] = sdf_value(sdf[sdf_index(i % W, i / W, sdf_skip(H))]);
}

hipError_t launch_sdf_untile(const uint8_t *sdf, float *out, int W, int H, hipStream_t s)
{
    const int n = W * H;
    hipLaunchKernelGGL(k_sdf_untile, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, sdf, out, W, H);
    return hipGetLastError();
}

hipError_t launch_seed_init(const float *depth, const DepthConv &dc, const float *emissive, uint32_t *seed,
                            float4 *emis4, int W, int H, hipStream_t s, const int *gate)
{
    const int bxc = (W + 63) / 64, byc = (H + kSeedRows - 1) / kSeedRows;
    hipLaunchKernelGGL(k_seed_init, dim3(bxc * byc), dim3(kBlock), 0, s, depth, dc, emissive, seed, emis4, W, H, bxc,
                       gate);
    return hipGetLastError();
}

template <bool FINAL>
static void launch_pass_rows(const uint32_t *in, uint32_t *out, uint8_t *sdf, int W, int H, int step, int rows,
                             hipStream_t s, const int *gate)
{
    const int bxc = (W + 63) / 64, byc = (H + 4 * rows - 1) / (4 * rows);
    const dim3 grid(bxc * byc), block(kBlock);
    switch (rows) {
    case 2: hipLaunchKernelGGL((k_jfa_pass<FINAL, 2>), grid, block, 0, s, in, out, sdf, W, H, step, bxc, gate); break;
    case 4: hipLaunchKernelGGL((k_jfa_pass<FINAL, 4>), grid, block, 0, s, in, out, sdf, W, H, step, bxc, gate); break;
    default: hipLaunchKernelGGL((k_jfa_pass<FINAL, 1>), grid, block, 0, s, in, out, sdf, W, H, step, bxc, gate); break;
    }
}

hipError_t launch_jfa_pass(const uint32_t *in, uint32_t *out, uint8_t *sdf, int W, int H, int step, bool final_pass,
                           hipStream_t s, const int *gate, int rows)
{
    if (rows != 1 && rows != 2 && rows != 4) return hipErrorInvalidValue;
    if (final_pass) launch_pass_rows<true>(in, out, sdf, W, H, step, rows, s, gate);
    else launch_pass_rows<false>(in, out, sdf, W, H, step, rows, s, gate);
    return hipGetLastError();
}

template <int S> static void launch_tail(const uint32_t *in, uint32_t *seeds_out, uint8_t *sdf, int W, int H,
                                         hipStream_t s, const int *gate)
{
    constexpr int T = jfa_tail_tile<S>();
    const int bxc = (W + T - 1) / T, byc = (H + T - 1) / T;
    hipLaunchKernelGGL(k_jfa_tail<S>, dim3(bxc * byc), dim3(kTailBlock), 0, s, in, seeds_out, sdf, W, H, bxc, gate);
}

hipError_t launch_jfa_tail(const uint32_t *in, uint32_t *seeds_out, uint8_t *sdf, int W, int H, int tail_step,
                           hipStream_t s, const int *gate)
{
    switch (tail_step) {
    case 0: {
        const int n = W * H;
        hipLaunchKernelGGL(k_jfa_none, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, in, seeds_out, sdf, W, n,
                           gate);
        break;
    }
    case 1: launch_tail<1>(in, seeds_out, sdf, W, H, s, gate); break;
    case 2: launch_tail<2>(in, seeds_out, sdf, W, H, s, gate); break;
    case 4: launch_tail<4>(in, seeds_out, sdf, W, H, s, gate); break;
    case 8: launch_tail<8>(in, seeds_out, sdf, W, H, s, gate); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rc
