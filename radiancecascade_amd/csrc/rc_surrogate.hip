// rc_surrogate.hip -- the learned far field: the levels >= 2 of a frame are
// replaced by a cone-traced estimate of the radiance arriving from beyond the
// level-1 interval, corrected per level-2 probe by the MLP (rc_mlp.hip).
//
// None of this exists in the reference (it has no global illumination at
// all; its only pass is the forward raster at src/renderer.rs:559-613).  The
// frame is
//   k_seed_surr                 occupancy bits + mip levels 1..6
//   k_mip_top_near              mip levels 7..top (one workgroup) beside the
//                               exact distance field up to 10 px
//   k_cone_pair                 16 cones per level-3 probe (far segment) and
//                               per level-2 probe (near segment)
//   k_surr_mlp                  far segment appended; the network's merged level 2
//   k_level (rc_cascade.hip)    level 1, marched on the near distance field
//   k_level0 (rc_cascade.hip)   level 0 + gather
// (a profiled frame runs k_near_sdf, k_mip_top and k_cone stage by stage).
// Cone tracing through a prefiltered pyramid is Crassin et al.'s voxel cone
// tracing (2011) in screen space; here it is the MLP's input feature, not the
// answer.
//
// Cone model (per level-2 probe c and level-1 direction k, 16 cones of
// half-angle pi/16; beyond kConeSplit from the level-3 probes, interpolated
// as the merge does): samples at the schedule's distances t_n, bilinear in the
// mip level m_n whose cells match the cone width.  The share of the cone's
// parallel rays that a cell stops is read from its column and row fractions
// with the direction's weights, c = dy^2 x + dx^2 y: a thin line across the
// cone stops all of it, whatever its share of the cell's area.  Emitters and
// dark occluders give a_e = 1 - (1 - min(c_e, 0.999))^re_n and a_o = 1 -
// (1 - min(c_o, 0.999))^ro_n (re_n = the step's share of a cell, ro_n = 3
// re_n), a = min(a_e + a_o, 1); L += T a_e Ebar with Ebar the mean radiance
// of the footprint's emitter pixels (mean E*[emitter] / emitter area),
// T *= 1 - a; a cone stops once T < 1e-4 or once its bilinear footprint
// lies wholly outside the image.
#include <math.h>

#include "rc_internal.hpp"
#include "rc_march.hpp"

namespace rc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---- packed mip cells ------------------------------------------------------------
// A cell of 2^m x 2^m pixels is 8 fp16 values (16 bytes, rounded to nearest
// even): the mean E*[emitter] rgb, the emitter area fraction, and the
// fractions of the cell's pixel columns (x) and rows (y) that hold an
// emitter (ex, ey) or a dark occluder (ox, oy) pixel; a pixel is a cell with
// these 0 or 1.  A parent is computed in fp32 from its 4 stored children TL,
// TR, BL, BR: means ((TL + TR) + (BL + BR)) / 4, x fractions (max(TL, BL) +
// max(TR, BR)) / 2, y fractions (max(TL, TR) + max(BL, BR)) / 2 -- a column
// of the left half is held if either child holds it; the max is that union
// whenever the children's held columns nest (a line, a block).
struct MipCell {
    float v[8];  // er eg eb ea ex ey ox oy
};
__device__ inline uint4 pack_cell(const MipCell &c)
{
    return make_uint4(f32_to_f16_rne(c.v[0]) | (f32_to_f16_rne(c.v[1]) << 16),
                      f32_to_f16_rne(c.v[2]) | (f32_to_f16_rne(c.v[3]) << 16),
                      f32_to_f16_rne(c.v[4]) | (f32_to_f16_rne(c.v[5]) << 16),
                      f32_to_f16_rne(c.v[6]) | (f32_to_f16_rne(c.v[7]) << 16));
}
__device__ inline MipCell unpack_cell(uint4 u)
{
    MipCell c;
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        c.v[2 * i] = f16_to_f32(w[i] & 0xFFFFu);
        c.v[2 * i + 1] = f16_to_f32(w[i] >> 16);
    }
    return c;
}
__device__ inline MipCell combine4(const MipCell &tl, const MipCell &tr, const MipCell &bl, const MipCell &br)
{
    MipCell p;
#pragma unroll
    for (int i = 0; i < 4; ++i) p.v[i] = 0.25f * ((tl.v[i] + tr.v[i]) + (bl.v[i] + br.v[i]));
#pragma unroll
    for (int i = 4; i < 8; i += 2) {
        p.v[i] = 0.5f * (fmaxf(tl.v[i], bl.v[i]) + fmaxf(tr.v[i], br.v[i]));
        p.v[i + 1] = 0.5f * (fmaxf(tl.v[i + 1], tr.v[i + 1]) + fmaxf(bl.v[i + 1], br.v[i + 1]));
    }
    return p;
}
__device__ inline uint4 combine4(uint4 tl, uint4 tr, uint4 bl, uint4 br)
{
    return pack_cell(combine4(unpack_cell(tl), unpack_cell(tr), unpack_cell(bl), unpack_cell(br)));
}

// The pixel of a mip cell: emitter (kind 1: rgb, area 1, columns and rows
// 1) or dark occluder (kind 2) or empty (kind 0), from (rgb, kind bits).
__device__ inline MipCell pixel_cell(float4 p)
{
    const uint32_t kind = __float_as_uint(p.w);
    const float f = kind == 1u ? 1.0f : 0.0f, o = kind == 2u ? 1.0f : 0.0f;
    return MipCell{{p.x, p.y, p.z, f, f, f, o, o}};
}

// ---- near-field distance field ---------------------------------------------
// The distance code of the nearest occupied pixel wherever it is at most
// kNearR pixels away (exact there), the far code elsewhere.  Levels 0 and 1
// end their intervals at 2 and 10 px, so every step they take from such a
// field (a hit needs distance 0; a value of 10 or more ends the interval in
// one step either way) equals the step they take from the full field.  Per
// 32x32 tile: the occupancy row masks of the tile plus a 10-px halo (two
// words of the bit field per row), the
// distance to the nearest occupied pixel of the row within +-10 (count of
// zero bits), then d^2 = min over |dy| <= 10 of g(y+dy)^2 + dy^2.
constexpr int kNearR = 10, kNearT = 32, kNearS = kNearT + 2 * kNearR;

__device__ inline void near_sdf_block(int blk, const uint64_t *__restrict__ bits, int words, uint8_t *__restrict__ sdf,
                                      int W, int H, int tiles_x)
{
    __shared__ uint64_t mask[kNearS];
    __shared__ uint8_t g[kNearS][kNearT];
    const int x0 = (blk % tiles_x) * kNearT, y0 = (blk / tiles_x) * kNearT;
    if (threadIdx.x < kNearS) {
        const int r = threadIdx.x, y = y0 - kNearR + r;
        uint64_t mk = 0;
        if (y >= 0 && y < H) {
            const uint64_t *rowp = bits + (size_t)y * words;
            const int xs = x0 - kNearR;  // bit j of mk <-> column xs + j
            if (xs < 0) {
                mk = rowp[0] << (-xs);
            } else {
                const int wi = xs >> 6, sh = xs & 63;
                const uint64_t lo = rowp[wi], hi = wi + 1 < words ? rowp[wi + 1] : 0ull;
                mk = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
            }
        }
        mask[r] = mk;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kNearS * kNearT; i += kBlock) {
        const int r = i / kNearT, j = i % kNearT, b = j + kNearR;
        const uint64_t mk = mask[r];
        const uint64_t right = mk >> b;         // bit d: column b + d
        const uint64_t left = mk << (63 - b);   // bit 63 - d: column b - d
        int d = kNearR + 1;
        if (right) d = min(d, (int)__builtin_ctzll(right));
        if (left) d = min(d, (int)__builtin_clzll(left));
        g[r][j] = (uint8_t)d;
    }
    __syncthreads();
    const int skip = sdf_skip(H);
    for (int i = threadIdx.x; i < kNearT * kNearT; i += kBlock) {
        const int j = i % kNearT, r = i / kNearT, x = x0 + j, y = y0 + r;
        if (x >= W || y >= H) continue;
        int best = kNearR * kNearR + 1;
#pragma unroll
        for (int dy = -kNearR; dy <= kNearR; ++dy) {
            const int gv = g[r + kNearR + dy][j];
            best = min(best, gv * gv + dy * dy);
        }
        sdf[sdf_index(x, y, skip)] = (uint8_t)(best <= kNearR * kNearR ? isqrt_u32((uint32_t)best) : kSdfFar);
    }
}

__global__ __launch_bounds__(kBlock) void k_near_sdf(const uint64_t *__restrict__ bits, int words,
                                                     uint8_t *__restrict__ sdf, int W, int H, int tiles_x)
{
    near_sdf_block(blockIdx.x, bits, words, sdf, W, H, tiles_x);
}

hipError_t launch_near_sdf(const uint64_t *bits, uint8_t *sdf, int W, int H, hipStream_t s)
{
    const int tx = (W + kNearT - 1) / kNearT, ty = (H + kNearT - 1) / kNearT;
    hipLaunchKernelGGL(k_near_sdf, dim3(tx * ty), dim3(kBlock), 0, s, bits, (W + 63) / 64, sdf, W, H, tx);
    return hipGetLastError();
}

// ---- mip pyramid, levels 3.. -----------------------------------------------------
int cone_top_level(int W, int H)
{
    int top = 1;
    while ((1 << top) < W || (1 << top) < H) ++top;
    return top < kMipMax ? top : kMipMax;
}

void mip_layout(int W, int H, int top, MipArgs &m)
{
    m.levels = top;
    size_t off = 0;
    for (int l = 0; l <= kMipMax; ++l) {
        if (l == 0 || l > top) {
            m.w[l] = m.h[l] = 0;
            m.off[l] = off;
            continue;
        }
        m.w[l] = (W + (1 << l) - 1) >> l;
        m.h[l] = (H + (1 << l) - 1) >> l;
        m.off[l] = off;
        off += (size_t)(m.w[l] + 2) * (m.h[l] + 2);
    }
}

size_t mip_cells(const MipArgs &m) { return m.off[m.levels] + (size_t)(m.w[m.levels] + 2) * (m.h[m.levels] + 2); }

// Levels 3..6 of a 16x16 block of level-2 cells (64x64 pixels) in LDS.
constexpr int kMidTile = 16, kMidLast = 6;

// Levels 3..6 of one 16x16 tile of level-2 cells held in LDS (se, row-major;
// cells outside the image are zero), written to the pyramid.
__device__ inline void mip_mid_levels(uint4 *se, const MipArgs &m, int tx, int ty)
{
    int n = kMidTile;
    const int lastl = m.levels < kMidLast ? m.levels : kMidLast;
    for (int l = 3; l <= lastl; ++l) {
        const int nn = n >> 1;
        const bool have = (int)threadIdx.x < nn * nn;
        uint4 t = make_uint4(0u, 0u, 0u, 0u);
        if (have) {
            const int cx = threadIdx.x % nn, cy = threadIdx.x / nn, i00 = (2 * cy) * n + 2 * cx;
            t = combine4(se[i00], se[i00 + 1], se[i00 + n], se[i00 + n + 1]);
        }
        __syncthreads();
        if (have) {
            se[threadIdx.x] = t;
            const int cx = threadIdx.x % nn, cy = threadIdx.x / nn;
            const int gx = tx * nn + cx, gy = ty * nn + cy;
            if (gx < m.w[l] && gy < m.h[l]) m.tex[mip_at(m, l, gx, gy)] = t;
        }
        __syncthreads();
        n = nn;
    }
}

// ---- occupancy bits + mip levels 1..6 ------------------------------------------
// One 64x64-pixel block of 16 waves: wave w covers rows 4w..4w+3 of 64
// consecutive columns (a lane's 4 pixels have their loads in flight at once;
// the depth rows above and below a lane's middle rows are its own loads).
// The occupancy rule of the jump flood's seed pass (emissive > 0 in any
// channel, or the near side of a depth edge), one 64-bit ballot per row into
// bits[y][x/64]; mip level 1 (2x2-px cells, the lane pair's pixels by
// shuffles), level 2 (4x4 px, the lane quad's level-1 cells) and levels 3..6
// of the block's 16x16 level-2 cells in LDS.  Every cell is combined from its
// stored (fp16-rounded) children in the order TL, TR, BL, BR.  Pixels outside
// the image are empty.
constexpr int kSeedMipBlock = 1024, kSeedMipPx = 64;
static_assert(kSeedMipBlock / 64 * 4 == kSeedMipPx && kSeedMipPx / 4 == kMidTile, "a block is one mid tile");

__device__ inline float4 shfl_xor4(float4 v, int mask)
{
    return make_float4(__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64), __shfl_xor(v.z, mask, 64),
                       __shfl_xor(v.w, mask, 64));
}
__device__ inline uint4 shfl_xor4(uint4 v, int mask)
{
    return make_uint4(__shfl_xor(v.x, mask, 64), __shfl_xor(v.y, mask, 64), __shfl_xor(v.z, mask, 64),
                      __shfl_xor(v.w, mask, 64));
}

__global__ __launch_bounds__(kSeedMipBlock) void k_seed_surr(const float *__restrict__ depth, DepthConv dc,
                                                             const float *__restrict__ emissive,
                                                             uint64_t *__restrict__ bits, int words, MipArgs m, int W,
                                                             int H, int bx_count)
{
    __shared__ uint4 se[kMidTile * kMidTile];
    // row bands per XCD: the depth rows above and below a block are mostly
    // read by the same XCD, from its L2
    const int b = xcd_band(blockIdx.x, gridDim.x);
    const int bx = b % bx_count, by = b / bx_count;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = bx * 64 + lane, y0 = by * kSeedMipPx + 4 * w;
    const int xc = min(x, W - 1);
    const size_t plane = (size_t)W * H;
    float e[4][3], dp[4], dl[4], dr[4];
    // all loads first (clamped addresses; out-of-image pixels are masked below)
    float du0 = depth[(size_t)min(max(y0 - 1, 0), H - 1) * W + xc];
    float dd3 = depth[(size_t)min(y0 + 4, H - 1) * W + xc];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = min(y0 + j, H - 1);
        const size_t p = (size_t)y * W + xc;
        e[j][0] = emissive[p];
        e[j][1] = emissive[plane + p];
        e[j][2] = emissive[2 * plane + p];
        dp[j] = depth[p];
        dl[j] = depth[p - (xc > 0)];
        dr[j] = depth[p + (xc + 1 < W)];
    }
    if (dc.ndc) {
        du0 = linear_depth(dc, du0);
        dd3 = linear_depth(dc, dd3);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            dp[j] = linear_depth(dc, dp[j]);
            dl[j] = linear_depth(dc, dl[j]);
            dr[j] = linear_depth(dc, dr[j]);
        }
    }
    float4 px[4];  // (rgb, kind bits) of the lane's 4 pixels
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = y0 + j;
        const bool in = x < W && y < H;
        // the rows above and below: this lane's neighbouring rows, or the two
        // extra loads at the ends (where y + 1 < H, row y + 1 is not clamped)
        const float du = j == 0 ? du0 : dp[j - 1], dd = j == 3 ? dd3 : dp[j + 1];
        const float thr = 0.05f * dp[j];
        const bool em = in && (e[j][0] > 0.0f || e[j][1] > 0.0f || e[j][2] > 0.0f);
        const bool occ = em || (in && ((x > 0 && dl[j] - dp[j] > thr) || (x + 1 < W && dr[j] - dp[j] > thr) ||
                                       (y > 0 && du - dp[j] > thr) || (y + 1 < H && dd - dp[j] > thr)));
        const uint64_t bw = __ballot(occ);
        if (lane == 0 && y < H) bits[(size_t)y * words + bx] = bw;
        px[j] = em ? make_float4(e[j][0], e[j][1], e[j][2], __uint_as_float(1u))
                   : make_float4(0.f, 0.f, 0.f, __uint_as_float(occ ? 2u : 0u));
    }
    // level 1: the cell of rows (2r, 2r + 1) and columns (lane, lane + 1) of
    // an even lane; TR, BR come from the odd lane
    uint4 l1[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const float4 tr = shfl_xor4(px[2 * r], 1), br = shfl_xor4(px[2 * r + 1], 1);
        l1[r] = pack_cell(combine4(pixel_cell(px[2 * r]), pixel_cell(tr), pixel_cell(px[2 * r + 1]), pixel_cell(br)));
        const int gx = bx * 32 + (lane >> 1), gy = by * (kSeedMipPx / 2) + 2 * w + r;
        if ((lane & 1) == 0 && gx < m.w[1] && gy < m.h[1]) m.tex[mip_at(m, 1, gx, gy)] = l1[r];
    }
    // level 2: the cell of level-1 rows (2w, 2w + 1) and columns of lanes
    // (4c, 4c + 2); TR, BR come from lane 4c + 2
    if (m.levels >= 2) {
        const uint4 tr = shfl_xor4(l1[0], 2), br = shfl_xor4(l1[1], 2);
        const uint4 l2 = combine4(l1[0], tr, l1[1], br);
        if ((lane & 3) == 0) {
            se[w * kMidTile + (lane >> 2)] = l2;
            const int gx = bx * 16 + (lane >> 2), gy = by * kMidTile + w;
            if (gx < m.w[2] && gy < m.h[2]) m.tex[mip_at(m, 2, gx, gy)] = l2;
        }
    }
    __syncthreads();
    if (m.levels >= 3) mip_mid_levels(se, m, bx, by);
}

hipError_t launch_seed_surr(const float *depth, const DepthConv &dc, const float *emissive, uint64_t *bits,
                            const MipArgs &m, int W, int H, hipStream_t s)
{
    const int bxc = (W + 63) / 64, byc = (H + kSeedMipPx - 1) / kSeedMipPx;
    hipLaunchKernelGGL(k_seed_surr, dim3(bxc * byc), dim3(kSeedMipBlock), 0, s, depth, dc, emissive, bits,
                       (W + 63) / 64, m, W, H, bxc);
    return hipGetLastError();
}

// Levels 7..top from level 6, one workgroup; also clears the fallback counter
// of the frame (a kernel store, not a memset node, so graph replays stay
// exact).  When level 6 fits in LDS (up to 4096 cells: images up to
// 4096x4096) it is read once and the levels are reduced in LDS.
constexpr int kTopLds = 4096;

__global__ __launch_bounds__(kBlock) void k_mip_top(MipArgs m, int *zero_me)
{
    __shared__ uint4 sa[kTopLds], sb[kTopLds / 4 + 160];  // ceil(w/2)ceil(h/2) <= (wh + w + h + 1) / 4
    if (zero_me && threadIdx.x == 0) *zero_me = 0;
    if (m.levels <= kMidLast) return;
    const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
    const int w6 = m.w[kMidLast], h6 = m.h[kMidLast];
    if (w6 * h6 <= kTopLds) {
        for (int i = threadIdx.x; i < w6 * h6; i += kBlock) sa[i] = m.tex[mip_at(m, kMidLast, i % w6, i / w6)];
        __syncthreads();
        uint4 *src = sa, *dst = sb;
        for (int l = kMidLast + 1; l <= m.levels; ++l) {
            const int w = m.w[l], h = m.h[l], cw = m.w[l - 1], ch = m.h[l - 1];
            for (int i = threadIdx.x; i < w * h; i += kBlock) {
                const int x = i % w, y = i / w;
                uint4 v[4];
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const int cx = 2 * x + (d & 1), cy = 2 * y + (d >> 1);
                    v[d] = (cx < cw && cy < ch) ? src[cy * cw + cx] : zero;
                }
                const uint4 t = combine4(v[0], v[1], v[2], v[3]);
                dst[i] = t;
                m.tex[mip_at(m, l, x, y)] = t;
            }
            __syncthreads();
            uint4 *tmp = src;
            src = dst;
            dst = tmp;
        }
        return;
    }
    for (int l = kMidLast + 1; l <= m.levels; ++l) {
        const int w = m.w[l], h = m.h[l], cw = m.w[l - 1], ch = m.h[l - 1];
        for (int i = threadIdx.x; i < w * h; i += kBlock) {
            const int x = i % w, y = i / w;
            uint4 v[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int cx = 2 * x + (d & 1), cy = 2 * y + (d >> 1);
                v[d] = (cx < cw && cy < ch) ? m.tex[mip_at(m, l - 1, cx, cy)] : zero;
            }
            m.tex[mip_at(m, l, x, y)] = combine4(v[0], v[1], v[2], v[3]);
        }
        __threadfence_block();
        __syncthreads();
    }
}

// Levels 7..top in one workgroup with little LDS: level 7 from level 6 in
// global memory (written by the previous launch), each later level from the
// previous one in LDS when that fits in kTopCap cells, else from global
// memory after a device-scope fence (images above 8192 px a side).
constexpr int kTopCap = 512;

__device__ inline void mip_top_small(const MipArgs &m, uint4 *__restrict__ buf_a, uint4 *__restrict__ buf_b)
{
    const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
    const uint4 *src = nullptr;  // level l - 1 in LDS, or nullptr: in global memory
    for (int l = kMidLast + 1; l <= m.levels; ++l) {
        const int w = m.w[l], h = m.h[l], cw = m.w[l - 1], ch = m.h[l - 1];
        uint4 *dst = w * h <= kTopCap ? (src == buf_a ? buf_b : buf_a) : nullptr;
        for (int i = threadIdx.x; i < w * h; i += kBlock) {
            const int x = i % w, y = i / w;
            uint4 v[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int cx = 2 * x + (d & 1), cy = 2 * y + (d >> 1);
                v[d] = (cx < cw && cy < ch) ? (src ? src[cy * cw + cx] : m.tex[mip_at(m, l - 1, cx, cy)]) : zero;
            }
            const uint4 t = combine4(v[0], v[1], v[2], v[3]);
            m.tex[mip_at(m, l, x, y)] = t;
            if (dst) dst[i] = t;
        }
        if (!dst) __threadfence();
        __syncthreads();
        src = dst;
    }
}

// Mip levels 7..top (block 0, a latency chain of small reductions) beside
// the near distance field (blocks 1..): both read only what earlier launches
// wrote.  Block 0 also clears the fallback counter of the frame (a kernel
// store, not a memset node, so graph replays stay exact).
__global__ __launch_bounds__(kBlock) void k_mip_top_near(MipArgs m, int *zero_me, const uint64_t *__restrict__ bits,
                                                         int words, uint8_t *__restrict__ sdf, int W, int H,
                                                         int near_tiles_x)
{
    // row bands per XCD; hardware block 0 is logical block 0, the top levels
    const int lb = xcd_band(blockIdx.x, gridDim.x);
    if (lb == 0) {
        __shared__ uint4 buf_a[kTopCap], buf_b[kTopCap];
        if (zero_me && threadIdx.x == 0) *zero_me = 0;
        mip_top_small(m, buf_a, buf_b);
        return;
    }
    near_sdf_block(lb - 1, bits, words, sdf, W, H, near_tiles_x);
}

hipError_t launch_mip_top_near(const MipArgs &m, int *zero_me, const uint64_t *bits, uint8_t *sdf, int W, int H,
                               hipStream_t s)
{
    const int ntx = (W + kNearT - 1) / kNearT, n_near = ntx * ((H + kNearT - 1) / kNearT);
    hipLaunchKernelGGL(k_mip_top_near, dim3(1 + n_near), dim3(kBlock), 0, s, m, zero_me, bits, (W + 63) / 64, sdf, W,
                       H, ntx);
    return hipGetLastError();
}

hipError_t launch_mip_top(const MipArgs &m, int *zero_me, hipStream_t s)
{
    hipLaunchKernelGGL(k_mip_top, dim3(1), dim3(kBlock), 0, s, m, zero_me);
    return hipGetLastError();
}


// Unpacked copy of mip level l (parity readback): emission [h][w][4] (mean
// E*[emitter] rgb, emitter area), coverage [h][w][4] (ex, ey, ox, oy).
__global__ __launch_bounds__(kBlock) void k_mip_unpack(MipArgs m, int l, float4 *__restrict__ emission,
                                                       float4 *__restrict__ coverage)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t n = (size_t)m.w[l] * m.h[l];
    if (i >= n) return;
    const MipCell c = unpack_cell(m.tex[mip_at(m, l, (int)(i % m.w[l]), (int)(i / m.w[l]))]);
    if (emission) emission[i] = make_float4(c.v[0], c.v[1], c.v[2], c.v[3]);
    if (coverage) coverage[i] = make_float4(c.v[4], c.v[5], c.v[6], c.v[7]);
}

hipError_t launch_mip_unpack(const MipArgs &m, int level, float *emission, float *coverage, hipStream_t s)
{
    const size_t n = (size_t)m.w[level] * m.h[level];
    hipLaunchKernelGGL(k_mip_unpack, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, m, level,
                       (float4 *)emission, (float4 *)coverage);
    return hipGetLastError();
}

// ---- cone trace --------------------------------------------------------------
// Schedule: t_0 = 10 px from the level-2 probe centre (where level 2's
// interval starts), cone diameter
// dia = max(2 t tan(pi/16), 1), mip level floor(log2 dia) clamped to
// [1, top], next t = t + 0.75 dia, up to the image diagonal + 2; the
// exponents are re = min(0.75 dia, 2^m) / 2^m and ro = kConeOccGain re.
void cone_schedule(int W, int H, const MipArgs &mip, ConeSchedule &near, ConeSchedule &far)
{
    const int top = mip.levels;
    const double tmax = sqrt((double)W * W + (double)H * H) + 2.0, tn = tan(M_PI / 16.0);
    double t = 10.0;  // the start of level 2's interval
    near.n = 0;
    far.n = 0;
    while (t < tmax && near.n + far.n < kConeMaxSteps) {
        const double dia = fmax(2.0 * t * tn, 1.0);
        int m = (int)floor(log2(dia));
        m = m < 1 ? 1 : (m > top ? top : m);
        const double step = 0.75 * dia;
        ConeSchedule &s = t < (double)kConeSplit ? near : far;
        ConeStep &c = s.s[s.n++];
        const double r = fmin(step, (double)(1 << m)) / (double)(1 << m);
        c.t_cell = (float)(t / (double)(1 << m));
        c.re = (float)r;
        c.ro = (float)(kConeOccGain * r);
        c.inv_cell = 1.0f / (float)(1 << m);
        c.w = mip.w[m];
        c.h = mip.h[m];
        c.off = (uint32_t)mip.off[m];
        c.wp = (uint32_t)(mip.w[m] + 2);
        t += step;
    }
}

// The far segment appended to a level-2 cone (L, T): the level-3 grid's cones
// interpolated at level-2 probe (qx, qy) with the merge's bilinear weights,
// L += T * F.rgb, T *= F.T.
__device__ inline float4 far_append(float4 c, const float4 *__restrict__ far, int FX, int FY, int qx, int qy, int k)
{
    const int bx = (qx - 1) >> 1, by = (qy - 1) >> 1;
    const int fx0 = min(max(bx, 0), FX - 1), fx1 = min(max(bx + 1, 0), FX - 1);
    const int fy0 = min(max(by, 0), FY - 1), fy1 = min(max(by + 1, 0), FY - 1);
    const float4 f00 = far[((size_t)fy0 * FX + fx0) * 16 + k], f10 = far[((size_t)fy0 * FX + fx1) * 16 + k];
    const float4 f01 = far[((size_t)fy1 * FX + fx0) * 16 + k], f11 = far[((size_t)fy1 * FX + fx1) * 16 + k];
    const float3 f = bilerp4(f00, f10, f01, f11, qx, qy);
    const float3 ft = bilerp4(make_float4(f00.w, f00.w, f00.w, 0.f), make_float4(f10.w, f10.w, f10.w, 0.f),
                              make_float4(f01.w, f01.w, f01.w, 0.f), make_float4(f11.w, f11.w, f11.w, 0.f), qx, qy);
    return make_float4(fmaf(c.w, f.x, c.x), fmaf(c.w, f.y, c.y), fmaf(c.w, f.z, c.z), c.w * ft.x);
}

// A workgroup is 64 consecutive level-2 probes of one row x 4 directions, a
// wave per direction: the lanes of a wave sample points 8 px apart at the same
// distance, so from mip level 3 up they share cells (and cache lines).  The 4
// taps of a step are one 32-bit offset and 4 range-checked buffer loads (the
// zero border stands in for taps beyond the image edge).
__device__ inline void cone_block(const ConeArgs &a, int lb, int seg_x)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // probe rows outermost: an XCD's band of blocks is a band of rows (all
    // four direction groups), so the fine mip levels it reads stay in its L2
    const int qy = lb / (seg_x * 4), rest = lb % (seg_x * 4);
    const int kgrp = rest / seg_x, seg = rest % seg_x;
    const int k = kgrp * 4 + wave;
    const int qx = seg * 64 + lane;
    if (qx >= a.QX) return;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(a.mip.tex, (short)0, (int)a.mip_bytes, kBufferDword3);
    const float cx = ((float)qx + 0.5f) * a.spacing, cy = ((float)qy + 0.5f) * a.spacing;
    const float2 d = a.dirs[k];
    const float dx2 = d.x * d.x, dy2 = d.y * d.y;
    float T = 1.0f, r = 0.f, g = 0.f, b = 0.f;
    const int nsteps = a.sch->n;
    for (int n = 0; n < nsteps; ++n) {
        const ConeStep st = a.sch->s[n];
        // sample position in cells of level m, less the half cell of the
        // bilinear taps: (t d + c) 2^-m - 1/2 as two FMAs per axis
        const float fx = fmaf(st.t_cell, d.x, fmaf(cx, st.inv_cell, -0.5f));
        const float fy = fmaf(st.t_cell, d.y, fmaf(cy, st.inv_cell, -0.5f));
        // cell index and fraction in one instruction each (v_cvt_flr_i32_f32,
        // v_fract_f32; the fraction is below 1 by construction)
        const int ix = floor_i32(fx), iy = floor_i32(fy);
        // the footprint has left the image: a straight ray does not come back
        if (!(ix >= -1 && ix < st.w && iy >= -1 && iy < st.h)) break;
        const float ax = __builtin_amdgcn_fractf(fx), ay = __builtin_amdgcn_fractf(fy);
        const uint32_t o00 = (st.off + __umul24((uint32_t)(iy + 1), st.wp) + (uint32_t)(ix + 1)) << 4;
        const uint32_t orow = st.wp << 4;
        const uint4 v[4] = {
            __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)o00, 0, 0)),
            __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(o00 + 16u), 0, 0)),
            __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(o00 + orow), 0, 0)),
            __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(o00 + orow + 16u), 0, 0))};
        const float wt[4] = {(1.0f - ax) * (1.0f - ay), ax * (1.0f - ay), (1.0f - ax) * ay, ax * ay};
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tap = 0; tap < 4; ++tap) {
            const MipCell c = unpack_cell(v[tap]);
#pragma unroll
            for (int i = 0; i < 8; ++i) acc[i] = fmaf(wt[tap], c.v[i], acc[i]);
        }
        const float ce = fmaf(dy2, acc[4], dx2 * acc[5]), co = fmaf(dy2, acc[6], dx2 * acc[7]);
        // 1 - (1 - c)^r; a step of one whole cell (r = 1, occluder r = 3) needs
        // no transcendental, the others use the hardware log2 / exp2
        // (argument in [1e-3, 1]); the branch is uniform (one step for all lanes)
        float ae, ao;
        static_assert(kConeOccGain == 3.0, "the whole-cell path cubes");
        if (st.re == 1.0f) {
            const float xo = 1.0f - fminf(co, 0.999f);
            ae = fminf(ce, 0.999f);
            ao = 1.0f - xo * xo * xo;
        } else {
            ae = 1.0f - __builtin_amdgcn_exp2f(st.re * __builtin_amdgcn_logf(1.0f - fminf(ce, 0.999f)));
            ao = 1.0f - __builtin_amdgcn_exp2f(st.ro * __builtin_amdgcn_logf(1.0f - fminf(co, 0.999f)));
        }
        const float al = fminf(ae + ao, 1.0f);
        // T a_e / emitter area, with the hardware reciprocal (1 ulp)
        const float we = acc[3] > 0.0f ? T * ae * __builtin_amdgcn_rcpf(acc[3]) : 0.0f;
        r = fmaf(we, acc[0], r);
        g = fmaf(we, acc[1], g);
        b = fmaf(we, acc[2], b);
        T = T * (1.0f - al);
        if (T < 1e-4f) break;
    }
    if (a.far) {
        const float4 v = far_append(make_float4(r, g, b, T), a.far, a.FX, a.FY, qx, qy, k);
        r = v.x;
        g = v.y;
        b = v.z;
        T = v.w;
    }
    a.cone[((size_t)qy * a.QX + qx) * 16 + k] = make_float4(r, g, b, T);
}

__global__ __launch_bounds__(kBlock) void k_cone(ConeArgs a, int seg_x)
{
    cone_block(a, xcd_band(blockIdx.x, gridDim.x), seg_x);
}

// Far and near cones in one launch: blocks [0, nb_far) (a multiple of 8, so
// the near blocks keep the XCD pattern) are far cones, dispatched first as
// they take more steps; each part is dealt to the XCDs in contiguous runs.
__global__ __launch_bounds__(kBlock) void k_cone_pair(ConeArgs near, int seg_near, int nb_near, ConeArgs far, int seg_far,
                                                      int nb_far_used, int nb_far)
{
    const int b = blockIdx.x;
    if (b < nb_far) {
        const int lb = xcd_band(b, nb_far);
        if (lb < nb_far_used) cone_block(far, lb, seg_far);
    } else {
        cone_block(near, xcd_band(b - nb_far, nb_near), seg_near);
    }
}

hipError_t launch_cone(const ConeArgs &a, hipStream_t s)
{
    const int seg_x = (a.QX + 63) / 64;
    hipLaunchKernelGGL(k_cone, dim3(seg_x * a.QY * 4), dim3(kBlock), 0, s, a, seg_x);
    return hipGetLastError();
}

hipError_t launch_cone_pair(const ConeArgs &near, const ConeArgs &far, hipStream_t s)
{
    const int seg_n = (near.QX + 63) / 64, seg_f = (far.QX + 63) / 64;
    const int nb_n = seg_n * near.QY * 4, nb_f_used = seg_f * far.QY * 4, nb_f = (nb_f_used + 7) / 8 * 8;
    hipLaunchKernelGGL(k_cone_pair, dim3(nb_f + nb_n), dim3(kBlock), 0, s, near, seg_n, nb_n, far, seg_f, nb_f_used,
                       nb_f);
    return hipGetLastError();
}

// ---- the network's level 2 -----------------------------------------------------------
// A workgroup owns 8x8 level-2 probes per iteration (persistent over tiles).
// Lane (probe p, quad ko) builds the 16 inputs of directions 4ko..4ko+3 of p:
// log(1 + c) of the probe's cone radiance and the cone's transmittance; wave
// w then computes the 32x32 output block (rows 32(w&1), columns 32(w>>1)) of
// each MLP layer on MFMA.
// The merged level 2 is A_2 = max(c + s * z, 0) per direction (4th channel
// 1; level 1 reads only rgb), written in the storage format of the merged
// levels; level 1 and level 0 are then the plain kernels.  The error head
// exp(z_48), averaged over the tile, is the tile's predicted mean squared
// error (normalised units).
constexpr int kSStrideIn = kMlpIn + 8, kSStrideHid = kMlpHid + 8;

// Input feature of a cone radiance value: log(1 + c) through the hardware
// log2 (rounded to bf16 before the network, far below fp32 log1p error).
__device__ inline float feat_log(float c) { return __logf(1.0f + c); }

__device__ inline bf16x8 ld8s(const uint16_t *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ inline int acc_row_s(int v, int lane) { return ((v >> 2) << 3) + ((lane >> 5) << 2) + (v & 3); }

// The wave's B fragments of one layer (columns c0..c0+31, all K rows), held
// in registers across the tiles a workgroup processes.
template <int K> struct BFrag {
    bf16x8 f[K / 16];
};
template <int K> __device__ inline BFrag<K> load_bfrag(const uint16_t *w, int c0)
{
    const int lane = threadIdx.x & 63, r = lane & 31, kh = (lane >> 5) * 8;
    BFrag<K> b;
#pragma unroll
    for (int i = 0; i < K / 16; ++i) b.f[i] = ld8s(w + (c0 + r) * K + 16 * i + kh);
    return b;
}
template <int K>
__device__ inline floatx16 mfma_block(const uint16_t *in, int sin, int rb, const BFrag<K> &b)
{
    const int lane = threadIdx.x & 63, r = lane & 31, kh = (lane >> 5) * 8;
    floatx16 c = {};
#pragma unroll
    for (int i = 0; i < K / 16; ++i)
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8s(in + (rb + r) * sin + 16 * i + kh), b.f[i], c, 0, 0, 0);
    return c;
}

template <Storage S>
__global__ __launch_bounds__(kBlock) void k_surr_mlp(SurrMlpArgs a, MlpDev m)
{
    using TT = Texel<S>;
    __shared__ float bias_s[4 * 64], mean_s[kMlpIn], istd_s[kMlpIn];
    // xs (layer-1 input) and hb (layer-3 input) are dead when the output layer
    // writes zs: one buffer holds all three.
    __shared__ __attribute__((aligned(16))) uint16_t xs_hb[64 * kSStrideIn + 64 * kSStrideHid];
    __shared__ __attribute__((aligned(16))) uint16_t ha[64 * kSStrideHid];
    __shared__ float red[kBlock / 64][2];
    static_assert(64 * (kMlpErr + 2) * 4 <= (64 * kSStrideIn + 64 * kSStrideHid) * 2, "zs fits over xs + hb");
    uint16_t *const xs = xs_hb;
    uint16_t *const hb = xs_hb + 64 * kSStrideIn;
    float(*const zs)[kMlpErr + 2] = reinterpret_cast<float(*)[kMlpErr + 2]>(xs_hb);

    const int p = threadIdx.x >> 2, ko = threadIdx.x & 3;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rb = (wave & 1) * 32, c0 = (wave >> 1) * 32;
    // tiles are dealt XCD by XCD in contiguous ranges
    const int per_xcd = (int)gridDim.x >> 3, xcd = blockIdx.x & 7, idx = blockIdx.x >> 3;
    const int t_begin = (int)((long long)a.tiles * xcd / 8), t_end = (int)((long long)a.tiles * (xcd + 1) / 8);
    // the whole cones of this lane's 4 directions of its probe in a tile (near
    // segment + appended far segment); the first tile's are in flight while
    // the constants are staged
    auto load_cones = [&](int tile, float4 *cf) {
        const int qx = (tile % a.tiles_x) * 8 + (p & 7), qy = (tile / a.tiles_x) * 8 + (p >> 3);
        const size_t q = (size_t)min(qy, a.QY - 1) * a.QX + min(qx, a.QX - 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) cf[j] = a.cone[q * 16 + 4 * ko + j];
        if (a.far) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                cf[j] = far_append(cf[j], a.far, a.FX, a.FY, min(qx, a.QX - 1), min(qy, a.QY - 1), 4 * ko + j);
        }
    };
    // B operands from the bf16 weights in global memory (32 KB, held in every
    // CU's L1 and the L2): all four layers' fragments of this wave, loaded
    // once (64 VGPRs) and issued first, so their latency, the first tile's
    // cones' and the constants' overlap
    const uint16_t *wl[4] = {m.wt, m.wt + 64 * kMlpIn, m.wt + 64 * kMlpIn + 64 * kMlpHid,
                             m.wt + 64 * kMlpIn + 2 * 64 * kMlpHid};
    const BFrag<kMlpIn> b0 = load_bfrag<kMlpIn>(wl[0], c0);
    const BFrag<kMlpHid> b1 = load_bfrag<kMlpHid>(wl[1], c0), b2 = load_bfrag<kMlpHid>(wl[2], c0),
                         b3 = load_bfrag<kMlpHid>(wl[3], c0);
    static_assert(4 * 64 == kBlock && kMlpIn <= kBlock, "one bias (and at most one feature constant) per thread");
    const float bias_v = m.bias[threadIdx.x];
    const float mean_v = threadIdx.x < kMlpIn ? m.in_mean[threadIdx.x] : 0.f;
    const float istd_v = threadIdx.x < kMlpIn ? m.in_istd[threadIdx.x] : 0.f;
    float4 cf[4];
    if (t_begin + idx < t_end) load_cones(t_begin + idx, cf);
    bias_s[threadIdx.x] = bias_v;
    if (threadIdx.x < kMlpIn) {
        mean_s[threadIdx.x] = mean_v;
        istd_s[threadIdx.x] = istd_v;
    }
    __syncthreads();

#pragma unroll 1
    for (int tile = t_begin + idx; tile < t_end; tile += per_xcd) {
        const int qx = (tile % a.tiles_x) * 8 + (p & 7), qy = (tile / a.tiles_x) * 8 + (p >> 3);
        const bool valid = qx < a.QX && qy < a.QY;
        const size_t q = (size_t)min(qy, a.QY - 1) * a.QX + min(qx, a.QX - 1);
        if (a.cone_out && valid) {
#pragma unroll
            for (int j = 0; j < 4; ++j) a.cone_out[q * 16 + 4 * ko + j] = cf[j];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = kMlpFeat * (4 * ko + j);
            const float f[kMlpFeat] = {feat_log(cf[j].x), feat_log(cf[j].y), feat_log(cf[j].z), cf[j].w};
#pragma unroll
            for (int e = 0; e < kMlpFeat; ++e)
                xs[p * kSStrideIn + col + e] =
                    (uint16_t)f32_to_bf16_rne(valid ? (f[e] - mean_s[col + e]) * istd_s[col + e] : 0.f);
        }
        __syncthreads();

        // MLP, 4 layers; wave (row block, column block)
        {
            floatx16 acc = mfma_block<kMlpIn>(xs, kSStrideIn, rb, b0);
            const float bb = bias_s[0 * 64 + c0 + (lane & 31)];
#pragma unroll
            for (int v = 0; v < 16; ++v)
                ha[(rb + acc_row_s(v, lane)) * kSStrideHid + c0 + (lane & 31)] =
                    (uint16_t)f32_to_bf16_rne(fmaxf(acc[v] + bb, 0.0f));
        }
        __syncthreads();
        {
            floatx16 acc = mfma_block<kMlpHid>(ha, kSStrideHid, rb, b1);
            const float bb = bias_s[1 * 64 + c0 + (lane & 31)];
#pragma unroll
            for (int v = 0; v < 16; ++v)
                hb[(rb + acc_row_s(v, lane)) * kSStrideHid + c0 + (lane & 31)] =
                    (uint16_t)f32_to_bf16_rne(fmaxf(acc[v] + bb, 0.0f));
        }
        __syncthreads();
        {
            floatx16 acc = mfma_block<kMlpHid>(hb, kSStrideHid, rb, b2);
            const float bb = bias_s[2 * 64 + c0 + (lane & 31)];
#pragma unroll
            for (int v = 0; v < 16; ++v)
                ha[(rb + acc_row_s(v, lane)) * kSStrideHid + c0 + (lane & 31)] =
                    (uint16_t)f32_to_bf16_rne(fmaxf(acc[v] + bb, 0.0f));
        }
        __syncthreads();
        {
            floatx16 acc = mfma_block<kMlpHid>(ha, kSStrideHid, rb, b3);
            const int col = c0 + (lane & 31);
            const float bb = bias_s[3 * 64 + col];
            if (col <= kMlpErr) {
#pragma unroll
                for (int v = 0; v < 16; ++v) zs[rb + acc_row_s(v, lane)][col] = acc[v] + bb;
            }
        }
        __syncthreads();

        // the predicted merged level 2: 4 texels of directions 4ko..4ko+3
        if (valid) {
            typename TT::T *out = (typename TT::T *)a.out;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 4 * ko + j;
                const float4 t = make_float4(fmaxf(fmaf(a.out_scale, zs[p][3 * k + 0], cf[j].x), 0.0f),
                                             fmaxf(fmaf(a.out_scale, zs[p][3 * k + 1], cf[j].y), 0.0f),
                                             fmaxf(fmaf(a.out_scale, zs[p][3 * k + 2], cf[j].z), 0.0f), 1.0f);
                out[q * 16 + k] = TT::pack(t);
            }
        }
        // the next tile's cones are in flight during the error reduction
        if (tile + per_xcd < t_end) load_cones(tile + per_xcd, cf);

        // tile error estimate
        float e = (valid && ko == 0) ? expf(zs[p][kMlpErr]) : 0.0f;
        float cnt = (valid && ko == 0) ? 1.0f : 0.0f;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            e += __shfl_xor(e, off, 64);
            cnt += __shfl_xor(cnt, off, 64);
        }
        if (lane == 0) {
            red[wave][0] = e;
            red[wave][1] = cnt;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const float se = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
            const float sc = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
            const float mean = sc > 0.f ? se / sc : 0.f;
            a.tile_err[tile] = mean;
            if (!(mean <= a.threshold)) {  // a NaN estimate counts as flagged
                const int slot = atomicAdd(a.flag_count, 1);
                a.flag_list[slot] = tile;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_surr_mlp(const SurrMlpArgs &a, const MlpDev &m, Storage st, hipStream_t s)
{
    if (a.tiles < 1) return hipSuccess;
    // a multiple of 8 workgroups (one share per XCD), at most 4 per CU
    const int grid = a.tiles < 1024 ? ((a.tiles + 7) / 8) * 8 : 1024;
    switch (st) {
    case Storage::F32: hipLaunchKernelGGL(k_surr_mlp<Storage::F32>, dim3(grid), dim3(kBlock), 0, s, a, m); break;
    case Storage::F16: hipLaunchKernelGGL(k_surr_mlp<Storage::F16>, dim3(grid), dim3(kBlock), 0, s, a, m); break;
    case Storage::BF16: hipLaunchKernelGGL(k_surr_mlp<Storage::BF16>, dim3(grid), dim3(kBlock), 0, s, a, m); break;
    }
    return hipGetLastError();
}

// ---- training samples ------------------------------------------------------------
// One row per level-2 probe: inputs per direction log(1 + c) rgb and the
// cone's transmittance T, target residual A_2 - c, where c is the probe's cone
// radiance and A_2 the plain path's merged level 2 (16 entries).
template <Storage S>
__global__ __launch_bounds__(kBlock) void k_surr_sample(SurrSampleArgs a)
{
    using TT = Texel<S>;
    const int gi = blockIdx.x * kBlock + threadIdx.x;
    if (gi >= a.QX * a.QY * 16) return;
    const int k = gi & 15, q = gi >> 4;
    const typename TT::T *up = (const typename TT::T *)a.upper;
    const float4 c = a.cone[gi];
    const float4 u = TT::unpack(up[gi]);
    float *f = a.feats + (size_t)q * kMlpIn + kMlpFeat * k;
    float *t = a.targets + (size_t)q * kMlpRad + 3 * k;
    f[0] = feat_log(c.x);
    f[1] = feat_log(c.y);
    f[2] = feat_log(c.z);
    f[3] = c.w;
    t[0] = u.x - c.x;
    t[1] = u.y - c.y;
    t[2] = u.z - c.z;
}

hipError_t launch_surr_sample(const SurrSampleArgs &a, Storage st, hipStream_t s)
{
    const int n = a.QX * a.QY * 16;
    const dim3 g((n + kBlock - 1) / kBlock);
    switch (st) {
    case Storage::F32: hipLaunchKernelGGL(k_surr_sample<Storage::F32>, g, dim3(kBlock), 0, s, a); break;
    case Storage::F16: hipLaunchKernelGGL(k_surr_sample<Storage::F16>, g, dim3(kBlock), 0, s, a); break;
    case Storage::BF16: hipLaunchKernelGGL(k_surr_sample<Storage::BF16>, g, dim3(kBlock), 0, s, a); break;
    }
    return hipGetLastError();
}

// ---- tile fallback: the probes the flagged tiles need -------------------------
// A tile is 8x8 level-2 probes = 16x16 level-1 probes (64x64 px).  Level 1
// needs the tiles' own probes; level i + 1 the probes level i's bilinear
// merge reads, [(lo - 1) >> 1, ((hi - 1) >> 1) + 1] per axis, clamped to the
// grid.  The first kernel clears the masks, the second marks one flagged
// tile's rectangle at every level per workgroup (overlapping rectangles store
// the same byte).  Both do nothing when no tile is flagged.
__global__ __launch_bounds__(kBlock) void k_fallback_clear(FallbackMaskArgs a)
{
    if (*a.count == 0) return;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < a.need_bytes; i += (size_t)gridDim.x * kBlock)
        a.need[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_fallback_mark(FallbackMaskArgs a)
{
    const int n = *a.count;
    for (int t = blockIdx.x; t < n && t < a.max_tiles; t += gridDim.x) {
        const int tile = a.list[t];
        int x0 = (tile % a.tiles_x) * 16, y0 = (tile / a.tiles_x) * 16;
        int x1 = min(x0 + 15, a.PX[1] - 1), y1 = min(y0 + 15, a.PY[1] - 1);
        for (int i = 1; i < a.L; ++i) {
            if (i > 1) {
                x0 = max((x0 - 1) >> 1, 0);
                y0 = max((y0 - 1) >> 1, 0);
                x1 = min(((x1 - 1) >> 1) + 1, a.PX[i] - 1);
                y1 = min(((y1 - 1) >> 1) + 1, a.PY[i] - 1);
            }
            const int w = x1 - x0 + 1, h = y1 - y0 + 1;
            uint8_t *m = a.need + a.need_off[i];
            for (int j = threadIdx.x; j < w * h; j += kBlock) m[(size_t)(y0 + j / w) * a.PX[i] + x0 + j % w] = 1;
        }
    }
}

hipError_t launch_fallback_masks(const FallbackMaskArgs &a, hipStream_t s)
{
    if (a.need_bytes == 0 || a.max_tiles < 1) return hipSuccess;
    const size_t per = (size_t)kBlock * 16;
    const int clear_blocks = (int)((a.need_bytes + per - 1) / per < 256 ? (a.need_bytes + per - 1) / per : 256);
    hipLaunchKernelGGL(k_fallback_clear, dim3(clear_blocks), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_fallback_mark, dim3(min(a.max_tiles, 1024)), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

__global__ void k_fallback_tally(const int *__restrict__ count, unsigned long long *__restrict__ tally)
{
    const int n = *count;
    tally[0] += (unsigned long long)n;
    tally[1] += n > 0 ? 1ull : 0ull;
}

hipError_t launch_fallback_tally(const int *count, unsigned long long *tally, hipStream_t s)
{
    hipLaunchKernelGGL(k_fallback_tally, dim3(1), dim3(1), 0, s, count, tally);
    return hipGetLastError();
}

// ---- column statistics (FP64, fixed order) and scaling --------------------------
// partial[b][0..cols) = sum, partial[b][cols..2 cols) = sum of squares over
// the rows of block b's contiguous chunk.
__global__ __launch_bounds__(kBlock) void k_col_stats(const float *__restrict__ x, size_t rows, int cols,
                                                      double *__restrict__ partial)
{
    __shared__ double s1[kBlock], s2[kBlock];
    const int groups = kBlock / cols;
    const int col = threadIdx.x % cols, grp = threadIdx.x / cols;
    const size_t chunk = (rows + gridDim.x - 1) / gridDim.x;
    const size_t r0 = (size_t)blockIdx.x * chunk, r1 = r0 + chunk < rows ? r0 + chunk : rows;
    double a = 0.0, b = 0.0;
    if (grp < groups)
        for (size_t r = r0 + grp; r < r1; r += groups) {
            const double v = x[r * cols + col];
            a += v;
            b += v * v;
        }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    if ((int)threadIdx.x < cols) {
        double ta = 0.0, tb = 0.0;
        for (int g = 0; g < groups; ++g) {
            ta += s1[g * cols + threadIdx.x];
            tb += s2[g * cols + threadIdx.x];
        }
        partial[(size_t)blockIdx.x * 2 * cols + threadIdx.x] = ta;
        partial[(size_t)blockIdx.x * 2 * cols + cols + threadIdx.x] = tb;
    }
}

hipError_t launch_col_stats(const float *x, size_t rows, int cols, double *partial, int parts, hipStream_t s)
{
    if (cols < 1 || cols > kBlock) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_col_stats, dim3(parts), dim3(kBlock), 0, s, x, rows, cols, partial);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_scale(float *__restrict__ x, size_t n, float k)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) x[i] *= k;
}

hipError_t launch_scale(float *x, size_t n, float k, hipStream_t s)
{
    size_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) return hipSuccess;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)blocks), dim3(kBlock), 0, s, x, n, k);
    return hipGetLastError();
}

}  // namespace rc
