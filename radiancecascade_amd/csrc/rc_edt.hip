// rc_edt.hip -- the exact distance field: occupancy bits, per-column vertical
// distances and per-row lower envelopes.
//
// The reference has no distance field (its only pass is the forward raster,
// src/renderer.rs:559-613).  The field the march reads is the one-byte code
// (rc_common.hpp) of d = sqrt(d^2), where
//   d^2(x, y) = min over occupied (x', y') of (x - x')^2 + (y - y')^2
// is the exact squared Euclidean distance transform of the occupancy (the
// occupancy rule of the seed kernel).  It is computed in its separable form
// (Felzenszwalb & Huttenlocher, "Distance Transforms of Sampled Functions",
// Theory of Computing 8, 2012, who build the row envelope with a stack):
//   g(x, y)   = min |y - y'| over occupied (x, y')          (per column)
//   d^2(x, y) = min over x' of (x - x')^2 + g(x', y)^2       (per row)
// d^2 is unique, so any exact method gives the same field; the kernels below
// are built for a 64-wide SIMT machine, not after that stack algorithm:
//   k_occ_bits : the occupancy of 64 x 64-pixel blocks as column words,
//                bits[y / 64][x] bit y % 64 (one ballot-free LDS transpose of
//                16 waves' 4-row nibbles), and the interleaved emission of
//                occupied pixels for the merge kernels;
//   k_edt_ends : per 64-row band and column, the nearest occupied rows above
//                and below the band (one sweep per column and direction, 8
//                words in flight): with the band's word they give g(x, y) in
//                a few bit operations, so g itself is never stored;
//   k_edt_rows : one workgroup per row, the row's g (from the words and the
//                band ends) in LDS.  The smallest argmin
//                opt(x) of (x - x')^2 + g(x')^2 never decreases with x (for
//                x1 < x2 and j < opt(x1), f_j(x2) - f_opt(x2) = f_j(x1) -
//                f_opt(x1) + 2 (x2 - x1)(opt - j) > 0), so it is found by divide
//                and conquer: with 2^K > W, level k solves x = (2q + 1) 2^(K-1-k)
//                - 1 over the candidates between the argmins the earlier levels
//                fixed, about W + 2^k pairs per level, balanced over the lanes.
// Every step is exact integer arithmetic, so the field equals any exact
// transform's bit for bit.
#include "rc_internal.hpp"

namespace rc {

constexpr uint32_t kGNone = 0xFFFFu;          // g of a column with no occupied pixel
constexpr int kOccBlock = 1024;               // k_occ_bits: 16 waves x 4 rows = 64 rows
constexpr int kRowLanes = 256;                // k_edt_rows: lanes per row (4 waves)
// g of an empty column in the row kernel's LDS: its square (~2^30) is above
// every real value dx^2 + g^2 <= 2 * 16383^2 < 2^29 and dx^2 + 32767^2 < 2^32,
// so the pair value needs no clamp
constexpr uint32_t kGEmpty = 32767u;

__host__ __device__ inline int edt_bands(int H) { return (H + 63) >> 6; }
size_t edt_bits_bytes(int W, int H) { return (size_t)edt_bands(H) * W * sizeof(uint64_t); }
size_t edt_ends_bytes(int W, int H) { return 2 * (size_t)edt_bands(H) * W * sizeof(uint16_t); }

// ---- occupancy as column words --------------------------------------------
__device__ __forceinline__ void occ_bits_body(const float *__restrict__ depth, const DepthConv &dc,
                                              const float *__restrict__ emissive, uint64_t *__restrict__ bits,
                                              float4 *__restrict__ emis4, int W, int H, int bx_count, int b)
{
    __shared__ uint8_t nib[16][64];
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const int x = (b % bx_count) * 64 + lane;
    const int band = b / bx_count;
    const int y0 = band * 64 + 4 * w;
    uint32_t m = 0;
    if (x < W && y0 < H) {
        bool occ[4];
        float4 e[4];
        occupancy4(depth, dc, emissive, W, H, x, y0, occ, e);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (occ[j]) {
                m |= 1u << j;
                if (emis4) emis4[(size_t)(y0 + j) * W + x] = e[j];
            }
    }
    nib[w][lane] = (uint8_t)m;
    __syncthreads();
    if (w == 0 && x < W) {
        uint64_t word = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) word |= (uint64_t)nib[k][lane] << (4 * k);
        bits[(size_t)band * W + x] = word;
    }
}

__global__ __launch_bounds__(kOccBlock) void k_occ_bits(const float *__restrict__ depth, DepthConv dc,
                                                        const float *__restrict__ emissive,
                                                        uint64_t *__restrict__ bits, float4 *__restrict__ emis4, int W,
                                                        int H, int bx_count, const int *gate)
{
    if (gate && *gate == 0) return;
    occ_bits_body(depth, dc, emissive, bits, emis4, W, H, bx_count, xcd_band(blockIdx.x, gridDim.x));
}

__global__ __launch_bounds__(kOccBlock) void k_occ_batch(const OccJob *__restrict__ jobs,
                                                         const int *__restrict__ begins, int njobs)
{
    const int b = xcd_band(blockIdx.x, gridDim.x);
    const int j = batch_job(begins, njobs, b);
    const OccJob &J = jobs[j];
    occ_bits_body(J.depth, J.dc, J.emissive, J.bits, J.emis4, J.W, J.H, J.bx_count, b - begins[j]);
}

// ---- band ends -------------------------------------------------------------
// For every 64-row band w and column x, the nearest occupied row above the
// band and below it: ends[w][x] (rows < 64 w) and ends[nb + w][x] (rows >=
// 64 (w + 1)), 0xFFFF for none.  One lane sweeps one column, down for the
// rows above (lanes [0, W)) or up for the rows below (lanes [W, 2W)), its
// band words loaded 8 at a time.  With the band's own word, these give any
// pixel's vertical distance g in a few bit operations (edt_g).
__device__ __forceinline__ void edt_ends_body(const uint64_t *__restrict__ bits, uint16_t *__restrict__ ends, int W,
                                              int H, int i)
{
    const int nb = edt_bands(H);
    if (i >= 2 * W) return;
    const bool down = i < W;
    const int x = down ? i : i - W;
    uint32_t near = kGNone;  // the occupied row nearest the band boundary reached so far
    for (int k0 = 0; k0 < nb; k0 += 8) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int w = down ? k0 + k : nb - 1 - (k0 + k);
            v[k] = k0 + k < nb ? bits[(size_t)w * W + x] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k0 + k >= nb) break;
            const int w = down ? k0 + k : nb - 1 - (k0 + k);
            ends[(size_t)(down ? w : nb + w) * W + x] = (uint16_t)near;
            if (v[k]) near = (uint32_t)(down ? w * 64 + 63 - __builtin_clzll(v[k]) : w * 64 + __builtin_ctzll(v[k]));
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_edt_ends(const uint64_t *__restrict__ bits, uint16_t *__restrict__ ends,
                                                     int W, int H, const int *gate)
{
    if (gate && *gate == 0) return;
    edt_ends_body(bits, ends, W, H, (int)(blockIdx.x * kBlock + threadIdx.x));
}

__global__ __launch_bounds__(kBlock) void k_ends_batch(const EndsJob *__restrict__ jobs,
                                                       const int *__restrict__ begins, int njobs)
{
    const int b = (int)blockIdx.x;
    const int j = batch_job(begins, njobs, b);
    const EndsJob &J = jobs[j];
    edt_ends_body(J.bits, J.ends, J.W, J.H, (b - begins[j]) * kBlock + (int)threadIdx.x);
}

// Vertical distance of pixel (x, y) to the nearest occupied pixel of its
// column (kGNone: the column has none) from the band's word, else the band
// ends (r = y % 64).
__device__ __forceinline__ uint32_t edt_g_of(uint64_t word, uint32_t ua, uint32_t db, int r, int y)
{
    const uint64_t le = word & ((2ull << r) - 1);  // rows <= r of the band (all 64 for r = 63)
    const uint64_t ge = word >> r;                 // rows >= r
    // both candidates, then a select (no branch for the loads to sink into;
    // the sentinel bits keep clz / ctz defined, their value unused then)
    const uint32_t up_in = (uint32_t)(r - 63 + __builtin_clzll(le | 1ull));
    const uint32_t dn_in = (uint32_t)__builtin_ctzll(ge | (1ull << 63));
    const uint32_t up_out = ua != kGNone ? (uint32_t)y - ua : kGNone;
    const uint32_t dn_out = db != kGNone ? db - (uint32_t)y : kGNone;
    const uint32_t up = le ? up_in : up_out, dn = ge ? dn_in : dn_out;
    return up < dn ? up : dn;
}

// ---- per-row lower envelope -------------------------------------------------------
// floor(sqrt(d2)) for d2 < 2^29: the hardware square root (within an ulp, so
// within 0.01 of the root here) truncated, then one integer correction step
// each way, which is all the truncation can be off by.
__device__ __forceinline__ uint32_t edt_isqrt(uint32_t d2)
{
    uint32_t r = (uint32_t)__builtin_amdgcn_sqrtf((float)d2);  // r < 2^15: 24-bit squares
    r -= __umul24(r, r) > d2 ? 1u : 0u;
    r += __umul24(r + 1, r + 1) <= d2 ? 1u : 0u;
    return r;
}

// The argmin of a solved pixel x: the low 16 bits of its 64-bit key
// (value << 16 | argmin), read through a 16-bit view of the key array.
__device__ __forceinline__ int edt_opt(const uint16_t *key16, int x) { return key16[4 * x]; }
// Candidate range [lo, hi] of query q (at x = (2q + 1) s - 1) at a level of
// step s: the argmins the earlier levels fixed at x -+ s (0 and W - 1 beyond
// the ends).
__device__ __forceinline__ int edt_lo(int q, int s, const uint16_t *key16)
{
    return q == 0 ? 0 : edt_opt(key16, 2 * q * s - 1);
}
__device__ __forceinline__ int edt_hi(int q, int s, int W, const uint16_t *key16)
{
    return (2 * q + 2) * s > W ? W - 1 : edt_opt(key16, (2 * q + 2) * s - 1);
}

// Row y of a W-pixel image (W <= WMAX) by one workgroup of RL lanes.
// Consecutive queries of a level share an end (hi(q) = lo(q + 1)), so the
// level's (query, candidate) pairs form one staircase of T = hi(nq - 1) + nq
// steps: query q owns the steps [lo(q) + q, hi(q) + q], step t of it being
// candidate j = t - q.  The work per query is very uneven (a query whose
// neighbours' argmins lie on either side of a wide gap spans the gap), so
// the lanes take equal runs of steps instead: a lane finds the query of its
// first step by binary search on lo(q) + q (the argmins in LDS), walks its
// run in one loop of uniform trip count, keeping the first smallest value
// (ties go to the smallest candidate), and meets the other lanes of a query
// by a 64-bit LDS atomic minimum of (value << 16 | candidate) at the query's
// own pixel.  Every pixel is solved at exactly one level, so the keys start
// at all ones and need no reset between levels, and the next level reads the
// argmins straight from them: one barrier per level.
template <int WMAX, int RL>
__device__ __forceinline__ void edt_row_body(const uint64_t *__restrict__ bits, const uint16_t *__restrict__ ends,
                                             uint8_t *__restrict__ sdf, int W, int H, int y)
{
    __shared__ uint16_t gs[WMAX];
    __shared__ unsigned long long key[WMAX];  // key[x] = value << 16 | smallest argmin, all ones until solved
    const uint16_t *opt = reinterpret_cast<const uint16_t *>(key);
    const int tid = (int)threadIdx.x;
    const int nb = edt_bands(H);
    {
        // the row's g: every load of a batch of columns issued before any is
        // used (one round trip per batch, not one or two per column)
        constexpr int GB = 8;
        const int w = y >> 6, r = y & 63;
        const uint64_t *brow = bits + (size_t)w * W;
        const uint16_t *urow = ends + (size_t)w * W, *drow = ends + (size_t)(nb + w) * W;
        for (int x0 = tid; x0 < W; x0 += GB * RL) {
            uint64_t wd[GB];
            uint32_t ua[GB], db[GB];
#pragma unroll
            for (int i = 0; i < GB; ++i) {
                const int x = min(x0 + i * RL, W - 1);
                wd[i] = brow[x];
                ua[i] = urow[x];
                db[i] = drow[x];
            }
            // the values are pinned here, so the compiler neither sinks a
            // load into the branch below nor issues it again there
#pragma unroll
            for (int i = 0; i < GB; ++i) asm volatile("" : "+v"(wd[i]), "+v"(ua[i]), "+v"(db[i]));
#pragma unroll
            for (int i = 0; i < GB; ++i) {
                const int x = x0 + i * RL;
                if (x < W) {
                    gs[x] = (uint16_t)min(edt_g_of(wd[i], ua[i], db[i], r, y), kGEmpty);
                    key[x] = ~0ull;
                }
            }
        }
    }
    __syncthreads();
    const int K = 32 - __builtin_clz((uint32_t)W);  // smallest K with 2^K > W
    for (int k = 0; k < K; ++k) {
        const int ls = K - 1 - k, s = 1 << ls;
        const int nq = ((W >> ls) + 1) >> 1;  // odd multiples of s up to W
        const int T = edt_hi(nq - 1, s, W, opt) + nq;
        const int c = (T + RL - 1) / RL;
        const int t0 = tid * c, t1 = min(t0 + c, T);
        if (t0 < t1) {
            // the last query starting at or before t0: a lower bound of
            // uniform width (a + n <= nq throughout, so m <= nq - 1)
            int a = 0;
            for (int n = nq; n > 1;) {
                const int h = n >> 1, m = a + h;
                if (edt_opt(opt, (m << (ls + 1)) - 1) + m <= t0) a = m;  // lo(m) + m
                n -= h;
            }
            // branch-free values: |x - j| < 2^14 and g <= kGEmpty fit the
            // 24-bit multiplier, and dx^2 + g^2 the 32 bits
            int q = a, x = (2 * q + 1) * s - 1;
            int end = edt_hi(q, s, W, opt) + q + 1;  // the first step of query q + 1
            uint32_t bv = 0xFFFFFFFFu, bj = 0;
            // key = value << 16 | candidate from 32-bit halves (value < 2^30)
            auto flush = [&]() {
                atomicMin(&key[x], ((unsigned long long)(bv >> 16) << 32) | ((bv << 16) | bj));
            };
            // candidate j = t - q and dx = x - j carried step to step: the
            // next step of a query is j + 1, dx - 1; entering the next query
            // keeps j (t and q both advance) and moves x on by 2 s
            int j = t0 - q, dx = x - j;
            for (int t = t0; t < t1; ++t, ++j, --dx) {
                if (t >= end) {  // the run enters the next query (sizes >= 1: one step)
                    flush();
                    bv = 0xFFFFFFFFu;
                    ++q;
                    x += 2 * s;
                    --j;
                    dx += 2 * s + 1;
                    end = edt_hi(q, s, W, opt) + q + 1;
                }
                const uint32_t gj = gs[j], g2 = (uint32_t)__umul24(gj, gj);
                const uint32_t v = (uint32_t)__mul24(dx, dx) + g2;
                if (v < bv) {
                    bv = v;
                    bj = (uint32_t)j;
                }
            }
            flush();
        }
        __syncthreads();
    }
    // the byte of (x, y) at x + 16 y + (x >> 4) skip (sdf_index) in 32-bit
    // arithmetic: x >> 4 < 2^10 and skip < 2^18 fit the 24-bit multiplier
    const uint32_t skip = (uint32_t)sdf_skip(H), row = (uint32_t)y << 4;
    for (int x = tid; x < W; x += RL) {
        const int j = edt_opt(opt, x);
        const uint32_t gj = gs[j];
        uint32_t code = kSdfFar;
        if (gj != kGEmpty) {  // |x - j| < 2^14 and g < 2^14: d^2 < 2^29
            const int dx = x - j;
            code = sdf_code_of_distance(edt_isqrt((uint32_t)__mul24(dx, dx) + (uint32_t)__umul24(gj, gj)));
        }
        sdf[__umul24((uint32_t)x >> 4, skip) + (uint32_t)x + row] = (uint8_t)code;
    }
}

template <int WMAX>
__global__ __launch_bounds__(kRowLanes) void k_edt_rows(const uint64_t *__restrict__ bits,
                                                        const uint16_t *__restrict__ ends, uint8_t *__restrict__ sdf,
                                                        int W, int H, const int *gate)
{
    if (gate && *gate == 0) return;
    edt_row_body<WMAX, kRowLanes>(bits, ends, sdf, W, H, xcd_band(blockIdx.x, gridDim.x));
}

template <int WMAX>
__global__ __launch_bounds__(kRowLanes) void k_rows_batch(const RowJob *__restrict__ jobs,
                                                          const int *__restrict__ begins, int njobs)
{
    const int b = xcd_band(blockIdx.x, gridDim.x);
    const int j = batch_job(begins, njobs, b);
    const RowJob &J = jobs[j];
    edt_row_body<WMAX, kRowLanes>(J.bits, J.ends, J.sdf, J.W, J.H, b - begins[j]);
}

// ---- launchers ---------------------------------------------------------------------
int occ_blocks(int W, int H) { return ((W + 63) / 64) * edt_bands(H); }
int occ_bx(int W) { return (W + 63) / 64; }
int ends_blocks(int W) { return (2 * W + kBlock - 1) / kBlock; }
int edt_row_wmax(int W)
{
    int m = 1024;
    while (m < W) m <<= 1;
    return m;
}

template <int WMAX>
static void launch_rows_t(const uint64_t *bits, const uint16_t *ends, uint8_t *sdf, int W, int H, hipStream_t s,
                          const int *gate)
{
    hipLaunchKernelGGL(k_edt_rows<WMAX>, dim3(H), dim3(kRowLanes), 0, s, bits, ends, sdf, W, H, gate);
}

hipError_t launch_edt(const float *depth, const DepthConv &dc, const float *emissive, uint64_t *bits, uint16_t *ends,
                      uint8_t *sdf, float4 *emis4, int W, int H, hipStream_t s, const int *gate, int stage)
{
    if (W < 1 || H < 1 || W > RC_MAX_DIM || H > RC_MAX_DIM) return hipErrorInvalidValue;
    const int bx = occ_bx(W);
    if (stage & 1)
        hipLaunchKernelGGL(k_occ_bits, dim3(occ_blocks(W, H)), dim3(kOccBlock), 0, s, depth, dc, emissive, bits, emis4,
                           W, H, bx, gate);
    if (stage & 2) {
        const uint64_t *cb = bits;
        hipLaunchKernelGGL(k_edt_ends, dim3(ends_blocks(W)), dim3(kBlock), 0, s, cb, ends, W, H, gate);
        switch (edt_row_wmax(W)) {
        case 1024: launch_rows_t<1024>(cb, ends, sdf, W, H, s, gate); break;
        case 2048: launch_rows_t<2048>(cb, ends, sdf, W, H, s, gate); break;
        case 4096: launch_rows_t<4096>(cb, ends, sdf, W, H, s, gate); break;
        case 8192: launch_rows_t<8192>(cb, ends, sdf, W, H, s, gate); break;
        default: launch_rows_t<16384>(cb, ends, sdf, W, H, s, gate); break;
        }
    }
    return hipGetLastError();
}

hipError_t launch_occ_batch(const BatchLaunch &b, hipStream_t s)
{
    if (b.blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_occ_batch, dim3(b.blocks), dim3(kOccBlock), 0, s, (const OccJob *)b.jobs, b.begins, b.njobs);
    return hipGetLastError();
}

hipError_t launch_ends_batch(const BatchLaunch &b, hipStream_t s)
{
    if (b.blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ends_batch, dim3(b.blocks), dim3(kBlock), 0, s, (const EndsJob *)b.jobs, b.begins, b.njobs);
    return hipGetLastError();
}

hipError_t launch_rows_batch(const BatchLaunch &b, int wmax, hipStream_t s)
{
    if (b.blocks <= 0) return hipSuccess;
    const RowJob *j = (const RowJob *)b.jobs;
    switch (wmax) {
    case 1024: hipLaunchKernelGGL(k_rows_batch<1024>, dim3(b.blocks), dim3(kRowLanes), 0, s, j, b.begins, b.njobs); break;
    case 2048: hipLaunchKernelGGL(k_rows_batch<2048>, dim3(b.blocks), dim3(kRowLanes), 0, s, j, b.begins, b.njobs); break;
    case 4096: hipLaunchKernelGGL(k_rows_batch<4096>, dim3(b.blocks), dim3(kRowLanes), 0, s, j, b.begins, b.njobs); break;
    case 8192: hipLaunchKernelGGL(k_rows_batch<8192>, dim3(b.blocks), dim3(kRowLanes), 0, s, j, b.begins, b.njobs); break;
    case 16384:
        hipLaunchKernelGGL(k_rows_batch<16384>, dim3(b.blocks), dim3(kRowLanes), 0, s, j, b.begins, b.njobs);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rc
