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
//   k_edt_cols : g of one 64-row band of 64 columns per wave: the band's word
//                and the nearest non-empty words above and below (8 words per
//                round in flight), 16-bit g rows (0xFFFF: the column is empty);
//   k_edt_rows : one workgroup per row, the row's g in LDS.  The smallest argmin
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
constexpr uint32_t kEdtBig = 1u << 29;        // value of an empty column: above any real 2 * 16383^2

__host__ __device__ inline int edt_bands(int H) { return (H + 63) >> 6; }
size_t edt_bits_bytes(int W, int H) { return (size_t)edt_bands(H) * W * sizeof(uint64_t); }
size_t edt_g_bytes(int W, int H) { return (size_t)W * H * sizeof(uint16_t); }

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

// ---- vertical distances -------------------------------------------------------
// Wave `wid` of a view: columns 64 (wid % bx_count) + lane of band wid / bx_count.
__device__ __forceinline__ void edt_cols_body(const uint64_t *__restrict__ bits, uint16_t *__restrict__ g, int W,
                                              int H, int bx_count, int wid)
{
    const int nb = edt_bands(H);
    const int band = wid / bx_count;
    const int x = (wid % bx_count) * 64 + (int)(threadIdx.x & 63);
    if (band >= nb || x >= W) return;  // no cross-lane operation follows
    const uint64_t word = bits[(size_t)band * W + x];
    // nearest occupied row above the band and below it (-1: none)
    int above = -1, below = -1;
    for (int w0 = band - 1; w0 >= 0 && above < 0; w0 -= 8) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = w0 - k >= 0 ? bits[(size_t)(w0 - k) * W + x] : 0ull;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (above < 0 && v[k]) above = (w0 - k) * 64 + 63 - __builtin_clzll(v[k]);
    }
    for (int w0 = band + 1; w0 < nb && below < 0; w0 += 8) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = w0 + k < nb ? bits[(size_t)(w0 + k) * W + x] : 0ull;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (below < 0 && v[k]) below = (w0 + k) * 64 + __builtin_ctzll(v[k]);
    }
    const int y0 = band * 64, rows = min(64, H - y0);
    uint16_t *out = g + (size_t)y0 * W + x;
    for (int r = 0; r < rows; ++r) {
        const int y = y0 + r;
        const uint64_t le = word & ((2ull << r) - 1);  // rows <= r of the band (all 64 for r = 63)
        const uint64_t ge = word >> r;                 // rows >= r
        const uint32_t up = le ? (uint32_t)(r - 63 + __builtin_clzll(le)) : (above >= 0 ? (uint32_t)(y - above) : kGNone);
        const uint32_t dn = ge ? (uint32_t)__builtin_ctzll(ge) : (below >= 0 ? (uint32_t)(below - y) : kGNone);
        out[(size_t)r * W] = (uint16_t)min(up, dn);
    }
}

__global__ __launch_bounds__(kBlock) void k_edt_cols(const uint64_t *__restrict__ bits, uint16_t *__restrict__ g,
                                                     int W, int H, int bx_count, const int *gate)
{
    if (gate && *gate == 0) return;
    edt_cols_body(bits, g, W, H, bx_count, xcd_band(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6));
}

__global__ __launch_bounds__(kBlock) void k_cols_batch(const ColJob *__restrict__ jobs,
                                                       const int *__restrict__ begins, int njobs)
{
    const int b = xcd_band(blockIdx.x, gridDim.x);
    const int j = batch_job(begins, njobs, b);
    const ColJob &J = jobs[j];
    edt_cols_body(J.bits, J.g, J.W, J.H, J.bx_count, (b - begins[j]) * 4 + (int)(threadIdx.x >> 6));
}

// ---- per-row lower envelope -------------------------------------------------------
// Exclusive prefix sum of v over the RL lanes of the workgroup; `total` is the
// sum over all lanes.  wsum: RL / 64 words of LDS.
template <int RL>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t n = __shfl_up(inc, o, 64);
        if (lane >= o) inc += n;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < RL / 64; ++w) {
        const uint32_t sw = wsum[w];
        off += w < wave ? sw : 0u;
        total += sw;
    }
    return off + inc - v;
}

// Candidate range [lo, hi] of query q (at x = (2q + 1) s - 1) at a level of
// step s: the argmins the earlier levels fixed at x -+ s (0 and W - 1 beyond
// the ends).
__device__ __forceinline__ int edt_lo(int q, int s, const uint16_t *opt) { return q == 0 ? 0 : opt[2 * q * s - 1]; }
__device__ __forceinline__ int edt_hi(int q, int s, int W, const uint16_t *opt)
{
    return (2 * q + 2) * s > W ? W - 1 : opt[(2 * q + 2) * s - 1];
}

// Row y of a W-pixel image (W <= WMAX) by one workgroup of RL lanes.  A
// level's work is the list of its (query, candidate) pairs, about W + 2^k
// long but very uneven per query (a query whose neighbours' argmins lie on
// either side of a wide gap spans the gap), so it is balanced over the
// lanes: the queries' range sizes are prefix-summed, each lane takes an
// equal run of consecutive pairs (finding its first query by binary search
// over the prefix sums, then walking forward, keeping the first smallest
// value so ties go to the smallest candidate), and the per-query minima meet
// in LDS by a 64-bit atomic minimum of (value << 14 | candidate).
template <int WMAX, int RL>
__device__ __forceinline__ void edt_row_body(const uint16_t *__restrict__ g, uint8_t *__restrict__ sdf, int W, int y)
{
    __shared__ uint16_t gs[WMAX];
    __shared__ uint16_t opt[WMAX];                 // opt[x]: smallest argmin at x
    __shared__ unsigned long long best[WMAX / 2];  // per query of the current level (at most ceil(W / 2))
    __shared__ uint16_t pre[WMAX / 2 + 1];         // exclusive prefix of the range sizes (total <= W + W / 2)
    __shared__ uint32_t wsum[RL / 64];
    const int tid = (int)threadIdx.x;
    const uint16_t *gr = g + (size_t)y * W;
    for (int x = tid; x < W; x += RL) gs[x] = gr[x];
    __syncthreads();
    const int K = 32 - __builtin_clz((uint32_t)W);  // smallest K with 2^K > W
    for (int k = 0; k < K; ++k) {
        const int ls = K - 1 - k, s = 1 << ls;
        const int nq = ((W >> ls) + 1) >> 1;  // odd multiples of s up to W
        const int per = (nq + RL - 1) / RL;
        const int qa = min(tid * per, nq), qb = min(qa + per, nq);
        uint32_t local = 0;
        for (int q = qa; q < qb; ++q) local += (uint32_t)(edt_hi(q, s, W, opt) - edt_lo(q, s, opt) + 1);
        uint32_t total;
        uint32_t off = block_excl_scan<RL>(local, wsum, total);
        for (int q = qa; q < qb; ++q) {
            pre[q] = (uint16_t)off;
            off += (uint32_t)(edt_hi(q, s, W, opt) - edt_lo(q, s, opt) + 1);
            best[q] = ~0ull;
        }
        if (tid == 0) pre[nq] = (uint16_t)total;
        __syncthreads();
        const uint32_t c = (total + RL - 1) / RL;
        const uint32_t t0 = (uint32_t)tid * c, t1 = min(t0 + c, total);
        if (t0 < t1) {
            int a = 0, b = nq;  // largest q with pre[q] <= t0
            while (b - a > 1) {
                const int m = (a + b) >> 1;
                if (pre[m] <= t0) a = m;
                else b = m;
            }
            // one uniform-count loop over the lane's run (the step into the
            // next query is a short divergent branch); branch-free values:
            // |x - j| < 2^14 and g <= 0xFFFF fit the 24-bit multiplier, and an
            // empty column's g^2 (> 2^29) is clamped to kEdtBig, above every
            // real value
            int q = a;
            int x = (2 * q + 1) * s - 1, lo = edt_lo(q, s, opt);
            uint32_t start = pre[q], end = pre[q + 1];
            uint32_t bv = 0xFFFFFFFFu, bj = 0;
            for (uint32_t t = t0; t < t1; ++t) {
                if (t >= end) {
                    do {  // flush, then enter the query holding step t
                        atomicMin(&best[q], ((unsigned long long)bv << 14) | bj);
                        bv = 0xFFFFFFFFu;
                        ++q;
                        start = end;
                        end = pre[q + 1];
                    } while (t >= end);
                    x = (2 * q + 1) * s - 1;
                    lo = opt[2 * q * s - 1];
                }
                const int j = lo + (int)(t - start), dx = x - j;
                const uint32_t gj = gs[j], g2 = (uint32_t)__umul24(gj, gj);
                const uint32_t v = (uint32_t)__mul24(dx, dx) + (g2 < kEdtBig ? g2 : kEdtBig);
                if (v < bv) {
                    bv = v;
                    bj = (uint32_t)j;
                }
            }
            atomicMin(&best[q], ((unsigned long long)bv << 14) | bj);
        }
        __syncthreads();
        for (int q = qa; q < qb; ++q) opt[(2 * q + 1) * s - 1] = (uint16_t)(best[q] & 0x3FFFu);
        __syncthreads();
    }
    const int tx = sdf_tiles_x(W);
    for (int x = tid; x < W; x += RL) {
        const int j = opt[x];
        const uint32_t gj = gs[j];
        uint32_t code = kSdfFar;
        if (gj != kGNone) {
            const int dx = x - j;
            code = sdf_code_of_distance(isqrt_u32((uint32_t)(dx * dx) + gj * gj));
        }
        sdf[sdf_index(x, y, tx)] = (uint8_t)code;
    }
}

template <int WMAX>
__global__ __launch_bounds__(kRowLanes) void k_edt_rows(const uint16_t *__restrict__ g, uint8_t *__restrict__ sdf,
                                                        int W, const int *gate)
{
    if (gate && *gate == 0) return;
    edt_row_body<WMAX, kRowLanes>(g, sdf, W, xcd_band(blockIdx.x, gridDim.x));
}

template <int WMAX>
__global__ __launch_bounds__(kRowLanes) void k_rows_batch(const RowJob *__restrict__ jobs,
                                                          const int *__restrict__ begins, int njobs)
{
    const int b = xcd_band(blockIdx.x, gridDim.x);
    const int j = batch_job(begins, njobs, b);
    const RowJob &J = jobs[j];
    edt_row_body<WMAX, kRowLanes>(J.g, J.sdf, J.W, b - begins[j]);
}

// ---- launchers ---------------------------------------------------------------------
int occ_blocks(int W, int H) { return ((W + 63) / 64) * edt_bands(H); }
int occ_bx(int W) { return (W + 63) / 64; }
int cols_blocks(int W, int H) { return (occ_blocks(W, H) + 3) / 4; }
int edt_row_wmax(int W)
{
    int m = 1024;
    while (m < W) m <<= 1;
    return m;
}

template <int WMAX>
static void launch_rows_t(const uint16_t *g, uint8_t *sdf, int W, int H, hipStream_t s, const int *gate)
{
    hipLaunchKernelGGL(k_edt_rows<WMAX>, dim3(H), dim3(kRowLanes), 0, s, g, sdf, W, gate);
}

hipError_t launch_edt(const float *depth, const DepthConv &dc, const float *emissive, uint64_t *bits, uint16_t *g,
                      uint8_t *sdf, float4 *emis4, int W, int H, hipStream_t s, const int *gate, int stage)
{
    if (W < 1 || H < 1 || W > RC_MAX_DIM || H > RC_MAX_DIM) return hipErrorInvalidValue;
    const int bx = occ_bx(W);
    if (stage & 1)
        hipLaunchKernelGGL(k_occ_bits, dim3(occ_blocks(W, H)), dim3(kOccBlock), 0, s, depth, dc, emissive, bits, emis4,
                           W, H, bx, gate);
    if (stage & 2) {
        hipLaunchKernelGGL(k_edt_cols, dim3(cols_blocks(W, H)), dim3(kBlock), 0, s, (const uint64_t *)bits, g, W, H,
                           bx, gate);
        switch (edt_row_wmax(W)) {
        case 1024: launch_rows_t<1024>(g, sdf, W, H, s, gate); break;
        case 2048: launch_rows_t<2048>(g, sdf, W, H, s, gate); break;
        case 4096: launch_rows_t<4096>(g, sdf, W, H, s, gate); break;
        case 8192: launch_rows_t<8192>(g, sdf, W, H, s, gate); break;
        default: launch_rows_t<16384>(g, sdf, W, H, s, gate); break;
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

hipError_t launch_cols_batch(const BatchLaunch &b, hipStream_t s)
{
    if (b.blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cols_batch, dim3(b.blocks), dim3(kBlock), 0, s, (const ColJob *)b.jobs, b.begins, b.njobs);
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
