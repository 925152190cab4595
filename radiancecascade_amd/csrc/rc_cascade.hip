// rc_cascade.hip -- per-level interval march, top-down bilinear merge,
// 4-child direction reduction, and the level-0 gather to irradiance.
//
// None of this exists in the reference (its only pass is the forward raster,
// src/renderer.rs:559-613); the maths is Sannikov's radiance cascades (2023):
//
//   march  : from c = ((px+1/2)s, (py+1/2)s) along theta_k, t from t0: sample
//            q = floor(c + t dir) (fused multiply-add per axis); off-image ->
//            (0, alpha 1); SDF(q) < 1/2 -> (E(q), alpha 0); else
//            t += max(SDF(q), 1); t >= t1 -> (0, alpha 1).
//   merge  : R_i(p,k) = march + alpha * bilinear_b A_{i+1}(b, k) over the 4
//            upper probes around p (upper coordinate p/2 - 1/4, indices
//            clamped, weights 3/4 and 1/4: x first, then y).
//   reduce : A_i(p,k') = 1/4 ((R(4k') + R(4k'+1)) + (R(4k'+2) + R(4k'+3))),
//            the only form in which level i-1 reads level i.
//   gather : irradiance(pixel) = bilinear over the level-0 fluence A_0.
//
// Work mapping: a workgroup (4 wave64) owns TX x TY probes x KD consecutive
// directions, lanes direction-fastest.  The 4 children of a level-(i-1)
// direction sit in one lane quad, so the reduction is two DPP quad
// permutations and no LDS.  Before marching, the workgroup stages the upper
// level's probe neighbourhood for its direction slice in LDS (one global read
// per texel instead of 16 L2 reads per ray).  Rays that hit skip the merge.
#include "rc_internal.hpp"
#include "rc_march.hpp"

namespace rc {

// One ray with SPEC-deep speculation.  After a sample with distance d the
// next sample is exactly t + max(d, 1); a ray grazing a surface keeps taking
// the same step, so each round also loads the SPEC-1 samples that follow if
// the step stays unchanged (positions built with the same fp32 additions the
// sequential march would make).  The samples are then checked in order; the
// run of accepted ones ends at the first sample whose step differs, which
// starts the next round.  The visited sample sequence, and so every result
// bit, is that of the one-sample-per-step march.
struct Ray {
    float cx, cy, dx, dy, t, guess;
    int code;
    bool alive;
};

__device__ inline Ray ray_begin(float cx, float cy, float dx, float dy, float t0, float t1, bool active)
{
    return Ray{cx, cy, dx, dy, t0, 1.0f, kRayT1, active && t0 < t1};
}

template <int SPEC>
__device__ inline void ray_round(Ray &r, __amdgpu_buffer_rsrc_t sdf, int W, int H, float t1)
{
    float p[SPEC], d[SPEC];
    int q[SPEC];
    bool ins[SPEC];
    p[0] = r.t;
#pragma unroll
    for (int j = 1; j < SPEC; ++j) p[j] = p[j - 1] + r.guess;
#pragma unroll
    for (int j = 0; j < SPEC; ++j) {
        const int ix = floor_i32(fmaf(p[j], r.dx, r.cx));
        const int iy = floor_i32(fmaf(p[j], r.dy, r.cy));
        ins[j] = (uint32_t)ix < (uint32_t)W && (uint32_t)iy < (uint32_t)H;
        q[j] = pixel_index(ix, iy, W);
        d[j] = sdf_load(sdf, ix, iy, sdf_skip(H));
    }
    bool go = true;
#pragma unroll
    for (int j = 0; j < SPEC; ++j) {
        if (!go) continue;
        if (j > 0 && !(p[j] < t1)) {  // the previous step reached the interval end
            r.alive = false;
            go = false;
            continue;
        }
        if (!ins[j]) {
            r.code = kRayExit;
            r.alive = false;
            go = false;
            continue;
        }
        if (d[j] < 0.5f) {
            r.code = q[j];
            r.alive = false;
            go = false;
            continue;
        }
        const float step = march_step(d[j]);
        if (j + 1 < SPEC && step == r.guess) continue;  // p[j+1] == p[j] + step
        r.t = p[j] + step;
        r.guess = step;
        r.alive = r.t < t1;
        go = false;
    }
}

// Level i >= 1.  Workgroup tile: TX x (TY*RPT) probes x KD directions; lane
// (probe column, direction) owns the RPT probes stacked TY rows apart.
template <int TX, int TY, int KD, int RPT, Storage S, bool UPPER, bool FROM_CODES>
__device__ __forceinline__ void level_body(const LevelArgs &a, int lb)
{
    static_assert(TX * TY * KD == kBlock, "a tile row block holds 256 rays");
    using TT = Texel<S>;
    constexpr int TYR = TY * RPT;
    constexpr int UCX = TX / 2 + 2, UCY = TYR / 2 + 2;
    __shared__ float4 su[UPPER ? UCX * UCY * KD : 1];

    const int tiles_x = (a.PX + TX - 1) / TX;
    const int slices = a.N / KD;
    const int slice = lb % slices, tile = lb / slices;
    const int px0 = (tile % tiles_x) * TX, py0 = (tile / tiles_x) * TYR, k0 = slice * KD;

    const int kk = threadIdx.x % KD, pl = threadIdx.x / KD;
    const int px = px0 + pl % TX, pyb = py0 + pl / TX, k = k0 + kk;

    // the tile fallback: nothing when no tile is flagged; only the probes
    // the flagged tiles need (a workgroup without one leaves at once)
    if (a.gate && *a.gate == 0) return;
    bool needed[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int py = pyb + j * TY;
        needed[j] = px < a.PX && py < a.PY && (!a.need || a.need[py * a.PX + px] != 0);
    }
    if (a.need) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < RPT; ++j) any = any || needed[j];
        if (!__syncthreads_or(any)) return;
    }

    // 32-bit indices: a level holds fewer than 2^32 rays (W, H <= 16384), N
    // is a power of two, and probe rows and columns are below 2^24
    const int log2n = 31 - __builtin_clz(a.N);
    const int ulx = (px0 - 1) >> 1, uly = (py0 - 1) >> 1;
    const uint32_t plane = __umul24((uint32_t)a.W, (uint32_t)a.H);
    bool active[RPT];
    int hitq[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) active[j] = needed[j];

    // The upper level's neighbourhood for this direction slice, staged in
    // LDS: UN texels per thread, loaded to registers first.
    constexpr int UTOT = UCX * UCY * KD, UN = UPPER ? (UTOT + kBlock - 1) / kBlock : 0;
    typename TT::T ureg[UN > 0 ? UN : 1];
    auto stage_loads = [&]() {
        const typename TT::T *up = (const typename TT::T *)a.upper;
#pragma unroll
        for (int n = 0; n < UN; ++n) {
            const int i = (int)threadIdx.x + n * kBlock;
            if (i < UTOT) {
                const int j = i % KD, sx = (i / KD) % UCX, sy = i / (KD * UCX);
                const int ux = min(max(ulx + sx, 0), a.UX - 1), uy = min(max(uly + sy, 0), a.UY - 1);
                ureg[n] = up[((__umul24((uint32_t)uy, (uint32_t)a.UX) + (uint32_t)ux) << log2n) + (uint32_t)(k0 + j)];
            }
        }
    };
    auto stage_stores = [&]() {
#pragma unroll
        for (int n = 0; n < UN; ++n) {
            const int i = (int)threadIdx.x + n * kBlock;
            if (i < UTOT) su[i] = TT::unpack(ureg[n]);
        }
        __syncthreads();
    };
    // the emission of each ray's hit pixel (rgb; zero for a miss)
    float3 ev[RPT];
    auto gather_hits = [&]() {
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            ev[j] = make_float3(0.f, 0.f, 0.f);
            if (active[j] && hitq[j] >= 0) {
                const uint32_t q = (uint32_t)hitq[j];
                if (a.emis4) {
                    const float4 e = a.emis4[q];
                    ev[j] = make_float3(e.x, e.y, e.z);
                } else {
                    ev[j] = make_float3(a.emissive[q], a.emissive[plane + q], a.emissive[2 * plane + q]);
                }
            }
        }
    };

    if (FROM_CODES) {
        // The rays were marched by k_march_all.  Their codes, the staging
        // loads and the hits' emission gathers are all issued before the
        // barrier, so the three round trips overlap.
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int py = min(pyb + j * TY, a.PY - 1);
            hitq[j] = a.codes[((__umul24((uint32_t)py, (uint32_t)a.PX) + (uint32_t)min(px, a.PX - 1)) << log2n) +
                              (uint32_t)k];
        }
        if (UPPER) stage_loads();
        gather_hits();
        if (UPPER) stage_stores();
    } else {
        if (UPPER) {
            stage_loads();
            stage_stores();
        }
        const float2 dir = a.dirs[k];
        float cx[RPT], cy[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int py = pyb + j * TY;
            cx[j] = ((float)px + 0.5f) * a.spacing;
            cy[j] = ((float)py + 0.5f) * a.spacing;
        }
        march_rays<RPT>(sdf_rsrc(a.sdf, a.W, a.H), a.W, a.H, cx, cy, dir.x, dir.y, a.t0, a.t1, active, hitq);
        if (a.dbg_codes) {
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int py = pyb + j * TY;
                if (active[j])
                    a.dbg_codes[((__umul24((uint32_t)py, (uint32_t)a.PX) + (uint32_t)px) << log2n) + (uint32_t)k] =
                        hitq[j];
            }
        }
        gather_hits();
    }

#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int py = pyb + j * TY;
        float r = 0.f, g = 0.f, b = 0.f, al = 0.f;
        if (active[j]) {
            if (hitq[j] >= 0) {
                r = ev[j].x;
                g = ev[j].y;
                b = ev[j].z;
            } else {
                al = 1.0f;
                if (UPPER) {
                    const int sx = ((px - 1) >> 1) - ulx, sy = ((py - 1) >> 1) - uly;
                    const float4 u00 = su[(sy * UCX + sx) * KD + kk];
                    const float4 u10 = su[(sy * UCX + sx + 1) * KD + kk];
                    const float4 u01 = su[((sy + 1) * UCX + sx) * KD + kk];
                    const float4 u11 = su[((sy + 1) * UCX + sx + 1) * KD + kk];
                    const float3 v = bilerp4(u00, u10, u01, u11, px, py);
                    r = v.x;
                    g = v.y;
                    b = v.z;
                }
            }
        }
        r = quad_sum(r) * 0.25f;
        g = quad_sum(g) * 0.25f;
        b = quad_sum(b) * 0.25f;
        al = quad_sum(al) * 0.25f;
        if (active[j] && (kk & 3) == 0) {
            typename TT::T *out = (typename TT::T *)a.out;
            out[((__umul24((uint32_t)py, (uint32_t)a.PX) + (uint32_t)px) << (log2n - 2)) + (uint32_t)(k >> 2)] =
                TT::pack(make_float4(r, g, b, al));
        }
    }
}

template <int TX, int TY, int KD, int RPT, Storage S, bool UPPER, bool FROM_CODES>
__global__ __launch_bounds__(kBlock) void k_level(LevelArgs a)
{
    level_body<TX, TY, KD, RPT, S, UPPER, FROM_CODES>(a, xcd_band(blockIdx.x, gridDim.x));
}

// One level over a work list of views (each job a view's LevelArgs; the top
// level of a view has no upper level, so UPPER is taken per job).
template <int TX, int TY, int KD, int RPT, Storage S, bool FROM_CODES>
__global__ __launch_bounds__(kBlock) void k_level_batch(const LevelArgs *__restrict__ jobs,
                                                        const int *__restrict__ begins, int njobs)
{
    const int b = xcd_band(blockIdx.x, gridDim.x);
    const int j = batch_job(begins, njobs, b);
    const LevelArgs a = jobs[j];
    const int lb = b - begins[j];
    if (a.upper) level_body<TX, TY, KD, RPT, S, true, FROM_CODES>(a, lb);
    else level_body<TX, TY, KD, RPT, S, false, FROM_CODES>(a, lb);
}

// Level 0 (4 directions, interval [0,2)) fused with the gather.  A tile owns
// 16x12 level-0 probes = 32x24 pixels (each output row one 128-byte line); a
// thread marches one probe of the tile plus its 1-probe halo (18x14 = 252
// probes, the pixels' bilinear footprint), all 4 directions at once, and
// writes the probe's fluence to LDS; the pixels then read it bilinearly and
// write planar fp32 irradiance.
//
// Every direction of a probe starts at the probe centre (t = 0), so the first
// sample is shared.  Distances are whole pixels below 128, hence a first
// value of 0 is a hit for all 4 rays, a value >= 2 ends all 4 intervals, and
// only a value of exactly 1 takes a second, per-direction sample at t = 1,
// after which every ray has ended (1 + max(d, 1) >= 2).  These are the steps
// the generic interval march takes, so the codes are the same; the fluence
// is the same sum ((v0 + v1) + (v2 + v3)) / 4 of the per-direction values.
template <Storage S, bool UPPER>
__device__ __forceinline__ void level0_body(const LevelArgs &a, float *__restrict__ irr, float *__restrict__ dbg,
                                            int lb)
{
    using TT = Texel<S>;
    constexpr int TX = kL0TX, TY = kL0TY, NX = TX + 2, NY = TY + 2, UCX = TX / 2 + 2, UCY = TY / 2 + 2;
    static_assert(NX * NY <= kBlock, "one probe per thread");
    __shared__ float4 su[UPPER ? UCX * UCY * 4 : 1];
    __shared__ float4 fl[NX * NY];

    const int tiles_x = (a.PX + TX - 1) / TX;
    const int px0 = (lb % tiles_x) * TX, py0 = (lb / tiles_x) * TY;
    const int ulx = (px0 - 2) >> 1, uly = (py0 - 2) >> 1;
    // 32-bit indices (W, H <= 16384: three planes hold fewer than 2^32 floats)
    const uint32_t plane = __umul24((uint32_t)a.W, (uint32_t)a.H);
    const int slot = threadIdx.x;
    const int sx = slot % NX, sy = slot / NX;
    const int px = px0 - 1 + sx, py = py0 - 1 + sy;
    const bool probe = slot < NX * NY && px >= 0 && py >= 0 && px < a.PX && py < a.PY;
    const __amdgpu_buffer_rsrc_t sdf = sdf_rsrc(a.sdf, a.W, a.H);
    const int skip = sdf_skip(a.H);
    const float cx = ((float)px + 0.5f) * 2.0f, cy = ((float)py + 0.5f) * 2.0f;
    const int ix = 2 * px + 1, iy = 2 * py + 1;  // floor of the probe centre (exact)
    const bool in0 = probe && (uint32_t)ix < (uint32_t)a.W && (uint32_t)iy < (uint32_t)a.H;
    // the first distance sample is issued before the upper level's loads, so
    // the march does not wait for them; they land in LDS after the march
    const float d0 = in0 ? sdf_load(sdf, ix, iy, skip) : 0.0f;
    // the emission of the first sample's pixel (the usual hit) in the same wave
    // of loads, so a hit there needs no dependent load
    const int q0 = in0 ? pixel_index(ix, iy, a.W) : 0;
    float ep0 = 0.f, ep1 = 0.f, ep2 = 0.f;
    if (in0) {
        ep0 = a.emissive[q0];
        ep1 = a.emissive[plane + q0];
        ep2 = a.emissive[2 * plane + q0];
    }
    constexpr int UN = UPPER ? (UCX * UCY * 4 + kBlock - 1) / kBlock : 1;
    typename TT::T ureg[UN];
    if (UPPER) {
        const typename TT::T *up = (const typename TT::T *)a.upper;
#pragma unroll
        for (int n = 0; n < UN; ++n) {
            const int i = threadIdx.x + n * kBlock;
            if (i < UCX * UCY * 4) {
                const int j = i & 3, ux0 = (i >> 2) % UCX, uy0 = (i >> 2) / UCX;
                const int ux = min(max(ulx + ux0, 0), a.UX - 1), uy = min(max(uly + uy0, 0), a.UY - 1);
                ureg[n] = up[((__umul24((uint32_t)uy, (uint32_t)a.UX) + (uint32_t)ux) << 2) + (uint32_t)j];
            }
        }
    }

    int code[4] = {kRayT1, kRayT1, kRayT1, kRayT1};
    if (probe) {
        if (!in0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) code[k] = kRayExit;
        } else if (d0 < 0.5f) {
#pragma unroll
            for (int k = 0; k < 4; ++k) code[k] = q0;
        } else if (march_step(d0) < 2.0f) {
            const float t = march_step(d0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float2 dir = a.dirs[k];
                const int jx = floor_i32(fmaf(t, dir.x, cx)), jy = floor_i32(fmaf(t, dir.y, cy));
                if (!((uint32_t)jx < (uint32_t)a.W && (uint32_t)jy < (uint32_t)a.H)) {
                    code[k] = kRayExit;
                } else if (sdf_load(sdf, jx, jy, skip) < 0.5f) {
                    code[k] = pixel_index(jx, jy, a.W);
                }
            }
        }
    }
    if (a.dbg_codes && probe && sx >= 1 && sx <= TX && sy >= 1 && sy <= TY) {
#pragma unroll
        for (int k = 0; k < 4; ++k) a.dbg_codes[((size_t)py * a.PX + px) * 4 + k] = code[k];
    }
    float r[4], g[4], b[4], al[4];
    const bool one_hit = code[0] >= 0 && code[0] == code[1] && code[0] == code[2] && code[0] == code[3];
    if (one_hit) {
        // one hit pixel for all 4 directions: its emission once (already
        // loaded when it is the first sample's pixel)
        const bool first = code[0] == q0;
        const uint32_t c0 = (uint32_t)code[0];
        const float e0 = first ? ep0 : a.emissive[c0];
        const float e1 = first ? ep1 : a.emissive[plane + c0];
        const float e2 = first ? ep2 : a.emissive[2 * plane + c0];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            r[k] = e0;
            g[k] = e1;
            b[k] = e2;
            al[k] = 0.f;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            r[k] = g[k] = b[k] = 0.f;
            al[k] = code[k] >= 0 ? 0.f : 1.0f;
            if (code[k] >= 0) {
                const bool first = code[k] == q0;
                const uint32_t ck = (uint32_t)code[k];
                r[k] = first ? ep0 : a.emissive[ck];
                g[k] = first ? ep1 : a.emissive[plane + ck];
                b[k] = first ? ep2 : a.emissive[2 * plane + ck];
            }
        }
    }
    if (UPPER) {
#pragma unroll
        for (int n = 0; n < UN; ++n) {
            const int i = threadIdx.x + n * kBlock;
            // direction-major (su[k][y][x]): neighbouring probes read
            // neighbouring texels, so the merge's LDS reads do not conflict
            if (i < UCX * UCY * 4) su[(i & 3) * (UCX * UCY) + (i >> 2)] = TT::unpack(ureg[n]);
        }
        __syncthreads();
        if (probe && !one_hit) {
            const int usx = ((px - 1) >> 1) - ulx, usy = ((py - 1) >> 1) - uly;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (code[k] < 0) {
                    const float4 *sk = su + k * (UCX * UCY) + usy * UCX + usx;
                    const float4 u00 = sk[0], u10 = sk[1], u01 = sk[UCX], u11 = sk[UCX + 1];
                    const float3 v = bilerp4(u00, u10, u01, u11, px, py);
                    r[k] = v.x;
                    g[k] = v.y;
                    b[k] = v.z;
                }
            }
        }
    }
    if (probe) {
        const float4 f = make_float4(((r[0] + r[1]) + (r[2] + r[3])) * 0.25f, ((g[0] + g[1]) + (g[2] + g[3])) * 0.25f,
                                     ((b[0] + b[1]) + (b[2] + b[3])) * 0.25f,
                                     ((al[0] + al[1]) + (al[2] + al[3])) * 0.25f);
        fl[slot] = f;
        if (dbg && sx >= 1 && sx <= TX && sy >= 1 && sy <= TY) reinterpret_cast<float4 *>(dbg)[(size_t)py * a.PX + px] = f;
    }
    __syncthreads();

    // Pixels in 2x2 quads: pixel x = 2m interpolates probes m-1, m and x =
    // 2m+1 probes m, m+1 (indices clamped to the grid; the same in y), so a
    // quad reads 3x3 fluence texels for its 4 pixels, and with W even (and an
    // 8-byte aligned output) each row of it is one 8-byte store per plane.
    static_assert(TX * TY <= kBlock, "one quad per lane");
    if ((int)threadIdx.x < TX * TY) {
        const int m = px0 + (int)threadIdx.x % TX, n = py0 + (int)threadIdx.x / TX;
        const int x = 2 * m, y = 2 * n;
        if (x < a.W && y < a.H) {
            int cx[3], cy[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                cx[d] = min(max(m - 1 + d, 0), a.PX - 1) - (px0 - 1);
                cy[d] = min(max(n - 1 + d, 0), a.PY - 1) - (py0 - 1);
            }
            float4 f[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) f[r][c] = fl[cy[r] * NX + cx[c]];
            float3 v[2][2];
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx)
                    v[dy][dx] = bilerp4(f[dy][dx], f[dy][dx + 1], f[dy + 1][dx], f[dy + 1][dx + 1], x + dx, y + dy);
            const bool x1 = x + 1 < a.W;
            const bool pair = x1 && (a.W & 1) == 0 && (reinterpret_cast<uintptr_t>(irr) & 7) == 0;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
                if (y + dy >= a.H) break;
                const uint32_t p = __umul24((uint32_t)(y + dy), (uint32_t)a.W) + (uint32_t)x;
                if (pair) {
                    *reinterpret_cast<float2 *>(irr + p) = make_float2(v[dy][0].x, v[dy][1].x);
                    *reinterpret_cast<float2 *>(irr + plane + p) = make_float2(v[dy][0].y, v[dy][1].y);
                    *reinterpret_cast<float2 *>(irr + 2 * plane + p) = make_float2(v[dy][0].z, v[dy][1].z);
                } else {
                    irr[p] = v[dy][0].x;
                    irr[plane + p] = v[dy][0].y;
                    irr[2 * plane + p] = v[dy][0].z;
                    if (x1) {
                        irr[p + 1] = v[dy][1].x;
                        irr[plane + p + 1] = v[dy][1].y;
                        irr[2 * plane + p + 1] = v[dy][1].z;
                    }
                }
            }
        }
    }
}

template <Storage S, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_level0(LevelArgs a, float *__restrict__ irr, float *__restrict__ dbg)
{
    level0_body<S, UPPER>(a, irr, dbg, xcd_band(blockIdx.x, gridDim.x));
}

template <Storage S>
__global__ __launch_bounds__(kBlock) void k_level0_batch(const Level0Job *__restrict__ jobs,
                                                         const int *__restrict__ begins, int njobs)
{
    const int b = xcd_band(blockIdx.x, gridDim.x);
    const int j = batch_job(begins, njobs, b);
    const LevelArgs a = jobs[j].a;
    float *irr = jobs[j].irr;
    const int lb = b - begins[j];
    if (a.upper) level0_body<S, true>(a, irr, nullptr, lb);
    else level0_body<S, false>(a, irr, nullptr, lb);
}

// Canonical ray r = (py*PX + px)*N + k -> (px, py, k): the direction is the
// low log2(N) bits; probe / PX by an fp32 reciprocal, corrected to the exact
// quotient (probe < 2^24, so the estimate is off by at most one).
__device__ inline void decode_ray(const MarchLevel &L, int ray, int &px, int &py, int &k)
{
    k = ray & (L.N - 1);
    const int probe = ray >> L.log2n;
    py = (int)((float)probe * L.inv_px);
    px = probe - py * L.PX;
    if (px < 0) { --py; px += L.PX; }
    else if (px >= L.PX) { ++py; px -= L.PX; }
}

// Split-march slot -> (px, py, k).  Canonical order: the slot is the
// canonical index.  Wave tiles: slot = T*64 + lane, T = (tile row * ntx +
// tile column) * slices + slice, lane = (probe in the tile, direction in the
// slice) with the direction fastest; false for the lanes of a ragged edge
// tile that fall outside the probe grid.  Rays with nearby start points and
// directions share a wave, so their march lengths are alike.
__device__ inline bool decode_slot(const MarchLevel &L, int slot, int &px, int &py, int &k)
{
    if (!L.map) {
        decode_ray(L, slot, px, py, k);
        return true;
    }
    const int lane = slot & 63, T = slot >> 6;
    const int ns_log = L.log2n - L.kd_log;
    const int slice = T & ((1 << ns_log) - 1);
    const int tt = T >> ns_log;
    int ty = (int)((float)tt * L.inv_ntx);
    int tx = tt - ty * L.ntx;
    if (tx < 0) { --ty; tx += L.ntx; }
    else if (tx >= L.ntx) { ++ty; tx -= L.ntx; }
    const int pl = lane >> L.kd_log;
    k = (slice << L.kd_log) | (lane & ((1 << L.kd_log) - 1));
    px = (tx << L.tx_log) | (pl & ((1 << L.tx_log) - 1));
    py = (ty << L.ty_log) | (pl >> L.tx_log);
    return px < L.PX && py < L.PY;
}

// All levels >= 1 marched by one launch; each ray's code (hit pixel, -1 left
// the image, -2 reached t1) is written in canonical order for the merge
// kernels.  The block list is level-major with the top level first, so the
// levels whose rays are longest start earliest, and the long dependent chains
// of different levels run concurrently instead of one kernel after another.
//
// A workgroup owns 256*M consecutive canonical rays ((py*Px+px)*N+k, so a
// wave is 64 consecutive directions of one probe, or 64/N whole probes) and
// lane l marches rays base+l, base+l+256, ... one after another without
// waiting for the other lanes: when its ray ends it starts the next one in the
// same loop iteration (march lengths differ by two orders of magnitude, so
// one-ray-per-lane waves idle most lanes while their longest ray runs).
// Blocks go round-robin to the XCDs, which interleaves heavy and light tiles.
// A_top texel of a lane quad, merged in the march: the level kernel's
// arithmetic for a level without an upper level (quad sums of the 4 rays,
// times 1/4), stored by the quad's first lane.
__device__ inline void merge_top_quad(const MarchLevel &L, const float4 *__restrict__ emis4, bool valid, int code,
                                      uint32_t at)
{
    float r = 0.f, g = 0.f, b = 0.f, al = 0.f;
    if (valid) {
        if (code >= 0) {
            const float4 e = emis4[code];
            r = e.x;
            g = e.y;
            b = e.z;
        } else {
            al = 1.0f;
        }
    }
    r = quad_sum(r) * 0.25f;
    g = quad_sum(g) * 0.25f;
    b = quad_sum(b) * 0.25f;
    al = quad_sum(al) * 0.25f;
    if (valid && (at & 3) == 0) {
        const float4 v = make_float4(r, g, b, al);
        const uint32_t i = at >> 2;
        if (L.storage == (int)Storage::F32) ((float4 *)L.merged)[i] = v;
        else if (L.storage == (int)Storage::F16) ((uint2 *)L.merged)[i] = Texel<Storage::F16>::pack(v);
        else ((uint2 *)L.merged)[i] = Texel<Storage::BF16>::pack(v);
    }
}

template <int SPEC, int M>
__device__ __forceinline__ void march_all_body(const uint8_t *sdf_codes, int W, int H, const MarchLevel &L, int lb,
                                               const float4 *__restrict__ emis4)
{
    const __amdgpu_buffer_rsrc_t sdf = sdf_rsrc(sdf_codes, W, H);
    const int base = lb * kMarchBlock * M + (int)threadIdx.x;
    const int total = L.map ? (int)L.slots : L.PX * L.PY * L.N;

    int j = 0, ray = base;
    bool valid = false;      // the slot holds a ray of the grid
    uint32_t at = 0;         // its canonical index (where its code goes)
    Ray r;
    auto start = [&]() {
        int px, py, k;
        valid = decode_slot(L, ray < total ? ray : 0, px, py, k) && ray < total;
        at = ((__umul24((uint32_t)py, (uint32_t)L.PX) + (uint32_t)px) << L.log2n) + (uint32_t)k;
        const float2 dir = L.dirs[k];
        r = ray_begin(((float)px + 0.5f) * L.spacing, ((float)py + 0.5f) * L.spacing, dir.x, dir.y, L.t0, L.t1,
                      valid);
    };
    if (M == 1 && SPEC > 1) {
        // One ray per lane, SPEC samples per round: the samples of
        // ray_round<SPEC> (p[0] = t, p[j] = p[j-1] + guess; sample j accepted
        // while its step equals the guess), with the lanes' paths as masked
        // updates and the ray's end read from where t stopped, as below: t is
        // left at the sample that ended the ray (off-image or a hit), or at or
        // past t1 when the interval ended.
        int px, py, k;
        valid = decode_slot(L, ray < total ? ray : 0, px, py, k) && ray < total;
        at = ((__umul24((uint32_t)py, (uint32_t)L.PX) + (uint32_t)px) << L.log2n) + (uint32_t)k;
        const float2 dir = L.dirs[k];
        const float cx = ((float)px + 0.5f) * L.spacing, cy = ((float)py + 0.5f) * L.spacing;
        const int skip = sdf_skip(H);
        const float t1 = L.t1;
        float t = L.t0, guess = 1.0f;
        bool alive = valid && t < t1;
        while (alive) {
            float p[SPEC];
            uint32_t c[SPEC];
            bool ins[SPEC];
            p[0] = t;
#pragma unroll
            for (int j = 1; j < SPEC; ++j) p[j] = p[j - 1] + guess;
#pragma unroll
            for (int j = 0; j < SPEC; ++j) {
                const int ix = floor_i32(fmaf(p[j], dir.x, cx)), iy = floor_i32(fmaf(p[j], dir.y, cy));
                ins[j] = (uint32_t)ix < (uint32_t)W && (uint32_t)iy < (uint32_t)H;
                c[j] = sdf_code_load(sdf, ix, iy, skip);
            }
            bool go = true;
#pragma unroll
            for (int j = 0; j < SPEC; ++j) {
                if (!go) continue;
                if ((j > 0 && !(p[j] < t1)) || !ins[j] || c[j] == 0) {
                    // the interval ended at the previous step, or this sample
                    // left the image or hit: the ray ends at p[j]
                    t = p[j];
                    alive = false;
                    go = false;
                    continue;
                }
                const float step = step_of_code(c[j]);
                if (j + 1 < SPEC && step == guess) continue;  // p[j+1] == p[j] + step
                t = p[j] + step;
                guess = step;
                alive = t < t1;
                go = false;
            }
        }
        int code = kRayT1;
        if (t < t1) {
            const int ix = floor_i32(fmaf(t, dir.x, cx)), iy = floor_i32(fmaf(t, dir.y, cy));
            code = ((uint32_t)ix < (uint32_t)W && (uint32_t)iy < (uint32_t)H) ? pixel_index(ix, iy, W) : kRayExit;
        }
        if (valid) L.codes[at] = code;
        if (L.merged) merge_top_quad(L, emis4, valid, code, at);
        return;
    }
    if (M == 1 && SPEC == 1) {
        // One ray per lane, one sample per step: a branch-free body in a
        // divergent loop (lanes leave as their rays end) on the raw byte
        // code.  The samples are those of ray_round<1> -- the same positions,
        // the same step -- and the ray's end is read from where t stopped: t
        // advances only on an in-image sample, by 0 on a hit (code 0), so a
        // ray that ends with t < t1 ended at the sample at t, which is
        // off-image (exit) or a hit; otherwise it reached t1.
        int px, py, k;
        valid = decode_slot(L, ray < total ? ray : 0, px, py, k) && ray < total;
        at = ((__umul24((uint32_t)py, (uint32_t)L.PX) + (uint32_t)px) << L.log2n) + (uint32_t)k;
        const float2 dir = L.dirs[k];
        const float cx = ((float)px + 0.5f) * L.spacing, cy = ((float)py + 0.5f) * L.spacing;
        const int skip = sdf_skip(H);
        const float t1 = L.t1;
        float t = L.t0;
        bool alive = valid && t < t1;
        while (alive) {
            const int ix = floor_i32(fmaf(t, dir.x, cx)), iy = floor_i32(fmaf(t, dir.y, cy));
            const uint32_t c = sdf_code_load(sdf, ix, iy, skip);
            const bool ins = (uint32_t)ix < (uint32_t)W && (uint32_t)iy < (uint32_t)H;
            const float tn = t + step_of_code(c);
            t = ins ? tn : t;
            alive = ins && c != 0 && tn < t1;
        }
        int code = kRayT1;
        if (t < t1) {
            const int ix = floor_i32(fmaf(t, dir.x, cx)), iy = floor_i32(fmaf(t, dir.y, cy));
            code = ((uint32_t)ix < (uint32_t)W && (uint32_t)iy < (uint32_t)H) ? pixel_index(ix, iy, W) : kRayExit;
        }
        if (valid) L.codes[at] = code;
        if (L.merged) merge_top_quad(L, emis4, valid, code, at);
        return;
    }
    start();
    for (;;) {
        if (r.alive) ray_round<SPEC>(r, sdf, W, H, L.t1);
        if (j < M && !r.alive) {
            if (valid) L.codes[at] = r.code;
            ray += kMarchBlock;
            if (++j < M) start();
        }
        if (!__any(j < M)) break;
    }
}

template <int SPEC, int M>
__global__ __launch_bounds__(kMarchBlock) void k_march_all(MarchAllArgs a)
{
    // the grid is (blocks of the largest level) x (levels): the level is
    // blockIdx.y, and workgroups are dispatched x-fastest, so the top level's
    // (longest) rays start first
    const MarchLevel &L = a.lev[blockIdx.y];
    if ((int)blockIdx.x >= L.blocks) return;
    march_all_body<SPEC, M>(a.sdf, a.W, a.H, L, (int)blockIdx.x, a.emis4);
}

// The split march over a work list of (view, level) jobs, blocks in launch
// order as k_march_all's (the job list puts every view's top level first).
template <int SPEC, int M>
__global__ __launch_bounds__(kMarchBlock) void k_march_batch(const MarchJob *__restrict__ jobs,
                                                        const int *__restrict__ begins, int njobs)
{
    const int b = (int)blockIdx.x;
    const int j = batch_job(begins, njobs, b);
    const MarchJob J = jobs[j];
    march_all_body<SPEC, M>(J.sdf, J.W, J.H, J.L, b - begins[j], J.emis4);
}

// Persistent form of the split march.  March lengths differ by two orders of
// magnitude (most rays take 1-4 samples, rays grazing a surface up to a few
// hundred), so with one ray per lane a wave idles most lanes while its
// longest ray runs.  Here a grid of resident waves walks every ray of every
// level: the rays are cut into packets of 64 consecutive codes, numbered
// from the top level down (the longest rays first), and wave w owns packets
// w, w + n_waves, w + 2 n_waves, ... (an even mix of levels and image
// regions per wave).  Each lane holds its current ray and the next one,
// prepared ahead (its direction load issued an iteration or more before the
// ray starts, so starting it costs no dependent load): every iteration a lane
// whose ray ended takes its prepared one, the lanes without a prepared ray
// take the next entries of the wave's packet sequence (a ballot and a lane
// rank; the position in the sequence is a scalar, no memory atomics), and
// every busy lane takes one march round.  Each ray's samples, and so every
// code, are those of k_march_all; only the order of the rays differs.
template <int SPEC>
__global__ __launch_bounds__(kBlock) void k_march_persist(MarchAllArgs a)
{
    // per-level constants, indexed per lane (a lane's ray may be of any level)
    __shared__ int tab_i[RC_MAX_LEVELS][4];     // code_begin, PX, log2n, dir_off
    __shared__ float tab_f[RC_MAX_LEVELS][4];   // inv_px, spacing, t0, t1
    if ((int)threadIdx.x < a.nlev) {
        const MarchLevel &L = a.lev[threadIdx.x];
        tab_i[threadIdx.x][0] = (int)L.code_begin;
        tab_i[threadIdx.x][1] = L.PX;
        tab_i[threadIdx.x][2] = L.log2n;
        tab_i[threadIdx.x][3] = L.dir_off;
        tab_f[threadIdx.x][0] = L.inv_px;
        tab_f[threadIdx.x][1] = L.spacing;
        tab_f[threadIdx.x][2] = L.t0;
        tab_f[threadIdx.x][3] = L.t1;
    }
    __syncthreads();
    const __amdgpu_buffer_rsrc_t sdf = sdf_rsrc(a.sdf, a.W, a.H);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    const uint32_t np = a.n_packets, nw = a.n_waves;
    // this wave's sequence length (waves past n_waves in the last workgroup have none)
    const uint32_t vend = (wave < nw && wave < np) ? ((np - 1 - wave) / nw + 1) * 64 : 0;
    uint32_t vnext = 0;  // wave-uniform
    Ray r = ray_begin(0.f, 0.f, 0.f, 0.f, 0.f, 0.f, false);
    float t1 = 0.f;
    uint32_t out = 0;
    bool have = false;
    // the prepared next ray
    float ncx = 0.f, ncy = 0.f, nt0 = 0.f, nt1 = 0.f;
    float2 ndir = make_float2(0.f, 0.f);
    uint32_t nout = 0;
    bool have_n = false;
    for (;;) {
        if (!have && have_n) {
            r = ray_begin(ncx, ncy, ndir.x, ndir.y, nt0, nt1, true);
            t1 = nt1;
            out = nout;
            have = true;
            have_n = false;
        }
        const uint64_t want = __ballot(!have_n);
        if (want != 0 && vnext < vend) {
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0u));
            const uint32_t v = vnext + rank;
            if (!have_n && v < vend) {
                const uint32_t pc = np - 1 - (wave + (v >> 6) * nw);  // packet in code order
                const uint32_t ray = pc * 64 + (v & 63);
                if (ray < a.total_rays) {
                    int lv = 0;  // levels are top first; code_begin falls with the level
#pragma unroll
                    for (int j = 0; j < RC_MAX_LEVELS; ++j)
                        if (j < a.nlev && ray < a.lev[j].code_begin) lv = j + 1;
                    const int rr = (int)(ray - (uint32_t)tab_i[lv][0]);
                    const int PX = tab_i[lv][1], log2n = tab_i[lv][2];
                    const int k = rr & ((1 << log2n) - 1), probe = rr >> log2n;
                    int py = (int)((float)probe * tab_f[lv][0]);
                    int px = probe - py * PX;
                    if (px < 0) { --py; px += PX; }
                    else if (px >= PX) { ++py; px -= PX; }
                    ndir = a.dirs_base[tab_i[lv][3] + k];  // used when the ray starts
                    const float sp = tab_f[lv][1];
                    ncx = ((float)px + 0.5f) * sp;
                    ncy = ((float)py + 0.5f) * sp;
                    nt0 = tab_f[lv][2];
                    nt1 = tab_f[lv][3];
                    nout = ray;
                    have_n = true;
                }
            }
            vnext += (uint32_t)__builtin_popcountll(want);
        }
        if (!__any(have || have_n)) {
            if (vnext >= vend) break;
            continue;
        }
        if (have) {
            ray_round<SPEC>(r, sdf, a.W, a.H, t1);
            if (!r.alive) {
                a.codes_base[out] = r.code;
                have = false;
            }
        }
    }
}

// The split march's decode of every slot of one level, recomposed into the
// canonical index it stands for (parity readback: equal to the slot number
// when the decode is right).
// For every slot of a level's split march, its ray's canonical index gets
// the slot number (the slot that marches it).
__global__ __launch_bounds__(kBlock) void k_ray_decode(MarchLevel l, int64_t *__restrict__ index)
{
    const int slot = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (slot >= (l.map ? (int)l.slots : l.PX * l.PY * l.N)) return;
    int px, py, k;
    if (decode_slot(l, slot, px, py, k)) index[((int64_t)py * l.PX + px) * l.N + k] = slot;
}

template <Storage S>
__global__ __launch_bounds__(kBlock) void k_to_f32(const void *__restrict__ in, size_t texels, float4 *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < texels) out[i] = Texel<S>::unpack(((const typename Texel<S>::T *)in)[i]);
}

int tile_kd(int shape)
{
    return shape == TILE_8x8x4 ? 4 : shape == TILE_4x4x16 ? 16 : shape == TILE_2x2x64 ? 64 : shape == TILE_8x4x8 ? 8 : 256;
}
int tile_tx(int shape)
{
    return shape == TILE_8x8x4 || shape == TILE_8x4x8 ? 8 : shape == TILE_4x4x16 ? 4 : shape == TILE_2x2x64 ? 2 : 1;
}
int tile_ty(int shape) { return shape == TILE_8x4x8 ? 4 : tile_tx(shape); }

template <int TX, int TY, int KD, int RPT, Storage S>
static void launch_level_shape(const LevelArgs &a, hipStream_t s)
{
    const int blocks = ((a.PX + TX - 1) / TX) * ((a.PY + TY * RPT - 1) / (TY * RPT)) * (a.N / KD);
    if (a.codes) {
        if (a.upper) hipLaunchKernelGGL((k_level<TX, TY, KD, RPT, S, true, true>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_level<TX, TY, KD, RPT, S, false, true>), dim3(blocks), dim3(kBlock), 0, s, a);
    } else {
        if (a.upper) hipLaunchKernelGGL((k_level<TX, TY, KD, RPT, S, true, false>), dim3(blocks), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_level<TX, TY, KD, RPT, S, false, false>), dim3(blocks), dim3(kBlock), 0, s, a);
    }
}

template <int RPT, Storage S>
static hipError_t launch_level_rpt(const LevelArgs &a, int shape, hipStream_t s)
{
    switch (shape) {
    case TILE_8x8x4: launch_level_shape<8, 8, 4, RPT, S>(a, s); break;
    case TILE_4x4x16: launch_level_shape<4, 4, 16, RPT, S>(a, s); break;
    case TILE_2x2x64: launch_level_shape<2, 2, 64, RPT, S>(a, s); break;
    case TILE_1x1x256: launch_level_shape<1, 1, 256, RPT, S>(a, s); break;
    case TILE_8x4x8: launch_level_shape<8, 4, 8, RPT, S>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <Storage S>
static hipError_t launch_level_st(const LevelArgs &a, int shape, int rpt, hipStream_t s)
{
    if (tile_kd(shape) > a.N) return hipErrorInvalidValue;
    switch (rpt) {
    case 1: return launch_level_rpt<1, S>(a, shape, s);
    case 2: return launch_level_rpt<2, S>(a, shape, s);
    case 4: return launch_level_rpt<4, S>(a, shape, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_level(const LevelArgs &a, Storage st, int shape, int rpt, hipStream_t s)
{
    switch (st) {
    case Storage::F32: return launch_level_st<Storage::F32>(a, shape, rpt, s);
    case Storage::F16: return launch_level_st<Storage::F16>(a, shape, rpt, s);
    case Storage::BF16: return launch_level_st<Storage::BF16>(a, shape, rpt, s);
    }
    return hipErrorInvalidValue;
}

template <Storage S>
static hipError_t launch_level0_st(const LevelArgs &a, float *irr, float *dbg, hipStream_t s)
{
    const int blocks = level0_blocks(a.PX, a.PY);
    if (a.upper) hipLaunchKernelGGL((k_level0<S, true>), dim3(blocks), dim3(kBlock), 0, s, a, irr, dbg);
    else hipLaunchKernelGGL((k_level0<S, false>), dim3(blocks), dim3(kBlock), 0, s, a, irr, dbg);
    return hipGetLastError();
}

hipError_t launch_level0_gather(const LevelArgs &a, Storage st, float *irr, float *dbg, hipStream_t s)
{
    switch (st) {
    case Storage::F32: return launch_level0_st<Storage::F32>(a, irr, dbg, s);
    case Storage::F16: return launch_level0_st<Storage::F16>(a, irr, dbg, s);
    case Storage::BF16: return launch_level0_st<Storage::BF16>(a, irr, dbg, s);
    }
    return hipErrorInvalidValue;
}

template <int SPEC>
static hipError_t launch_march_all_spec(const MarchAllArgs &a, int m, int blocks, hipStream_t s)
{
    const dim3 grid(blocks, a.nlev);
    switch (m) {
    case 1: hipLaunchKernelGGL((k_march_all<SPEC, 1>), grid, dim3(kMarchBlock), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_march_all<SPEC, 2>), grid, dim3(kMarchBlock), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_march_all<SPEC, 4>), grid, dim3(kMarchBlock), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_march_all<SPEC, 8>), grid, dim3(kMarchBlock), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_march_persist(const MarchAllArgs &a, int spec, hipStream_t s)
{
    if (a.nlev < 1 || a.total_rays == 0 || a.n_waves == 0) return hipSuccess;
    const dim3 grid((a.n_waves + 3) / 4);
    switch (spec) {
    case 1: hipLaunchKernelGGL(k_march_persist<1>, grid, dim3(kBlock), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_march_persist<2>, grid, dim3(kBlock), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_march_persist<4>, grid, dim3(kBlock), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_march_all(const MarchAllArgs &a, int spec, int m, hipStream_t s)
{
    if (a.nlev < 1) return hipSuccess;  // one-level frame: no level >= 1 to march
    int blocks = 0;
    for (int i = 0; i < a.nlev; ++i) blocks = a.lev[i].blocks > blocks ? a.lev[i].blocks : blocks;
    if (blocks <= 0) return hipSuccess;
    switch (spec) {
    case 1: return launch_march_all_spec<1>(a, m, blocks, s);
    case 2: return launch_march_all_spec<2>(a, m, blocks, s);
    case 4: return launch_march_all_spec<4>(a, m, blocks, s);
    }
    return hipErrorInvalidValue;
}



int level_blocks(int PX, int PY, int N, int shape, int rpt)
{
    const int tx = tile_tx(shape), ty = tile_ty(shape) * rpt, kd = tile_kd(shape);
    return ((PX + tx - 1) / tx) * ((PY + ty - 1) / ty) * (N / kd);
}

int level0_blocks(int PX, int PY) { return ((PX + kL0TX - 1) / kL0TX) * ((PY + kL0TY - 1) / kL0TY); }

template <int SPEC>
static hipError_t launch_march_batch_spec(const BatchLaunch &b, int m, hipStream_t s)
{
    const MarchJob *j = (const MarchJob *)b.jobs;
    switch (m) {
    case 1: hipLaunchKernelGGL((k_march_batch<SPEC, 1>), dim3(b.blocks), dim3(kMarchBlock), 0, s, j, b.begins, b.njobs); break;
    case 2: hipLaunchKernelGGL((k_march_batch<SPEC, 2>), dim3(b.blocks), dim3(kMarchBlock), 0, s, j, b.begins, b.njobs); break;
    case 4: hipLaunchKernelGGL((k_march_batch<SPEC, 4>), dim3(b.blocks), dim3(kMarchBlock), 0, s, j, b.begins, b.njobs); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_march_batch(const BatchLaunch &b, int spec, int m, hipStream_t s)
{
    if (b.blocks <= 0) return hipSuccess;
    switch (spec) {
    case 1: return launch_march_batch_spec<1>(b, m, s);
    case 2: return launch_march_batch_spec<2>(b, m, s);
    }
    return hipErrorInvalidValue;
}

template <int TX, int TY, int KD, int RPT, Storage S>
static void launch_level_batch_shape(const BatchLaunch &b, bool from_codes, hipStream_t s)
{
    const LevelArgs *j = (const LevelArgs *)b.jobs;
    if (from_codes)
        hipLaunchKernelGGL((k_level_batch<TX, TY, KD, RPT, S, true>), dim3(b.blocks), dim3(kBlock), 0, s, j, b.begins,
                           b.njobs);
    else
        hipLaunchKernelGGL((k_level_batch<TX, TY, KD, RPT, S, false>), dim3(b.blocks), dim3(kBlock), 0, s, j, b.begins,
                           b.njobs);
}

template <int RPT, Storage S>
static hipError_t launch_level_batch_rpt(const BatchLaunch &b, int shape, bool from_codes, hipStream_t s)
{
    switch (shape) {
    case TILE_8x8x4: launch_level_batch_shape<8, 8, 4, RPT, S>(b, from_codes, s); break;
    case TILE_4x4x16: launch_level_batch_shape<4, 4, 16, RPT, S>(b, from_codes, s); break;
    case TILE_2x2x64: launch_level_batch_shape<2, 2, 64, RPT, S>(b, from_codes, s); break;
    case TILE_1x1x256: launch_level_batch_shape<1, 1, 256, RPT, S>(b, from_codes, s); break;
    case TILE_8x4x8: launch_level_batch_shape<8, 4, 8, RPT, S>(b, from_codes, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <Storage S>
static hipError_t launch_level_batch_st(const BatchLaunch &b, int shape, int rpt, bool from_codes, hipStream_t s)
{
    switch (rpt) {
    case 1: return launch_level_batch_rpt<1, S>(b, shape, from_codes, s);
    case 2: return launch_level_batch_rpt<2, S>(b, shape, from_codes, s);
    case 4: return launch_level_batch_rpt<4, S>(b, shape, from_codes, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_level_batch(const BatchLaunch &b, Storage st, int shape, int rpt, bool from_codes, hipStream_t s)
{
    if (b.blocks <= 0) return hipSuccess;
    switch (st) {
    case Storage::F32: return launch_level_batch_st<Storage::F32>(b, shape, rpt, from_codes, s);
    case Storage::F16: return launch_level_batch_st<Storage::F16>(b, shape, rpt, from_codes, s);
    case Storage::BF16: return launch_level_batch_st<Storage::BF16>(b, shape, rpt, from_codes, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_level0_batch(const BatchLaunch &b, Storage st, hipStream_t s)
{
    if (b.blocks <= 0) return hipSuccess;
    const Level0Job *j = (const Level0Job *)b.jobs;
    switch (st) {
    case Storage::F32: hipLaunchKernelGGL(k_level0_batch<Storage::F32>, dim3(b.blocks), dim3(kBlock), 0, s, j, b.begins, b.njobs); break;
    case Storage::F16: hipLaunchKernelGGL(k_level0_batch<Storage::F16>, dim3(b.blocks), dim3(kBlock), 0, s, j, b.begins, b.njobs); break;
    case Storage::BF16: hipLaunchKernelGGL(k_level0_batch<Storage::BF16>, dim3(b.blocks), dim3(kBlock), 0, s, j, b.begins, b.njobs); break;
    }
    return hipGetLastError();
}

hipError_t launch_level_to_f32(const void *level, Storage st, size_t texels, float *out, hipStream_t s)
{
    const unsigned blocks = (unsigned)((texels + kBlock - 1) / kBlock);
    if (blocks == 0) return hipSuccess;
    switch (st) {
    case Storage::F32: hipLaunchKernelGGL(k_to_f32<Storage::F32>, dim3(blocks), dim3(kBlock), 0, s, level, texels, (float4 *)out); break;
    case Storage::F16: hipLaunchKernelGGL(k_to_f32<Storage::F16>, dim3(blocks), dim3(kBlock), 0, s, level, texels, (float4 *)out); break;
    case Storage::BF16: hipLaunchKernelGGL(k_to_f32<Storage::BF16>, dim3(blocks), dim3(kBlock), 0, s, level, texels, (float4 *)out); break;
    }
    return hipGetLastError();
}

bool march_map_legal(int level, int map)
{
    static const int kd_log[kMarchMaps] = {0, 4, 2, 3, 5, 0};
    return level >= 1 && map >= 0 && map < kMarchMaps && (map == 0 || (1 << kd_log[map]) <= num_dirs(level));
}

void march_level_map(MarchLevel &l, int map)
{
    // (tx_log, ty_log, kd_log) of each wave tile
    static const int tile[kMarchMaps][3] = {{0, 0, 0}, {1, 1, 4}, {2, 2, 2}, {2, 1, 3}, {1, 0, 5}, {3, 3, 0}};
    l.map = map;
    l.tx_log = tile[map][0];
    l.ty_log = tile[map][1];
    l.kd_log = tile[map][2];
    if (map == 0) {
        l.ntx = 0;
        l.inv_ntx = 0.0f;
        l.slots = (uint32_t)((int64_t)l.PX * l.PY * l.N);
        return;
    }
    l.ntx = (l.PX + (1 << l.tx_log) - 1) >> l.tx_log;
    const int nty = (l.PY + (1 << l.ty_log) - 1) >> l.ty_log;
    l.inv_ntx = 1.0f / (float)l.ntx;
    l.slots = (uint32_t)((int64_t)l.ntx * nty * (l.N >> l.kd_log) * 64);
}

hipError_t launch_ray_decode(const MarchLevel &l, int64_t *index, hipStream_t s)
{
    const int64_t total = l.map ? (int64_t)l.slots : (int64_t)l.PX * l.PY * l.N;
    const unsigned blocks = (unsigned)((total + kBlock - 1) / kBlock);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ray_decode, dim3(blocks), dim3(kBlock), 0, s, l, index);
    return hipGetLastError();
}

}  // namespace rc
