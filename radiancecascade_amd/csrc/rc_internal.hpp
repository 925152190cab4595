// rc_internal.hpp -- declarations shared by the library's translation units.
#pragma once

#include <string>

#include "../../include/rc_gi.h"
#include "rc_common.hpp"

namespace rc {

// ---- scene (rc_scene.hip) ---------------------------------------------------
void scene_generate(uint64_t seed, int n_boxes, int n_emissive, rc_box *boxes);
hipError_t launch_raster(const double *boxes_dev, int n, int W, int H, float *depth, float *albedo, float *emissive,
                         hipStream_t stream);

// ---- depth convention ---------------------------------------------------------
// Linear view depth as the occupancy rule reads it: the plane itself, or the
// reference renderer's NDC depth d converted by D = nf / (f - d * fmn) with
// nf = n*f and fmn = f - n rounded to fp32 on the host (each operation rounds
// once: the library is built without contraction).
struct DepthConv {
    int ndc;
    float nf, f, fmn;
};
__device__ inline float linear_depth(const DepthConv &c, float d) { return c.ndc ? c.nf / (c.f - d * c.fmn) : d; }

// Occupancy of pixels (x, y0 + j), j < 4, x < W: any emissive channel > 0, or
// a 4-neighbour q with D(q) - D(p) > 0.05 D(p) (the near side of a depth
// silhouette), in fp32.  All loads are unconditional at clamped
// addresses, so they are in flight together; the depth above and below the
// middle rows are the lane's own loads (two extra rows).  occ[j] is false for
// rows at or beyond H; e[j] is the pixel's emission (r, g, b, 0).
__device__ __forceinline__ void occupancy4(const float *__restrict__ depth, const DepthConv &dc,
                                           const float *__restrict__ emissive, int W, int H, int x, int y0,
                                           bool occ[4], float4 e[4])
{
    const size_t plane = (size_t)W * H;
    float du0 = depth[(size_t)max(y0 - 1, 0) * W + x];
    float dd3 = depth[(size_t)min(y0 + 4, H - 1) * W + x];
    float dp[4], dl[4], dr[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const size_t p = (size_t)min(y0 + j, H - 1) * W + x;
        e[j] = make_float4(emissive[p], emissive[plane + p], emissive[2 * plane + p], 0.0f);
        dp[j] = depth[p];
        dl[j] = depth[p - (x > 0)];
        dr[j] = depth[p + (x + 1 < W)];
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
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = y0 + j;
        // rows y - 1 and y + 1: this lane's neighbouring rows or the two extra
        // loads (where y + 1 < H, row y + 1 is not clamped)
        const float du = j == 0 ? du0 : dp[j - 1], dd = j == 3 ? dd3 : dp[j + 1];
        const float thr = 0.05f * dp[j];
        occ[j] = y < H && (e[j].x > 0.0f || e[j].y > 0.0f || e[j].z > 0.0f || (x > 0 && dl[j] - dp[j] > thr) ||
                           (x + 1 < W && dr[j] - dp[j] > thr) || (y > 0 && du - dp[j] > thr) ||
                           (y + 1 < H && dd - dp[j] > thr));
    }
}

// ---- occupancy + jump flood (rc_jfa.hip) --------------------------------------
int jfa_num_passes(int W, int H);
// largest step fused into the LDS tail for a requested fuse limit (0 = none)
int jfa_tail_step(int W, int H, int fuse);
hipError_t launch_sdf_untile(const uint8_t *sdf, float *out, int W, int H, hipStream_t s);
// `gate` (nullptr: always run): the kernel does nothing when *gate == 0 (the
// learned frame's tile fallback runs them only when some tile is flagged).
hipError_t launch_seed_init(const float *depth, const DepthConv &dc, const float *emissive, uint32_t *seed,
                            float4 *emis4, int W, int H, hipStream_t s, const int *gate = nullptr);
// rows: image rows per wave (1, 2 or 4; the result does not depend on it)
hipError_t launch_jfa_pass(const uint32_t *in, uint32_t *out, uint8_t *sdf, int W, int H, int step, bool final_pass,
                           hipStream_t s, const int *gate = nullptr, int rows = 1);
hipError_t launch_jfa_tail(const uint32_t *in, uint32_t *seeds_out, uint8_t *sdf, int W, int H, int tail_step,
                           hipStream_t s, const int *gate = nullptr);

// ---- exact distance transform (rc_edt.hip) ------------------------------------
// Buffers: bits = column words [ceil(H/64)][W] (u64), ends = the nearest
// occupied rows above / below each 64-row band [2][ceil(H/64)][W] (u16).
// stage: 1 = the occupancy kernel (bits + emis4), 2 = the band-end and row
// kernels (ends, then the SDF), 3 = all three.
size_t edt_bits_bytes(int W, int H);
size_t edt_ends_bytes(int W, int H);
hipError_t launch_edt(const float *depth, const DepthConv &dc, const float *emissive, uint64_t *bits, uint16_t *ends,
                      uint8_t *sdf, float4 *emis4, int W, int H, hipStream_t s, const int *gate = nullptr,
                      int stage = 3);

// ---- work lists of views (one launch per stage over many views) --------------
// A batched launch runs one kernel over the jobs of many views.  Job j owns
// the logical blocks [begins[j], begins[j+1]) (begins ascending from 0, no
// empty job); a workgroup takes its logical block b from the XCD banding of
// the whole grid, finds its job, and runs the single-view kernel body on
// block b - begins[j] of that job, so every view's result is the single-view
// one bit for bit.
struct BatchLaunch {
    const void *jobs;        // device array of the launch's job type
    const int *begins;       // device [njobs]
    int njobs, blocks;
};

// Job index of logical block b: a 64-ary search over begins, one wave-wide
// load and ballot per level (one level up to 64 jobs, two up to 4096).  Every
// lane of the wave must be active; the result is wave-uniform.
__device__ inline int batch_job(const int *__restrict__ begins, int njobs, int b)
{
    const int lane = (int)(threadIdx.x & 63);
    int lo = 0, stride = 1;
    while (stride < njobs / 64 + (njobs % 64 != 0)) stride *= 64;
    for (; stride >= 1; stride >>= 6) {
        const int idx = lo + lane * stride;
        const bool le = idx < njobs && begins[idx] <= b;
        lo += (__popcll(__ballot(le)) - 1) * stride;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

struct OccJob {
    const float *depth, *emissive;
    uint64_t *bits;
    float4 *emis4;
    DepthConv dc;
    int W, H, bx_count;
};
struct EndsJob {
    const uint64_t *bits;
    uint16_t *ends;
    int W, H;
};
struct RowJob {
    const uint64_t *bits;
    const uint16_t *ends;
    uint8_t *sdf;
    int W, H;
};
int occ_blocks(int W, int H);   // 1024-lane workgroups of the occupancy kernel
int occ_bx(int W);
int ends_blocks(int W);         // 256-lane workgroups of the band-end kernel
int edt_row_wmax(int W);        // LDS row capacity of the row kernel (1024 .. 16384)
hipError_t launch_occ_batch(const BatchLaunch &b, hipStream_t s);
hipError_t launch_ends_batch(const BatchLaunch &b, hipStream_t s);
hipError_t launch_rows_batch(const BatchLaunch &b, int wmax, hipStream_t s);  // one workgroup per row

// ---- cascades (rc_cascade.hip) ----------------------------------------------
// Workgroup tile of a level kernel: TX x TY probes x KD directions, TX*TY*KD = 256.
// 8x4x8 is the one non-square tile: per ray, fewer upper texels staged than
// 4x4x16 (a wider probe block shares its halo) with longer output runs than 8x8x4.
enum TileShape : int { TILE_8x8x4 = 0, TILE_4x4x16 = 1, TILE_2x2x64 = 2, TILE_1x1x256 = 3, TILE_8x4x8 = 4, TILE_COUNT = 5 };
// Default tile of level i of L: the top level has no upper neighbourhood to
// stage, so it takes whole 256-direction slices of one probe (no idle lanes
// at ragged grid edges); lower levels take 4x4 probes x 16 directions, whose
// staged upper block is exactly one texel per ray.
inline int default_level_tile(int level, int L)
{
    const int N = num_dirs(level);
    if (level == L - 1 && N >= 256) return TILE_1x1x256;
    if (N >= 16) return TILE_4x4x16;
    return TILE_8x8x4;
}
// First direction of level i in a table holding levels 0, 1, ... back to
// back: sum of N_j for j < i.
inline int dir_table_offset(int level) { return 4 * ((1 << (2 * level)) - 1) / 3; }
int tile_kd(int shape);
int tile_tx(int shape);
int tile_ty(int shape);

// Workgroups are handed to the 8 XCDs round-robin (block b to XCD b % 8);
// map b to a logical block so that each XCD gets one contiguous run of the
// nb logical blocks (a band of tiles sharing its L2).  A bijection on [0, nb).
__device__ inline int xcd_band(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return xcd * q + min(xcd, r) + idx;
}

struct LevelArgs {
    const uint8_t *sdf;      // distance codes, 16x8 tiles
    const float *emissive;   // [3][H][W]
    const float4 *emis4;     // the same, interleaved (r, g, b, 0) per pixel, or nullptr (read the planes)
    const float2 *dirs;      // [N] unit directions of this level
    const void *upper;       // merged level i+1 (storage texels), nullptr at the top level
    void *out;               // merged level i (storage texels)
    int W, H;
    int level, PX, PY, N;    // this level
    int UX, UY;              // probe grid of level i+1
    float t0, t1, spacing;
    const int *codes;        // pre-marched ray codes of this level, or nullptr (march in place)
    int *dbg_codes;          // in-place march: each ray's code also stored here (canonical order), or nullptr
    const int *gate;         // the kernel does nothing when *gate == 0 (nullptr: always runs)
    const uint8_t *need;     // [PY][PX] probes to compute (nullptr: all); the others are left untouched
};

// k_march_all: every level >= 1 in one launch (level-major, top level first).
struct MarchLevel {
    const float2 *dirs;
    int *codes;              // [PY][PX][N] ray codes
    int PX, PY, N, log2n;
    float inv_px;            // 1/PX in fp32 (probe -> row, corrected)
    float t0, t1, spacing;
    int blocks;              // workgroups of 256*M ray slots (the march grid is blocks x levels)
    uint32_t code_begin;     // first ray of this level in the code buffer (levels 1, 2, ... back to back)
    int dir_off;             // first direction of this level in dirs_base
    // Order of the rays over the lanes (k_march_all only): map 0 = canonical
    // order (slot = canonical index); map > 0 = wave tiles of
    // 2^tx_log x 2^ty_log probes x 2^kd_log directions (tx_log + ty_log +
    // kd_log = 6), tiles numbered (tile row, tile column, direction slice)
    // with the slice fastest; `slots` counts the padded tiles' lanes.
    int map, tx_log, ty_log, kd_log, ntx;
    float inv_ntx;
    uint32_t slots;
    // The top level (no upper level) merged by the march itself: each lane
    // quad holds the 4 child directions of one texel, so the march writes
    // A_top = quad mean of (hit ? (E, 0) : (0, 1)) here, in `storage`
    // (a Storage value), and no level kernel runs for it.  nullptr: not merged.
    // Needs one ray per lane and a lane order whose direction slice is at
    // least 4 (every order but 5).
    void *merged;
    int storage;
};
// Wave-tile orders of the split march: 0 canonical, 1 = 2x2 probes x 16
// directions, 2 = 4x4 x 4, 3 = 4x2 x 8, 4 = 2x1 x 32, 5 = 8x8 x 1.
constexpr int kMarchMaps = 6;
bool march_map_legal(int level, int map);
// fills map, the wave tile, ntx, inv_ntx and slots of a level (PX, PY, N,
// log2n set); the caller sizes blocks from slots
void march_level_map(MarchLevel &l, int map);
struct MarchAllArgs {
    const uint8_t *sdf;
    int W, H, nlev;
    MarchLevel lev[RC_MAX_LEVELS];  // top level first
    const float4 *emis4;     // interleaved emission (a merged level's hits)
    // persistent form: every ray of every level, in 64-ray packets
    int *codes_base;
    const float2 *dirs_base;
    uint32_t total_rays, n_packets, n_waves;
};
// The split march runs one-wave workgroups: march lengths differ by two
// orders of magnitude between lanes and waves, and a workgroup's wave slots
// are freed only when its last wave ends, so a wave is its own workgroup.
constexpr int kMarchBlock = 64;
hipError_t launch_march_all(const MarchAllArgs &a, int speculation, int rays_per_lane, hipStream_t s);
// one level of one view in a batched split march
struct MarchJob {
    const uint8_t *sdf;
    const float4 *emis4;     // the view's interleaved emission (a merged top level's hits)
    int W, H;
    MarchLevel L;
};
struct Level0Job {
    LevelArgs a;
    float *irr;
};
int level_blocks(int PX, int PY, int N, int shape, int rays_per_lane);
// level 0 + gather: a workgroup's tile of kL0TX x kL0TY level-0 probes
constexpr int kL0TX = 16, kL0TY = 12;
int level0_blocks(int PX, int PY);
// jobs: MarchJob (march), LevelArgs (levels, `upper` may be null per job), Level0Job
hipError_t launch_march_batch(const BatchLaunch &b, int speculation, int rays_per_lane, hipStream_t s);
hipError_t launch_level_batch(const BatchLaunch &b, Storage st, int shape, int rays_per_lane, bool from_codes,
                              hipStream_t s);
hipError_t launch_level0_batch(const BatchLaunch &b, Storage st, hipStream_t s);
// The persistent split march: a grid of `waves` waves (a multiple of 4), each
// walking its share of the packets and refilling finished lanes.
hipError_t launch_march_persist(const MarchAllArgs &a, int speculation, hipStream_t s);

hipError_t launch_level(const LevelArgs &a, Storage st, int shape, int rays_per_lane, hipStream_t s);
hipError_t launch_level0_gather(const LevelArgs &a, Storage st, float *irradiance, float *debug_fluence,
                                hipStream_t s);
hipError_t launch_level_to_f32(const void *level, Storage st, size_t texels, float *out, hipStream_t s);
// canonical ray index of each slot as the split march decodes it (parity readback)
hipError_t launch_ray_decode(const MarchLevel &l, int64_t *index, hipStream_t s);

// ---- error norms (rc_norms.hip) ---------------------------------------------
size_t norm_workspace_bytes();
void srgb_thresholds(float *t);   // 256 entries (rc_composite.hip)
hipError_t launch_composite(const float *albedo, const float *emissive, const float *irr, const float *direct,
                            const float *table, uint32_t *out, size_t n, hipStream_t s);
hipError_t launch_error_norms(const float *x, const void *ref, bool ref_f64, size_t n, double *workspace,
                              double *out, hipStream_t s);

// ---- surrogate MLP (rc_mlp.hip) ---------------------------------------------
// 64 -> 64 -> 64 -> 64 -> 64, ReLU on the hidden layers; inputs per direction
// (log(1 + c) rgb, cone transmittance T) x 16; outputs 0..47 the
// far-field residual A_2 - c (16 directions x rgb, normalised), 48 the log of the
// predicted squared error, 49..63 padding.
constexpr int kMlpIn = 64, kMlpHid = 64, kMlpOut = 64, kMlpRad = 48, kMlpErr = 48;
constexpr int kMlpFeat = 4;  // inputs per direction: log(1 + c) rgb, T
constexpr int kMlpParams = kMlpIn * 64 + 64 + 3 * (64 * 64 + 64);  // 16640 fp32 master parameters
constexpr int kMlpWt = 64 * kMlpIn + 3 * 64 * kMlpHid;            // bf16 weights per copy
constexpr float kMlpErrEps = 1e-2f, kMlpErrWeight = 0.05f;

struct MlpDev {
    float *master;          // [kMlpParams]: for l = 0..3, W_l [K_l][64] (input-major) then b_l [64]
    float *adam_m, *adam_v; // Adam moments, same layout
    uint16_t *wt;           // bf16 W_l^T [64][K_l] (output-major), layers back to back
    uint16_t *wk;           // bf16 W_l [K_l][64] (input-major)
    float *bias;            // [4][64]
    float *in_mean, *in_istd;  // [kMlpIn] input standardisation
};

hipError_t launch_mlp_infer(const MlpDev &m, const float *feats, int n, float *out, hipStream_t s);
hipError_t launch_mlp_train(const MlpDev &m, const float *feats, const float *targets, const int *idx, int batch,
                            int nb, float *partial, float *loss_part, int parts, hipStream_t s);
hipError_t launch_mlp_reduce(const float *partial, int parts, float *grad, hipStream_t s);
hipError_t launch_mlp_adam(const MlpDev &m, const float *grad, float lr, float b1, float b2, float eps, float c1,
                           float c2, hipStream_t s);
hipError_t launch_mlp_refresh(const MlpDev &m, hipStream_t s);
hipError_t launch_sum_f32(const float *x, int n, float *out, hipStream_t s);

// ---- surrogate far field (rc_surrogate.hip) ---------------------------------
// Emitter / occluder mip pyramid: level m (1..levels) has cells of 2^m x 2^m
// pixels, ceil(W/2^m) x ceil(H/2^m) of them; texel = (mean E*[emitter] rgb,
// coverage) over the cell's pixels (pixels outside the image are empty), the
// coverage word holding fp16 means of [emitter] (low) and [occupied, not
// emitter] (high).  Each level is stored with a one-cell border of zero
// texels ((w+2) x (h+2), cell (x, y) at (y+1)(w+2) + x+1), written once at
// allocation, so a bilinear footprint that overlaps the image edge needs no
// per-tap test.
constexpr int kMipMax = 14;
struct MipArgs {
    uint4 *tex;             // all levels back to back, 8 fp16 values per cell
    int levels;             // top level index (>= 1)
    int w[kMipMax + 1], h[kMipMax + 1];
    size_t off[kMipMax + 1];
};
__host__ __device__ inline size_t mip_at(const MipArgs &m, int l, int x, int y)
{
    return m.off[l] + (size_t)(y + 1) * (size_t)(m.w[l] + 2) + (size_t)(x + 1);
}
// Cone trace schedule: every cone takes the same distances t_n, so the mip
// level m_n and the opacity exponents are uniform over a launch; a step
// carries what its samples need (one scalar load of 32 bytes).
constexpr int kConeMaxSteps = 64;
constexpr double kConeOccGain = 3.0;  // occluder exponent over the emitter one
struct ConeStep {
    float t_cell, re, ro, inv_cell;  // distance in cells (t 2^-m), emitter / occluder exponents, 2^-m
    int w, h;                // size of the step's mip level in cells
    uint32_t off, wp;        // first (border) texel of the level, padded row length w+2
};
struct ConeSchedule {
    int n;
    int pad[7];
    ConeStep s[kConeMaxSteps];
};
// Cones are traced in two segments: the near one (t < kConeSplit) from the
// level-2 probes, the far one (t >= kConeSplit) from the level-3 probes (4x
// fewer cones; at that distance a cone is wider than the probe spacing); a
// level-2 cone ends with L += T * F.rgb, T *= F.T, F the far cones
// interpolated at its probe with the merge's bilinear weights.
constexpr float kConeSplit = 60.0f;
struct ConeArgs {
    MipArgs mip;
    uint32_t mip_bytes;     // size of the whole pyramid (range of the buffer loads)
    const ConeSchedule *sch;  // the launch's steps (device memory)
    const float2 *dirs;     // the 16 level-1 directions
    float4 *cone;           // [QY][QX][16]: radiance rgb, transmittance
    int QX, QY;             // probe grid of this launch (level 2, or level 3 for the far segment)
    float spacing;          // its probe spacing (8 or 16)
    const float4 *far;      // far-segment cones [FY][FX][16] to append, or nullptr
    int FX, FY;
};
struct SurrMlpArgs {
    const float4 *cone;     // [QY][QX][16] cones at the level-2 probes
    const float4 *far;      // far-segment cones [FY][FX][16] to append, or nullptr (cone already whole)
    int FX, FY;
    float4 *cone_out;       // whole cones written back (parity readback), or nullptr
    void *out;              // the network's merged level 2 (storage texels, [QY][QX][16])
    float *tile_err;        // [tiles] mean predicted squared error (normalised), tiles of 8x8 level-2 probes
    int *flag_count;        // tiles above threshold (counter)
    int *flag_list;         // their indices
    int QX, QY, tiles_x, tiles;
    float out_scale, threshold;
};
struct SurrSampleArgs {
    const float4 *cone;
    const void *upper;      // merged level 2 of the plain path (storage texels)
    float *feats;           // [QY*QX][64] per direction log(1 + cone radiance) rgb, T
    float *targets;         // [QY*QX][48] merged level 2 minus cone radiance
    int QX, QY;
};

void mip_layout(int W, int H, int top_level, MipArgs &m);   // fills w, h, off, levels
size_t mip_cells(const MipArgs &m);
// near: steps with t < kConeSplit; far: the rest of the same sequence
void cone_schedule(int W, int H, const MipArgs &m, ConeSchedule &near, ConeSchedule &far);
int cone_top_level(int W, int H);
hipError_t launch_seed_surr(const float *depth, const DepthConv &dc, const float *emissive, uint64_t *bits,
                            const MipArgs &m, int W, int H, hipStream_t s);  // occupancy bits + mip levels 1..6
hipError_t launch_near_sdf(const uint64_t *bits, uint8_t *sdf, int W, int H, hipStream_t s);
hipError_t launch_mip_top(const MipArgs &m, int *zero_me, hipStream_t s);    // levels 7..top
hipError_t launch_mip_unpack(const MipArgs &m, int level, float *emission, float *coverage, hipStream_t s);
hipError_t launch_cone(const ConeArgs &a, hipStream_t s);
hipError_t launch_cone_pair(const ConeArgs &near, const ConeArgs &far, hipStream_t s);   // one launch, both grids
hipError_t launch_mip_top_near(const MipArgs &m, int *zero_me, const uint64_t *bits, uint8_t *sdf, int W, int H,
                               hipStream_t s);  // levels 7..top beside the near field
hipError_t launch_surr_mlp(const SurrMlpArgs &a, const MlpDev &m, Storage st, hipStream_t s);
hipError_t launch_surr_sample(const SurrSampleArgs &a, Storage st, hipStream_t s);
// Tile fallback of the learned frame: the probes of each level that the
// flagged tiles need (level 1: the tiles' own 16x16 probes; level i+1: the
// bilinear neighbourhood of level i's), one byte per probe, levels 1..L-1
// back to back at need_off[i].
struct FallbackMaskArgs {
    const int *count;        // flagged tiles (device counter)
    const int *list;         // their indices (tiles of 8x8 level-2 probes)
    uint8_t *need;
    size_t need_off[RC_MAX_LEVELS];
    int PX[RC_MAX_LEVELS], PY[RC_MAX_LEVELS];
    int L, tiles_x, max_tiles;
    size_t need_bytes;
};
hipError_t launch_fallback_masks(const FallbackMaskArgs &a, hipStream_t s);
// tally[0] += flagged tiles of this frame, tally[1] += 1 if any
hipError_t launch_fallback_tally(const int *count, unsigned long long *tally, hipStream_t s);
hipError_t launch_col_stats(const float *x, size_t rows, int cols, double *partial, int parts, hipStream_t s);
hipError_t launch_scale(float *x, size_t n, float k, hipStream_t s);

}  // namespace rc
